// Fused Cholesky panel kernel: factor the kb x kb diagonal block and solve
// the rows below it against the factor while the factor is still in LDS.
//
// Replaces the reference HIP backend's per-16-column potrf_block_kernel +
// cuBLAS trsm pair (include/remora/kernels/hip/potrf.hpp:41-80, 113) with one
// launch per diagonal block.  Every workgroup loads the small diagonal block
// and factors it in LDS itself (right-looking: the thread that owns row i
// keeps the row in registers, each column step broadcasts the scaled pivot
// column through LDS), so no workgroup waits for another and no inverse is
// formed; it then solves its own 256 rows of the panel,
//   x_c = (a_c - sum_{t<c} x_t L_ct) / L_cc,
// one row per thread with the row in registers and L broadcast from LDS.
// The workgroup that holds the first rows writes the factored diagonal block.
// The arithmetic is FP64 whatever the storage type (the mixed-precision
// factorisation keeps its diagonal blocks in FP64).
#include "rmh_internal.hpp"

namespace rmh {
namespace {

constexpr int kRows = 256;  // panel rows per workgroup (one per thread)

// View: V(i, c) = A[i*rs + c*cs] for rows [0, rows) and columns [0, kb);
// V(0,0) is the diagonal element of the panel.
template <typename T, int NB>
__global__ __launch_bounds__(kRows) void potrf_panel_kernel(T* A, int64_t rs, int64_t cs, int64_t rows, int kb,
                                                            int64_t* info, int64_t info_base) {
    constexpr int LD = NB + 1;
    __shared__ double L[NB * LD];
    __shared__ double col[NB];
    __shared__ double piv;
    __shared__ int failed;
    if (*info != 0) return;  // an earlier panel failed: LAPACK stops there
    const int tid = threadIdx.x;

    // 1. the diagonal block's lower triangle into LDS
    for (int e = tid; e < kb * kb; e += kRows) {
        const int i = e % kb, c = e / kb;
        L[i * LD + c] = (c <= i) ? (double)A[i * rs + c * cs] : 0.0;
    }
    if (tid == 0) failed = -1;
    __syncthreads();

    // 2. right-looking factorisation, row `tid` in registers.  The loops are
    // fully unrolled (register arrays need static indices); steps beyond kb
    // or after a failed pivot are skipped with uniform predicates.
    // `failed` (LDS) is written only by the pivot thread before a barrier;
    // every thread copies it into `bad` right after that barrier, so the
    // uniform branches below never see a value another thread is writing.
    double r[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) r[c] = (tid < kb && c < kb) ? L[tid * LD + c] : 0.0;
    int bad = -1;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < kb && bad < 0) {
            if (tid == j) {
                const double d = r[j];
                if (!(d > 0.0)) failed = j;  // also catches NaN
                else piv = sqrt(d);
            }
            __syncthreads();
            bad = failed;
            if (bad < 0) {
                const double sj = piv;
                if (tid >= j && tid < kb) {
                    r[j] = (tid == j) ? sj : r[j] / sj;
                    col[tid] = r[j];
                }
                __syncthreads();
                if (tid > j && tid < kb) {
#pragma unroll
                    for (int c = j + 1; c < NB; ++c)
                        if (c <= tid) r[c] = fma(-r[j], col[c], r[c]);
                }
            }
        }
    }
    __syncthreads();
    if (bad >= 0) {
        if (blockIdx.x == 0 && tid == 0) *info = info_base + bad + 1;
        return;
    }
    if (tid < kb) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c <= tid) L[tid * LD + c] = r[c];
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid < kb) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c <= tid) A[tid * rs + c * cs] = (T)r[c];
    }

    // 3. rows below the diagonal block: X L^T = A21
    const int64_t i = kb + (int64_t)blockIdx.x * kRows + tid;
    if (i >= rows) return;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = (c < kb) ? (double)A[i * rs + c * cs] : 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < kb) {
            double acc = x[c];
#pragma unroll
            for (int t = 0; t < c; ++t) acc = fma(-x[t], L[c * LD + t], acc);
            x[c] = acc / L[c * LD + c];
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < kb) A[i * rs + c * cs] = (T)x[c];
}

}  // namespace

template <typename T>
rmh_status launch_potrf_panel(T* A, int64_t rs, int64_t cs, int64_t rows, int kb, int64_t* info, int64_t info_base,
                              hipStream_t stream) {
    if (kb <= 0) return RMH_STATUS_SUCCESS;
    if (rows < kb) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: fewer rows than columns");
    const int64_t below = rows - kb;
    const unsigned grid = (unsigned)std::max<int64_t>(1, (below + kRows - 1) / kRows);
    if (kb <= 16)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 16>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base);
    else if (kb <= 32)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 32>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base);
    else if (kb <= 64)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 64>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base);
    else
        return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: block wider than 64");
    return check_launch("potrf_panel_kernel");
}

template rmh_status launch_potrf_panel<double>(double*, int64_t, int64_t, int64_t, int, int64_t*, int64_t, hipStream_t);
template rmh_status launch_potrf_panel<float>(float*, int64_t, int64_t, int64_t, int, int64_t*, int64_t, hipStream_t);

}  // namespace rmh
