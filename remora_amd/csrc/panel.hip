// Fused Cholesky panel kernel: factor the kb x kb diagonal block and solve
// the rows below it against the factor while the factor is still in LDS.
//
// Replaces the reference HIP backend's per-16-column potrf_block_kernel +
// cuBLAS trsm pair (include/remora/kernels/hip/potrf.hpp:41-80, 113) with one
// launch per diagonal block.  Every workgroup loads the small diagonal block
// and factors it itself (right-looking, in one wavefront: lane i keeps row i
// in registers and pivot-column entries move lane to lane with readlane), so
// no workgroup waits for another and no inverse is formed; it then solves
// its own 256 rows of the panel,
//   x_c = (a_c - sum_{t<c} x_t L_ct) / L_cc,
// one row per thread with the row in registers and L broadcast from LDS.
// Every workgroup computes the same factor of the diagonal block; the one
// that is last to finish reading the block from memory writes the factor
// back (a counter in scratch memory, which it also resets), so no workgroup
// can read an already factored block.
// The arithmetic is FP64 whatever the storage type (the mixed-precision
// factorisation keeps its diagonal blocks in FP64).
#include "rmh_internal.hpp"

namespace rmh {
namespace {

constexpr int kRows = 256;  // panel rows per workgroup (one per thread)

// value of v in lane `lane` (wave-uniform lane index)
__device__ __forceinline__ double lane_bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// View: V(i, c) = A[i*rs + c*cs] for rows [0, rows) and columns [0, kb);
// V(0,0) is the diagonal element of the panel.
template <typename T, int NB>
__global__ __launch_bounds__(kRows) void potrf_panel_kernel(T* A, int64_t rs, int64_t cs, int64_t rows, int kb,
                                                            int64_t* info, int64_t info_base, unsigned* readers) {
    constexpr int LD = NB + 1;
    __shared__ double L[NB * LD];
    __shared__ double rdiag[NB];
    __shared__ int failed;
    __shared__ int last_reader;
    const int tid = threadIdx.x;
    // this thread's panel row, loaded first so the load overlaps the factorisation
    const int64_t i = kb + (int64_t)blockIdx.x * kRows + tid;
    const bool has_row = i < rows;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = (has_row && c < kb) ? (double)A[i * rs + c * cs] : 0.0;

    // 1. the diagonal block's lower triangle into LDS
    for (int e = tid; e < kb * kb; e += kRows) {
        const int i = e % kb, c = e / kb;
        L[i * LD + c] = (c <= i) ? (double)A[i * rs + c * cs] : 0.0;
    }
    if (tid == 0) failed = -1;
    __syncthreads();  // the block is in LDS: this workgroup is done reading it
    if (tid == 0) last_reader = (atomicAdd(readers, 1u) == gridDim.x - 1) ? 1 : 0;

    // 2. right-looking factorisation by wavefront 0, lane i holding row i in
    // registers; pivots and pivot-column entries are broadcast lane to lane
    // (readlane), so the factorisation needs no barrier.  Unrolled loops keep
    // the register indices static; steps past kb are skipped uniformly.
    int bad = -1;
    if (tid < 64) {
        double r[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) r[c] = (tid < kb && c < kb) ? L[tid * LD + c] : 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < kb && bad < 0) {
                const double d = lane_bcast(r[j], j);
                if (!(d > 0.0)) {  // also catches NaN
                    bad = j;
                } else {
                    const double sj = sqrt(d);
                    r[j] = (tid == j) ? sj : r[j] / sj;
#pragma unroll
                    for (int c = j + 1; c < NB; ++c) {
                        if (c < kb) r[c] = fma(-r[j], lane_bcast(r[j], c), r[c]);  // lanes < c touch only
                    }                                                                   // their unused upper part
                }
            }
        }
        if (tid < kb) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= tid) L[tid * LD + c] = r[c];
        }
        if (tid == 0) failed = bad;
    }
    __syncthreads();
    bad = failed;
    const bool writer = last_reader != 0;
    if (bad >= 0) {
        if (writer && tid == 0) {
            // keep the first failure: later panels run on garbage but must not overwrite it
            atomicCAS(reinterpret_cast<unsigned long long*>(info), 0ull, (unsigned long long)(info_base + bad + 1));
            *readers = 0u;
        }
        return;
    }
    if (tid < NB) rdiag[tid] = (tid < kb) ? 1.0 / L[tid * LD + tid] : 0.0;
    if (writer) {
        for (int e = tid; e < kb * kb; e += kRows) {
            const int i = e % kb, c = e / kb;
            if (c <= i) A[i * rs + c * cs] = (T)L[i * LD + c];
        }
        if (tid == 0) *readers = 0u;  // ready for the next panel launch
    }
    __syncthreads();

    // 3. rows below the diagonal block: X L^T = A21
    if (!has_row) return;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < kb) {
            double a0 = x[c], a1 = 0.0;  // two chains for instruction-level parallelism
#pragma unroll
            for (int t = 0; t + 1 < c; t += 2) {
                a0 = fma(-x[t], L[c * LD + t], a0);
                a1 = fma(-x[t + 1], L[c * LD + t + 1], a1);
            }
            if (c & 1) a0 = fma(-x[c - 1], L[c * LD + c - 1], a0);
            x[c] = (a0 + a1) * rdiag[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < kb) A[i * rs + c * cs] = (T)x[c];
}

}  // namespace

template <typename T>
rmh_status launch_potrf_panel(T* A, int64_t rs, int64_t cs, int64_t rows, int kb, int64_t* info, int64_t info_base,
                              unsigned* readers, hipStream_t stream) {
    if (kb <= 0) return RMH_STATUS_SUCCESS;
    if (rows < kb) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: fewer rows than columns");
    const int64_t below = rows - kb;
    const unsigned grid = (unsigned)std::max<int64_t>(1, (below + kRows - 1) / kRows);
    if (kb <= 16)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 16>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base, readers);
    else if (kb <= 32)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 32>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base, readers);
    else if (kb <= 64)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 64>), dim3(grid), dim3(kRows), 0, stream, A, rs, cs, rows, kb, info,
                           info_base, readers);
    else
        return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: block wider than 64");
    return check_launch("potrf_panel_kernel");
}

template rmh_status launch_potrf_panel<double>(double*, int64_t, int64_t, int64_t, int, int64_t*, int64_t, unsigned*,
                                               hipStream_t);
template rmh_status launch_potrf_panel<float>(float*, int64_t, int64_t, int64_t, int, int64_t*, int64_t, unsigned*,
                                              hipStream_t);

}  // namespace rmh
