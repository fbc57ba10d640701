// Fused left-looking Cholesky panel step.  For one block column of width
// kb (<= 64) inside a panel it
//   1. applies the updates of the panel's earlier columns (FP64 MFMA),
//   2. factors the kb x kb diagonal block in one wavefront, and
//   3. solves the rows below it against the factor while the factor is in LDS
// -- one launch per block column.  Replaces the reference HIP backend's
// per-16-column potrf_block_kernel + cuBLAS trsm pair
// (include/remora/kernels/hip/potrf.hpp:41-80, 113).
//
// View: V(i, c) = P[i*rs + c*cs], V(0, K) the diagonal element of the block
// column; columns [0, K) are the panel's already factored columns, columns
// [K, K+kb) the block column; rows [0, rows) run from the diagonal down.
//
// Each workgroup owns 64 rows below the diagonal block.  It streams K-chunks
// of the diagonal rows D = V(0:kb, 0:K) and of its own rows R through LDS and
// accumulates, on 16x16x4 FP64 MFMA tiles,
//     own  -= R D^T      (64 x kb)       diag -= D D^T   (lower blocks)
// Every workgroup forms the same updated diagonal block and factors it
// itself (right-looking, lane i holds row i, pivot entries move lane to lane
// by readlane), so no workgroup waits for another.  Workgroup 0 writes the
// factor to a side buffer rather than over the block, which the other
// workgroups may still be reading; the driver copies all diagonal factors
// into place once the factorisation is done.  Rows are then solved one per
// lane,
//     x_c = (a_c - sum_{t<c} x_t L_ct) / L_cc.
// The arithmetic is FP64 whatever the storage type (the mixed-precision
// factorisation keeps its diagonal blocks in FP64).
#include "rmh_internal.hpp"
#include "wave_chol.hpp"

namespace rmh {
namespace {

constexpr int kThreads = 256;
constexpr int kRW = 64;  // rows per workgroup
constexpr int kKc = 16;  // K chunk staged in LDS (keeps LDS at 58 KB for NB = 32)
// LDS row strides (doubles) of the k-major chunk buffers: == 16 (mod 32), so
// the two 16-lane halves of a 32-lane LDS pass hit different banks
constexpr int stride16(int w) { return ((w + 15) / 32) * 32 + 16; }

typedef double dx4 __attribute__((ext_vector_type(4)));

// d-th block (bi >= bj) of the lower block triangle, row by row
__device__ __forceinline__ void lower_block(int d, int& bi, int& bj) {
    bi = 0;
    while (d > bi) { d -= bi + 1; ++bi; }
    bj = d;
}

template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void potrf_panel_kernel(T* P, int64_t rs, int64_t cs, int64_t rows, int K,
                                                               int kb, int64_t* info, int64_t info_base, T* dfac,
                                                               int64_t ldf) {
    constexpr int NCB = NB / 16;              // 16-column blocks
    constexpr int NDB = NCB * (NCB + 1) / 2;  // lower blocks of the diagonal block
    constexpr int DPW = (NDB + 3) / 4;        // diagonal blocks per wave
    constexpr int SD = stride16(NB), SR = stride16(kRW);
    constexpr int LD = NB + 1;
    constexpr int PD = NB * kKc / kThreads, PR = kRW * kKc / kThreads;  // chunk elements per thread
    constexpr int CD = NB * NB / kThreads, CR = kRW * NB / kThreads;    // block-column elements per thread
    static_assert(NB % 16 == 0 && PD >= 1 && CD >= 1, "unsupported block width");

    // K chunks are double-buffered in LDS; two register sets keep the global
    // loads two chunks ahead of the MFMAs
    __shared__ double Dk[2][kKc * SD];  // Dk[b][k*SD + i]: diagonal rows, K chunk
    __shared__ double Rk[2][kKc * SR];  // Rk[b][k*SR + i]: own rows, K chunk
    __shared__ double L[NB * LD];       // updated diagonal block, then its factor
    __shared__ double X[kRW * LD];      // updated own rows
    __shared__ double rdiag[NB];
    __shared__ int failed;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t r0 = kb + (int64_t)blockIdx.x * kRW;  // first own row
    const int nown = (int)max((int64_t)0, min((int64_t)kRW, rows - r0));
    const int nch = (K + kKc - 1) / kKc;

    // block column: own rows and the diagonal block (padded with the
    // identity), loaded first
    double cur_r[CR], cur_d[CD];
#pragma unroll
    for (int q = 0; q < CR; ++q) {
        const int e = tid + q * kThreads, i = e % kRW, c = e / kRW;
        cur_r[q] = (i < nown && c < kb) ? (double)P[(r0 + i) * rs + (int64_t)(K + c) * cs] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CD; ++q) {
        const int e = tid + q * kThreads, i = e % NB, c = e / NB;
        cur_d[q] = (i < kb && c <= i) ? (double)P[i * rs + (int64_t)(K + c) * cs] : (i == c ? 1.0 : 0.0);
    }
    double pa_d[PD], pa_r[PR], pb_d[PD], pb_r[PR];
    auto load_chunk = [&](int ch, double(&pd)[PD], double(&pr)[PR]) {
        const int k0 = ch * kKc, kc = min(kKc, K - k0);
#pragma unroll
        for (int q = 0; q < PD; ++q) {
            const int e = tid + q * kThreads, i = e % NB, c = e / NB;
            pd[q] = (i < kb && c < kc) ? (double)P[i * rs + (int64_t)(k0 + c) * cs] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PR; ++q) {
            const int e = tid + q * kThreads, i = e % kRW, c = e / kRW;
            pr[q] = (i < nown && c < kc) ? (double)P[(r0 + i) * rs + (int64_t)(k0 + c) * cs] : 0.0;
        }
    };
    auto store_chunk = [&](int b, const double(&pd)[PD], const double(&pr)[PR]) {
#pragma unroll
        for (int q = 0; q < PD; ++q) {
            const int e = tid + q * kThreads;
            Dk[b][(e / NB) * SD + e % NB] = pd[q];
        }
#pragma unroll
        for (int q = 0; q < PR; ++q) {
            const int e = tid + q * kThreads;
            Rk[b][(e / kRW) * SR + e % kRW] = pr[q];
        }
    };
    if (nch > 0) load_chunk(0, pa_d, pa_r);
    if (nch > 1) load_chunk(1, pb_d, pb_r);

#pragma unroll
    for (int q = 0; q < CR; ++q) {
        const int e = tid + q * kThreads;
        X[(e % kRW) * LD + e / kRW] = cur_r[q];
    }
#pragma unroll
    for (int q = 0; q < CD; ++q) {
        const int e = tid + q * kThreads;
        L[(e % NB) * LD + e / NB] = cur_d[q];
    }
    if (nch > 0) store_chunk(0, pa_d, pa_r);
    if (nch > 2) load_chunk(2, pa_d, pa_r);
    if (tid == 0) failed = -1;
    __syncthreads();

    // 1. left-looking update on MFMA: wave w owns rows [16w, 16w+16) of X
    // (all column blocks) and diagonal blocks w, w+4, ...
    dx4 acc[NCB], dacc[DPW];
#pragma unroll
    for (int b = 0; b < NCB; ++b) acc[b] = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < DPW; ++b) dacc[b] = dx4{0.0, 0.0, 0.0, 0.0};
    const int li = lane & 15, lk = lane >> 4;
    // this wave's diagonal blocks (wave-uniform)
    int dbi[DPW], dbj[DPW];
#pragma unroll
    for (int q = 0; q < DPW; ++q) {
        dbi[q] = dbj[q] = 0;
        if (wave + 4 * q < NDB) lower_block(wave + 4 * q, dbi[q], dbj[q]);
    }
    // all LDS operands of a chunk are read before its MFMAs are issued; the
    // diagonal blocks' operands are column blocks of the same reads
    auto compute = [&](int buf) {
        const double* dk = Dk[buf];
        const double* rk = Rk[buf];
        double av[kKc / 4], bv[kKc / 4][NCB];
#pragma unroll
        for (int s4 = 0; s4 < kKc / 4; ++s4) {
            const int k = 4 * s4 + lk;
            av[s4] = rk[k * SR + 16 * wave + li];
#pragma unroll
            for (int b = 0; b < NCB; ++b) bv[s4][b] = dk[k * SD + 16 * b + li];
        }
#pragma unroll
        for (int s4 = 0; s4 < kKc / 4; ++s4) {
#pragma unroll
            for (int b = 0; b < NCB; ++b)
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s4], bv[s4][b], acc[b], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < DPW; ++q) {
                if (wave + 4 * q < NDB) {
                    double x = bv[s4][0], y = bv[s4][0];
#pragma unroll
                    for (int b = 1; b < NCB; ++b) {
                        x = (dbi[q] == b) ? bv[s4][b] : x;
                        y = (dbj[q] == b) ? bv[s4][b] : y;
                    }
                    dacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, dacc[q], 0, 0, 0);
                }
            }
        }
    };
    // chunk c sits in LDS buffer c % 2; registers a hold chunks 0, 2, 4, ...
    // and b chunks 1, 3, 5, ... until they are stored
    for (int ch = 0; ch < nch; ch += 2) {
        compute(0);
        if (ch + 1 < nch) {
            store_chunk(1, pb_d, pb_r);  // buffer 1 was last read before the previous barrier
            if (ch + 3 < nch) load_chunk(ch + 3, pb_d, pb_r);
        }
        __syncthreads();
        if (ch + 1 >= nch) break;
        compute(1);
        if (ch + 2 < nch) {
            store_chunk(0, pa_d, pa_r);
            if (ch + 4 < nch) load_chunk(ch + 4, pa_d, pa_r);
        }
        __syncthreads();
    }
    // subtract: lane l, item r of a 16x16 tile is element (4r + l/16, l%16)
    if (K > 0) {
#pragma unroll
        for (int b = 0; b < NCB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) X[(16 * wave + 4 * r + lk) * LD + 16 * b + li] -= acc[b][r];
#pragma unroll
        for (int q = 0; q < DPW; ++q) {
            if (wave + 4 * q < NDB) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * dbi[q] + 4 * r + lk, c = 16 * dbj[q] + li;
                    if (c <= i) L[i * LD + c] -= dacc[q][r];
                }
            }
        }
    }
    __syncthreads();

    // 2. factor the (identity-padded) diagonal block in wavefront 0
    if (tid < NB) {
        const int bad = wave_potrf<NB, LD>(L, rdiag, tid);
        if (tid == 0) failed = bad;
    }
    __syncthreads();
    const int bad = failed;
    if (bad >= 0) {
        // keep the first failure: later launches run on garbage but must not overwrite it
        if (blockIdx.x == 0 && tid == 0)
            atomicCAS(reinterpret_cast<unsigned long long*>(info), 0ull, (unsigned long long)(info_base + bad + 1));
        return;
    }
    // the factor of the diagonal block goes to the side buffer: every
    // workgroup of this launch still reads the block's original values
    if (blockIdx.x == 0) {
        for (int e = tid; e < kb * kb; e += kThreads) {
            const int i = e % kb, c = e / kb;
            if (c <= i) dfac[i + c * ldf] = (T)L[i * LD + c];
        }
    }

    // 3. own rows: X L^T = A21, one row per lane of wavefront 0 (the padded
    // columns come out zero and are not stored)
    if (tid >= nown) return;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = X[tid * LD + c];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        double sum[4] = {x[c], 0.0, 0.0, 0.0};  // four independent chains
#pragma unroll
        for (int t = 0; t < c; ++t) sum[t & 3] = fma(-x[t], L[c * LD + t], sum[t & 3]);
        x[c] = ((sum[0] + sum[1]) + (sum[2] + sum[3])) * rdiag[c];
    }
    const int64_t i = r0 + tid;
    if (kb == NB) {
#pragma unroll
        for (int c = 0; c < NB; ++c) P[i * rs + (int64_t)(K + c) * cs] = (T)x[c];
    } else {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < kb) P[i * rs + (int64_t)(K + c) * cs] = (T)x[c];
    }
}

}  // namespace

// Copy the diagonal-block factors from the side buffer into the matrix:
// element (i, c) of the block starting at j0 = (r / nbi) * nbi, r = j0 + i,
// lives at dfac[r + c * ldf].  Blocks at or after a failed pivot (info != 0)
// keep their original values.
template <typename T>
__global__ __launch_bounds__(256) void scatter_diag_kernel(T* A, int64_t rs, int64_t cs, int64_t n, int nbi,
                                                           const T* __restrict__ dfac, int64_t ldf,
                                                           const int64_t* __restrict__ info) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int64_t j0 = (r / nbi) * nbi;
    const int64_t failed = *info;
    if (failed != 0 && j0 + nbi >= failed) return;
    const int i = (int)(r - j0);
    for (int c = 0; c <= i; ++c) A[r * rs + (j0 + c) * cs] = dfac[r + c * ldf];
}

template <typename T>
rmh_status launch_potrf_panel(T* P, int64_t rs, int64_t cs, int64_t rows, int64_t K, int kb, int64_t* info,
                              int64_t info_base, T* dfac, int64_t ldf, hipStream_t stream) {
    if (kb <= 0) return RMH_STATUS_SUCCESS;
    if (rows < kb) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: fewer rows than columns");
    if (K < 0 || K > (int64_t)1 << 30) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: bad panel offset");
    const int64_t below = rows - kb;
    const unsigned grid = (unsigned)std::max<int64_t>(1, (below + kRW - 1) / kRW);
    const int k = (int)K;
    if (kb <= 16)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 16>), dim3(grid), dim3(kThreads), 0, stream, P, rs, cs, rows, k, kb,
                           info, info_base, dfac, ldf);
    else if (kb <= 32)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 32>), dim3(grid), dim3(kThreads), 0, stream, P, rs, cs, rows, k, kb,
                           info, info_base, dfac, ldf);
    else if (kb <= 64)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 64>), dim3(grid), dim3(kThreads), 0, stream, P, rs, cs, rows, k, kb,
                           info, info_base, dfac, ldf);
    else
        return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: block wider than 64");
    return check_launch("potrf_panel_kernel");
}

template <typename T>
rmh_status launch_scatter_diag(T* A, int64_t rs, int64_t cs, int64_t n, int nbi, const T* dfac, int64_t ldf,
                               const int64_t* info, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL((scatter_diag_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, A, rs, cs,
                       n, nbi, dfac, ldf, info);
    return check_launch("scatter_diag_kernel");
}

template rmh_status launch_potrf_panel<double>(double*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t,
                                               double*, int64_t, hipStream_t);
template rmh_status launch_potrf_panel<float>(float*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t,
                                              float*, int64_t, hipStream_t);
template rmh_status launch_scatter_diag<double>(double*, int64_t, int64_t, int64_t, int, const double*, int64_t,
                                                const int64_t*, hipStream_t);
template rmh_status launch_scatter_diag<float>(float*, int64_t, int64_t, int64_t, int, const float*, int64_t,
                                               const int64_t*, hipStream_t);

}  // namespace rmh
