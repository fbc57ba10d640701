// Mixed-precision SPD solve: factor a reduced-precision copy of A with the
// trailing updates on FP32 MFMA (diagonal blocks in FP64), then refine the
// FP64 solution with FP64 residuals until the backward error
//   eta = max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
// reaches the tolerance.  Falls back to the FP64 factorisation when the
// reduced-precision factor fails or refinement stalls, as LAPACK's dsposv
// does.  No counterpart in the reference (a new form beside the FP64
// drop-in); the FP64 pieces it uses are the same kernels as rmh_dpotrf.
#include <cstring>
#include <string>

#include "rmh_context.hpp"

namespace rmh {
namespace {

constexpr int kT = 64;  // conversion tile

// Convert the stored triangle of A (FP64) to FP32 in A32 and form the
// partial absolute row sums of the symmetric matrix: a stored off-diagonal
// a_ij counts for rows i and j.  One workgroup per 64 x 64 tile of the
// stored triangle (lower-tile order).  The tile on row block R and column
// block C writes, for each row r of block R, the sum over its columns into
// rsp[C * n + r], and for each column c of block C, the sum over its rows
// (off-diagonal entries) into rsp[R * n + c] -- on a diagonal tile both go
// into one slot.  Every slot (s, r), s < number of blocks, has exactly one
// writer, so rowsum_kernel adds them in a fixed order: no atomics, and the
// same ||A||_inf on every run.
// the lower-tile enumeration: tile t of a triangle as (ti, tj), ti >= tj
__device__ __forceinline__ void tri_tile64(int64_t t, int64_t& ti, int64_t& tj) {
    ti = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > t) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    tj = t - ti * (ti + 1) / 2;
}

// part 0: every tile; part 1: the tiles with tj < pt (the first panel of
// the factorisation, pt tile columns); part 2: the others
__global__ __launch_bounds__(256) void convert_sym_kernel(const double* __restrict__ A, int64_t lda,
                                                          float* __restrict__ A32, int64_t ld32, int64_t n,
                                                          bool upper, double* __restrict__ rsp, int part,
                                                          int64_t pt) {
    __shared__ double mag[kT][kT + 1];  // |a| of the stored off-diagonal entries, [row][col]
    __shared__ double rpart[4][kT], cpart[4][kT];
    // tile (ti, tj), ti >= tj, in the lower-tile enumeration
    const int64_t t = blockIdx.x;
    int64_t ti, tj;
    if (part == 1) {
        const int64_t tri = pt * (pt + 1) / 2;
        if (t < tri) {
            tri_tile64(t, ti, tj);
        } else {
            ti = pt + (t - tri) / pt;
            tj = (t - tri) % pt;
        }
    } else {
        tri_tile64(t, ti, tj);
        if (part == 2) {
            ti += pt;
            tj += pt;
        }
    }
    // lower storage: rows ti, cols tj.  upper storage: rows tj, cols ti.
    const int64_t rb = upper ? tj : ti, cb = upper ? ti : tj;
    const int64_t r0 = rb * kT, c0 = cb * kT;
    const int tid = threadIdx.x;
    const int lr = tid & 63, lc0 = tid >> 6;  // 64 rows x 4 column phases
    const int64_t gi = r0 + lr;
    constexpr int kQ = kT / 4;
    double v[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {  // every load in flight before any is used
        const int64_t gj = c0 + lc0 + 4 * q;
        const bool ok = gi < n && gj < n && (upper ? gi <= gj : gi >= gj);
        v[q] = ok ? A[gi + gj * lda] : 0.0;
    }
    double rsum = 0.0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int lc = lc0 + 4 * q;
        const int64_t gj = c0 + lc;
        if (gi < n && gj < n && (upper ? gi <= gj : gi >= gj)) A32[gi + gj * ld32] = (float)v[q];
        const double a = fabs(v[q]);
        rsum += a;
        mag[lr][lc] = gi != gj ? a : 0.0;  // an off-diagonal a_ij also counts for row j
    }
    rpart[lc0][lr] = rsum;
    __syncthreads();
    {  // column sums: thread (q, c) adds rows 16q .. 16q+15 of column c
        const int c = tid & 63, q = tid >> 6;
        double cs = 0.0;
#pragma unroll
        for (int r = 0; r < 16; ++r) cs += mag[16 * q + r][c];
        cpart[q][c] = cs;
    }
    __syncthreads();
    if (rb == cb) {
        if (tid < kT) {
            const double v2 = ((rpart[0][tid] + rpart[1][tid]) + (rpart[2][tid] + rpart[3][tid])) +
                              ((cpart[0][tid] + cpart[1][tid]) + (cpart[2][tid] + cpart[3][tid]));
            if (r0 + tid < n) rsp[cb * n + r0 + tid] = v2;
        }
    } else if (tid < kT) {
        const double v2 = (rpart[0][tid] + rpart[1][tid]) + (rpart[2][tid] + rpart[3][tid]);
        if (r0 + tid < n) rsp[cb * n + r0 + tid] = v2;
    } else if (tid < 2 * kT) {
        const int c = tid - kT;
        const double v2 = (cpart[0][c] + cpart[1][c]) + (cpart[2][c] + cpart[3][c]);
        if (c0 + c < n) rsp[rb * n + c0 + c] = v2;
    }
}

// rowsum[r] = sum over s < nb of rsp[s * n + r], s in order
__global__ __launch_bounds__(256) void rowsum_kernel(const double* __restrict__ rsp, int64_t n, int64_t nb,
                                                     double* __restrict__ rowsum) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double acc = 0.0;
    for (int64_t sl = 0; sl < nb; ++sl) acc += rsp[sl * n + r];
    rowsum[r] = acc;
}

__global__ __launch_bounds__(256) void max_kernel(const double* __restrict__ v, int64_t n, double* __restrict__ out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

template <typename Ti, typename To>
__global__ void cast2d_kernel(const Ti* __restrict__ src, int64_t lds, To* __restrict__ dst, int64_t ldd, int64_t rows,
                              int64_t cols) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        dst[i + j * ldd] = (To)src[i + j * lds];
    }
}

// X += double(C)
__global__ void add_cast_kernel(const float* __restrict__ C, int64_t ldc, double* __restrict__ X, int64_t ldx,
                                int64_t rows, int64_t cols) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        X[i + j * ldx] += (double)C[i + j * ldc];
    }
}

// eta = max_j ||r_j|| / (anorm ||x_j|| + ||b_j||); one workgroup per column,
// the maximum over columns through an integer atomic max on the (non-negative)
// double's bit pattern.  *eta_bits must be zero before the launch.
__global__ __launch_bounds__(256) void berr_kernel(const double* __restrict__ R, int64_t ldr, const double* __restrict__ X,
                                                   int64_t ldx, const double* __restrict__ B, int64_t ldb, int64_t n,
                                                   const double* __restrict__ anorm, unsigned long long* eta_bits) {
    __shared__ double red[3][256];
    const int64_t j = blockIdx.x;
    double r = 0.0, x = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        r = fmax(r, fabs(R[i + j * ldr]));
        x = fmax(x, fabs(X[i + j * ldx]));
        b = fmax(b, fabs(B[i + j * ldb]));
    }
    red[0][threadIdx.x] = r;
    red[1][threadIdx.x] = x;
    red[2][threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double den = *anorm * red[1][0] + red[2][0];
        double eta = den > 0.0 ? red[0][0] / den : (red[0][0] > 0.0 ? INFINITY : 0.0);
        if (!(eta == eta)) eta = INFINITY;  // NaN -> not converged
        atomicMax(eta_bits, (unsigned long long)__double_as_longlong(eta));
    }
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

unsigned grid1d(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65535LL * 16));
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Layout of the WS_MIXED slot.
struct MixedWs {
    float* A32;
    float* C32;
    double* R;
    double* rowsum;
    double* anorm;
    unsigned long long* eta;
    int64_t* info;
    void* inv_ws;  // block inverses of the FP32 factor (blocked solves), or NULL
    double* part;  // split-K partials of the residual product
    double* rsp;   // partial row sums of |A| (blocks x n) from the conversion
};

constexpr int kResidualSplits = 8;

static rmh_status mixed_ws(Context* h, int64_t n, int64_t nrhs, MixedWs* w) {
    const size_t a = align256((size_t)n * n * sizeof(float));
    const size_t c = align256((size_t)n * nrhs * sizeof(float));
    const size_t r = align256((size_t)n * nrhs * sizeof(double));
    const size_t rs = align256((size_t)n * sizeof(double));
    const size_t iv = trsm_blocked_applies(n, nrhs) ? align256(trsm_inverse_workspace<float>(n, nrhs)) : 0;
    const size_t pt = align256((size_t)kResidualSplits * n * nrhs * sizeof(double));
    const size_t rp = align256((size_t)((n + kT - 1) / kT) * n * sizeof(double));
    void* base = nullptr;
    RMH_TRY(h->reserve_slot(WS_MIXED, a + c + r + rs + 1024 + iv + pt + rp, &base));
    char* p = static_cast<char*>(base);
    w->A32 = reinterpret_cast<float*>(p); p += a;
    w->C32 = reinterpret_cast<float*>(p); p += c;
    w->R = reinterpret_cast<double*>(p); p += r;
    w->rowsum = reinterpret_cast<double*>(p); p += rs;
    w->anorm = reinterpret_cast<double*>(p); p += 256;
    w->eta = reinterpret_cast<unsigned long long*>(p); p += 256;
    w->info = reinterpret_cast<int64_t*>(p);
    p += 512;
    w->inv_ws = iv ? static_cast<void*>(p) : nullptr;
    p += iv;
    w->part = reinterpret_cast<double*>(p);
    p += pt;
    w->rsp = reinterpret_cast<double*>(p);
    return RMH_STATUS_SUCCESS;
}

// residual R = B - A X (A symmetric, stored triangle only) and its backward
// error, returned on the host
static rmh_status residual_berr(Context* h, bool upper, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                const double* B, int64_t ldb, const double* X, int64_t ldx, const MixedWs& w,
                                hipStream_t s, double* eta_out) {
    hipLaunchKernelGGL((cast2d_kernel<double, double>), dim3(grid1d(n * nrhs)), dim3(256), 0, s, B, ldb, w.R, n, n,
                       nrhs);
    RMH_TRY(check_launch("cast2d_kernel"));
    GemmDesc<double> g;
    g.M = n; g.N = nrhs; g.K = n;
    g.alpha = -1.0; g.beta = 1.0;
    g.A = A; g.lda = lda; g.B = X; g.ldb = ldx; g.C = w.R; g.ldc = n;
    // K = n is long and M x N = n x nrhs gives few tiles: split the depth
    if (n >= 2048) {
        g.splits = kResidualSplits;
        g.split_ws = w.part;
    }
    RMH_TRY(launch_symm<double>(g, upper, s));
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, s, w.eta);
    hipLaunchKernelGGL(berr_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, w.R, n, X, ldx, B, ldb, n, w.anorm, w.eta);
    RMH_TRY(check_launch("berr_kernel"));
    unsigned long long bits = 0;
    RMH_TRY_HIP(hipMemcpyAsync(&bits, w.eta, sizeof(bits), hipMemcpyDeviceToHost, s));
    RMH_TRY_HIP(hipStreamSynchronize(s));
    double eta;
    std::memcpy(&eta, &bits, sizeof(eta));
    *eta_out = eta;
    return RMH_STATUS_SUCCESS;
}

rmh_status dsposv_impl(Context* h, bool upper, rmh_factor_precision prec, int64_t n, int64_t nrhs, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double* X, int64_t ldx, double tol, int max_iter,
                       int* iters, double* berr, hipStream_t s) {
    if (prec != RMH_FACTOR_FP32 && prec != RMH_FACTOR_BF16X3)
        return fail(RMH_STATUS_INVALID_VALUE, "dsposv: unknown factor precision");
    MixedWs w;
    RMH_TRY(mixed_ws(h, n, nrhs, &w));
    // 1. FP32 copy of the stored triangle and ||A||_inf
    const int64_t nt = (n + kT - 1) / kT;
    // The first panel's tiles are converted on s; the rest, the row sums and
    // the norm run on the side stream beside the first panel's
    // factorisation, which waits for them before its first trailing update
    // (h->cols_ready).  The conversion is bandwidth-bound, the first panel's
    // chain latency-bound.
    const int64_t pt = std::min<int64_t>(nt, (potrf_panel_width(h, false, prec == RMH_FACTOR_BF16X3) + kT - 1) / kT);
    hipStream_t s2 = nullptr;
    hipEvent_t side_a = nullptr, side_b = nullptr;
    const bool overlap = pt < nt;
    if (overlap) {
        RMH_TRY(h->side_stream(&s2, &side_a, &side_b));
        if (h->ev_conv == nullptr) RMH_TRY_HIP(hipEventCreateWithFlags(&h->ev_conv, hipEventDisableTiming));
        if (h->ev_conv0 == nullptr) RMH_TRY_HIP(hipEventCreateWithFlags(&h->ev_conv0, hipEventDisableTiming));
        RMH_TRY_HIP(hipEventRecord(h->ev_conv0, s));  // the caller's earlier work on A
        RMH_TRY_HIP(hipStreamWaitEvent(s2, h->ev_conv0, 0));
        const int64_t rest = nt - pt;
        hipLaunchKernelGGL(convert_sym_kernel, dim3((unsigned)(rest * (rest + 1) / 2)), dim3(256), 0, s2, A, lda, w.A32,
                           n, n, upper, w.rsp, 2, pt);
        RMH_TRY(check_launch("convert_sym_kernel"));
        hipLaunchKernelGGL(convert_sym_kernel, dim3((unsigned)(pt * (pt + 1) / 2 + rest * pt)), dim3(256), 0, s, A, lda,
                           w.A32, n, n, upper, w.rsp, 1, pt);
        RMH_TRY(check_launch("convert_sym_kernel"));
        RMH_TRY_HIP(hipEventRecord(h->ev_conv0, s));  // the first panel's partial row sums
        RMH_TRY_HIP(hipStreamWaitEvent(s2, h->ev_conv0, 0));
    } else {
        hipLaunchKernelGGL(convert_sym_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, s, A, lda, w.A32, n,
                           n, upper, w.rsp, 0, pt);
        RMH_TRY(check_launch("convert_sym_kernel"));
    }
    hipStream_t sn = overlap ? s2 : s;
    hipLaunchKernelGGL(rowsum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sn, w.rsp, n, nt, w.rowsum);
    RMH_TRY(check_launch("rowsum_kernel"));
    hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, sn, w.rowsum, n, w.anorm);
    RMH_TRY(check_launch("max_kernel"));
    if (overlap) RMH_TRY_HIP(hipEventRecord(h->ev_conv, s2));
    // 2. factor in reduced precision (FP64 diagonal blocks, FP32 MFMA updates)
    {
        struct Bf16x3 {  // the split-BF16x3 trailing updates for this factorisation only
            Context* h;
            bool saved;
            Bf16x3(Context* c, bool on) : h(c), saved(c->spotrf_bf16x3) {
                h->bf16x3_trailing = on;
                h->spotrf_bf16x3 = false;  // dsposv's own choice wins
            }
            ~Bf16x3() {
                h->bf16x3_trailing = false;
                h->spotrf_bf16x3 = saved;
            }
        } mode(h, prec == RMH_FACTOR_BF16X3);
        struct ColsReady {  // the factorisation's wait for the side stream's conversion
            Context* h;
            ColsReady(Context* c, hipEvent_t e) : h(c) { h->cols_ready = e; }
            ~ColsReady() { h->cols_ready = nullptr; }
        } ready(h, overlap ? h->ev_conv : nullptr);
        RMH_TRY(potrf_impl<float>(h, upper, n, w.A32, n, w.info, s));
    }
    // a factorisation of one panel does not reach its wait
    if (overlap) RMH_TRY_HIP(hipStreamWaitEvent(s, h->ev_conv, 0));
    int64_t info = 0;
    RMH_TRY_HIP(hipMemcpyAsync(&info, w.info, sizeof(info), hipMemcpyDeviceToHost, s));
    RMH_TRY_HIP(hipStreamSynchronize(s));
    double eta = INFINITY;
    int it = 0;
    bool fallback = info != 0;
    int fallback_code = -3;
    // FP32 solves with the factor: the diagonal-block inverses are built once
    // and shared by the initial solve and every correction
    TrsmInverses<float> inv;
    if (!fallback && w.inv_ws != nullptr) RMH_TRY(trsm_prepare<float>(w.A32, n, n, upper, false, w.inv_ws, &inv, s));
    auto solve32 = [&](float* C) -> rmh_status {
        if (w.inv_ws == nullptr) return potrs_impl<float>(h, upper, n, nrhs, w.A32, n, C, n, s);
        RMH_TRY(trsm_apply<float>(inv, true, upper, n, nrhs, 1.0f, C, n, s));
        return trsm_apply<float>(inv, true, !upper, n, nrhs, 1.0f, C, n, s);
    };
    if (!fallback) {
        // 3. initial solve in FP32
        hipLaunchKernelGGL((cast2d_kernel<double, float>), dim3(grid1d(n * nrhs)), dim3(256), 0, s, B, ldb, w.C32, n,
                           n, nrhs);
        RMH_TRY(solve32(w.C32));
        hipLaunchKernelGGL((cast2d_kernel<float, double>), dim3(grid1d(n * nrhs)), dim3(256), 0, s, w.C32, n, X, ldx,
                           n, nrhs);
        RMH_TRY(check_launch("cast2d_kernel"));
        // 4. refine
        for (it = 0;; ++it) {
            RMH_TRY(residual_berr(h, upper, n, nrhs, A, lda, B, ldb, X, ldx, w, s, &eta));
            if (eta <= tol) break;
            if (it == max_iter) {
                fallback = true;
                fallback_code = -(max_iter + 1);
                break;
            }
            hipLaunchKernelGGL((cast2d_kernel<double, float>), dim3(grid1d(n * nrhs)), dim3(256), 0, s, w.R, n, w.C32,
                               n, n, nrhs);
            RMH_TRY(solve32(w.C32));
            hipLaunchKernelGGL(add_cast_kernel, dim3(grid1d(n * nrhs)), dim3(256), 0, s, w.C32, n, X, ldx, n, nrhs);
            RMH_TRY(check_launch("add_cast_kernel"));
        }
    }
    if (fallback) {
        // FP64 factorisation of a copy of A, as dsposv falls back to dpotrf
        void* f = nullptr;
        RMH_TRY(h->reserve_slot(WS_FACTOR, (size_t)n * n * sizeof(double), &f));
        double* F = static_cast<double*>(f);
        RMH_TRY_HIP(hipMemcpy2DAsync(F, n * sizeof(double), A, lda * sizeof(double), n * sizeof(double), n,
                                     hipMemcpyDeviceToDevice, s));
        RMH_TRY(potrf_impl<double>(h, upper, n, F, n, w.info, s));
        RMH_TRY_HIP(hipMemcpyAsync(&info, w.info, sizeof(info), hipMemcpyDeviceToHost, s));
        RMH_TRY_HIP(hipStreamSynchronize(s));
        if (info != 0) {
            if (iters) *iters = fallback_code;
            return fail(RMH_STATUS_NOT_POSITIVE_DEFINITE,
                        "dsposv: leading minor of order " + std::to_string(info) + " is not positive definite");
        }
        RMH_TRY_HIP(hipMemcpy2DAsync(X, ldx * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), nrhs,
                                     hipMemcpyDeviceToDevice, s));
        RMH_TRY(potrs_impl<double>(h, upper, n, nrhs, F, n, X, ldx, s));
        RMH_TRY(residual_berr(h, upper, n, nrhs, A, lda, B, ldb, X, ldx, w, s, &eta));
        it = fallback_code;
    }
    if (iters) *iters = it;
    if (berr) *berr = eta;
    return RMH_STATUS_SUCCESS;
}

}  // namespace rmh
