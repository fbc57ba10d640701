// Reverse-mode derivative of the Cholesky factorisation.
//
// For A = L L^T and the gradient Lbar of a scalar f with respect to the
// lower-triangular L, the gradient with respect to (symmetric) A is
//     P    = Phi(L^T Lbar)          Phi: lower triangle, diagonal halved
//     G    = L^-T P L^-1            two triangular solves
//     Abar = (G + G^T) / 2
// (dL = L Phi(L^-1 dA L^-T) gives <Lbar, dL> = <L^-T Phi(L^T Lbar) L^-1, dA>;
// Phi is self-adjoint.)  Upper storage (A = U^T U) is the same with L = U^T,
// Lbar = Ubar^T.
//
// Device steps: zero-padded copies U = L^T (upper) and tril(Lbar); P = U Lbar
// on the MFMA GEMM (lower part only, each tile skipping the depths below its
// first row; U as the plain, untransposed operand); Phi; the two solves
// with U (upper storage) on the TRSM, which rebuilds the inverses of the
// diagonal blocks each call instead of storing L^-1; an in-place
// symmetrisation.  No counterpart
// in the reference (a new form beside the FP64 drop-in).
#include "rmh_context.hpp"

namespace rmh {
namespace {

constexpr int kT = 32;

// dst = tril(op(src)), op = transpose for upper storage, zeros above
template <typename T>
__global__ __launch_bounds__(256) void tril_copy_kernel(const T* __restrict__ src, int64_t lds, int64_t n, bool upper,
                                                        T* __restrict__ dst, int64_t ldd) {
    __shared__ T tile[kT][kT + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kT, j0 = (int64_t)blockIdx.y * kT;
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;  // 32 x 8
    if (i0 < j0) {  // strictly above the diagonal: zeros
        for (int r = ty; r < kT; r += 8) {
            const int64_t i = i0 + tx, j = j0 + r;
            if (i < n && j < n) dst[i + j * ldd] = T(0);
        }
        return;
    }
    if (!upper) {
        for (int r = ty; r < kT; r += 8) {
            const int64_t i = i0 + tx, j = j0 + r;
            if (i < n && j < n) dst[i + j * ldd] = (i >= j) ? src[i + j * lds] : T(0);
        }
        return;
    }
    // upper: dst(i, j) = src(j, i); read the transposed tile coalesced
    for (int r = ty; r < kT; r += 8) {
        const int64_t si = j0 + tx, sj = i0 + r;  // src row, col
        tile[r][tx] = (si < n && sj < n) ? src[si + sj * lds] : T(0);
    }
    __syncthreads();
    for (int r = ty; r < kT; r += 8) {
        const int64_t i = i0 + tx, j = j0 + r;
        if (i < n && j < n) dst[i + j * ldd] = (i >= j) ? tile[tx][r] : T(0);
    }
}

// dst = tril(op(src))^T, an upper triangular matrix with zeros below:
// dst(i, j) = src(j, i) (lower storage) or src(i, j) (upper storage), i <= j
template <typename T>
__global__ __launch_bounds__(256) void triu_t_copy_kernel(const T* __restrict__ src, int64_t lds, int64_t n, bool upper,
                                                          T* __restrict__ dst, int64_t ldd) {
    __shared__ T tile[kT][kT + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kT, j0 = (int64_t)blockIdx.y * kT;
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;  // 32 x 8
    if (i0 > j0) {  // strictly below the diagonal: zeros
        for (int r = ty; r < kT; r += 8) {
            const int64_t i = i0 + tx, j = j0 + r;
            if (i < n && j < n) dst[i + j * ldd] = T(0);
        }
        return;
    }
    if (upper) {
        for (int r = ty; r < kT; r += 8) {
            const int64_t i = i0 + tx, j = j0 + r;
            if (i < n && j < n) dst[i + j * ldd] = (i <= j) ? src[i + j * lds] : T(0);
        }
        return;
    }
    // lower: dst(i, j) = src(j, i); read the transposed tile coalesced
    for (int r = ty; r < kT; r += 8) {
        const int64_t si = j0 + tx, sj = i0 + r;  // src row, col
        tile[r][tx] = (si < n && sj < n) ? src[si + sj * lds] : T(0);
    }
    __syncthreads();
    for (int r = ty; r < kT; r += 8) {
        const int64_t i = i0 + tx, j = j0 + r;
        if (i < n && j < n) dst[i + j * ldd] = (i <= j) ? tile[tx][r] : T(0);
    }
}

// P <- Phi(P): zeros above the diagonal, diagonal halved
template <typename T>
__global__ __launch_bounds__(256) void phi_kernel(T* __restrict__ P, int64_t ld, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x % kT;
    for (int64_t j = (int64_t)blockIdx.y * kT + threadIdx.x / kT; j < min(n, (int64_t)(blockIdx.y + 1) * kT); j += 8) {
        if (i >= n) continue;
        if (i < j) P[i + j * ld] = T(0);
        else if (i == j) P[i + j * ld] *= T(0.5);
    }
}

// S <- (S + S^T) / 2 in place: block (bi, bj), bi >= bj, handles both tiles
template <typename T>
__global__ __launch_bounds__(256) void symmetrize_kernel(T* __restrict__ S, int64_t ld, int64_t n) {
    __shared__ T a[kT][kT + 1], b[kT][kT + 1];
    const int64_t bi = blockIdx.x, bj = blockIdx.y;
    if (bi < bj) return;
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
    const int64_t i0 = bi * kT, j0 = bj * kT;
    for (int r = ty; r < kT; r += 8) {
        const int64_t i = i0 + tx, j = j0 + r;       // tile (bi, bj)
        a[r][tx] = (i < n && j < n) ? S[i + j * ld] : T(0);
        const int64_t i2 = j0 + tx, j2 = i0 + r;     // tile (bj, bi)
        b[r][tx] = (i2 < n && j2 < n) ? S[i2 + j2 * ld] : T(0);
    }
    __syncthreads();
    // element (i0+x, j0+y) of the result is (a[y][x] + b[x][y]) / 2, and its
    // mirror (j0+y, i0+x) takes the same value
    for (int r = ty; r < kT; r += 8) {
        const int64_t i = i0 + tx, j = j0 + r;
        const T v = T(0.5) * (a[r][tx] + b[tx][r]);
        if (i < n && j < n) S[i + j * ld] = v;
        const int64_t i2 = j0 + tx, j2 = i0 + r;
        const T w = T(0.5) * (b[r][tx] + a[tx][r]);
        if (i2 < n && j2 < n) S[i2 + j2 * ld] = w;
    }
}

}  // namespace

template <typename T>
rmh_status potrf_adjoint_impl(Context* h, bool upper, int64_t n, const T* L, int64_t ldl, const T* Lbar, int64_t ldlb,
                              T* Abar, int64_t ldab, hipStream_t s) {
    if (n == 0) return RMH_STATUS_SUCCESS;
    const size_t mat = (size_t)n * n * sizeof(T);
    void* w = nullptr;
    RMH_TRY(h->reserve_slot(WS_ADJOINT, 2 * ((mat + 255) / 256 * 256), &w));
    T* U = static_cast<T*>(w);
    T* Lbz = reinterpret_cast<T*>(static_cast<char*>(w) + (mat + 255) / 256 * 256);
    const unsigned nt = (unsigned)((n + kT - 1) / kT);
    hipLaunchKernelGGL((triu_t_copy_kernel<T>), dim3(nt, nt), dim3(256), 0, s, L, ldl, n, upper, U, n);
    RMH_TRY(check_launch("triu_t_copy_kernel"));
    hipLaunchKernelGGL((tril_copy_kernel<T>), dim3(nt, nt), dim3(256), 0, s, Lbar, ldlb, n, upper, Lbz, n);
    RMH_TRY(check_launch("tril_copy_kernel"));
    // P = lower(U Lbar), U = L^T, into Abar.  U is upper and Lbar lower
    // triangular, so for a tile whose rows start at r0 only k >= r0
    // contributes.
    GemmDesc<T> g;
    g.M = n; g.N = n; g.K = n;
    g.alpha = T(1); g.beta = T(0);
    g.A = U; g.lda = n; g.transA = false;
    g.B = Lbz; g.ldb = n; g.transB = false;
    g.C = Abar; g.ldc = ldab;
    g.part = Part::Lower;
    g.k_from_row = true;
    RMH_TRY(launch_gemm<T>(g, pick_tile(n, n, Part::Lower), s));
    hipLaunchKernelGGL((phi_kernel<T>), dim3(nt, nt), dim3(256), 0, s, Abar, ldab, n);
    RMH_TRY(check_launch("phi_kernel"));
    // G = L^-T P L^-1: U X = P, then X U^T = (.)
    RMH_TRY(trsm_impl<T>(h, true, true, false, false, n, n, T(1), U, n, Abar, ldab, s));
    RMH_TRY(trsm_impl<T>(h, false, true, true, false, n, n, T(1), U, n, Abar, ldab, s));
    hipLaunchKernelGGL((symmetrize_kernel<T>), dim3(nt, nt), dim3(256), 0, s, Abar, ldab, n);
    return check_launch("symmetrize_kernel");
}

template rmh_status potrf_adjoint_impl<double>(Context*, bool, int64_t, const double*, int64_t, const double*, int64_t,
                                               double*, int64_t, hipStream_t);
template rmh_status potrf_adjoint_impl<float>(Context*, bool, int64_t, const float*, int64_t, const float*, int64_t,
                                              float*, int64_t, hipStream_t);

}  // namespace rmh
