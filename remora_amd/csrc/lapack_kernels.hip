// Small-block kernels of the triangular solves: the diagonal-block
// triangular inverses, the in-place application of an inverted diagonal
// block, and BLAS-2 helpers.
#include "rmh_internal.hpp"

namespace rmh {
namespace {

constexpr int kDiagMax = 64;
constexpr int kLdsLd = kDiagMax + 1;  // row stride of the LDS blocks (odd: conflict-free columns)

// Lower-triangular inverse of the kb x kb factor held in L (LDS, row stride
// kLdsLd): lane c computes column c of X = L^-1 by forward substitution.
// No lane reads another lane's column, so no barrier is needed inside.
__device__ __forceinline__ void invert_lower_lds(const double* L, double* X, int kb, bool unit) {
    const int c = threadIdx.x;
    if (c < kb) X[c * kLdsLd + c] = unit ? 1.0 : 1.0 / L[c * kLdsLd + c];
    for (int i = 1; i < kb; ++i) {
        double acc = 0.0;
        for (int t = 0; t < i; ++t) {
            const double l = L[i * kLdsLd + t];
            if (t >= c && c < kb) acc = fma(l, X[t * kLdsLd + c], acc);
        }
        if (c < i) X[i * kLdsLd + c] = unit ? -acc : -acc / L[i * kLdsLd + i];
    }
}

// Invert the nb x nb diagonal blocks of op(A).  One wavefront per block.
// An upper block U is handled as the lower block U^T: inv(U) = inv(U^T)^T.
template <typename T>
__global__ __launch_bounds__(64) void trtri_blocks_kernel(const T* A, int64_t lda, int64_t n, int nb,
                                                          bool eff_upper, bool trans, bool unit, T* Dinv) {
    __shared__ double L[kDiagMax * kLdsLd];
    __shared__ double X[kDiagMax * kLdsLd];
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * nb;
    const int kb = (int)min((int64_t)nb, n - r0);
    const int lane = threadIdx.x;
    // element (i,j) of the effective block op(A)_bb is A[(r0+i) + (r0+j)*lda]
    // (no transpose) or A[(r0+j) + (r0+i)*lda] (transpose).  Load its lower
    // form Lf(i,j) = eff(i,j) if lower, eff(j,i) if upper.
    if (lane < kb) {
        for (int j = 0; j <= lane; ++j) {
            int64_t ei = lane, ej = j;       // Lf(i,j) -> eff index
            if (eff_upper) { ei = j; ej = lane; }
            int64_t ai = ei, aj = ej;        // eff -> A index
            if (trans) { ai = ej; aj = ei; }
            L[lane * kLdsLd + j] = (double)A[(r0 + ai) + (r0 + aj) * lda];
        }
    }
    __syncthreads();
    invert_lower_lds(L, X, kb, unit);
    __syncthreads();
    T* D = Dinv + b * (int64_t)nb * nb;
    // D(i,j) = inverse of eff block, column-major, full kb x kb (ld = nb)
    if (lane < kb) {
        for (int c = 0; c < kb; ++c) {
            double v;
            if (!eff_upper) v = lane >= c ? X[lane * kLdsLd + c] : 0.0;  // lower: D = X
            else v = c >= lane ? X[c * kLdsLd + lane] : 0.0;              // upper: D = X^T
            D[lane + c * (int64_t)nb] = (T)v;
        }
    }
}

// Inverses of the 64 x 64 diagonal blocks of the lower-triangular S
// (S = A, or A^T when s_is_At), written into the block-diagonal strip D:
// D holds inv(S_tt) of each nbd-row block t with element (i, j) of the
// inverse at D[i + (j mod nbd)*ldd].  64-block b lands at rows [64b, 64b+64)
// and local columns [(64b mod nbd), ...+64).  One wavefront per block.
template <typename T>
__global__ __launch_bounds__(64) void trtri_strip_kernel(const T* A, int64_t lda, int64_t n, bool s_is_At, bool unit,
                                                         T* D, int64_t ldd, int64_t nbd) {
    __shared__ double L[kDiagMax * kLdsLd];
    __shared__ double X[kDiagMax * kLdsLd];
    const int64_t r0 = (int64_t)blockIdx.x * kDiagMax;
    const int kb = (int)min((int64_t)kDiagMax, n - r0);
    const int lane = threadIdx.x;
    if (lane < kb)
        for (int j = 0; j <= lane; ++j)
            L[lane * kLdsLd + j] =
                (double)(s_is_At ? A[(r0 + j) + (r0 + lane) * lda] : A[(r0 + lane) + (r0 + j) * lda]);
    __syncthreads();
    invert_lower_lds(L, X, kb, unit);
    __syncthreads();
    T* Dblk = D + r0 + (r0 % nbd) * ldd;
    if (lane < kb)
        for (int c = 0; c < kb; ++c) Dblk[lane + c * ldd] = (T)(lane >= c ? X[lane * kLdsLd + c] : 0.0);
}

template <typename T>
__global__ void scale_kernel(T* B, int64_t ldb, int64_t rows, int64_t cols, T alpha) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        T* b = B + i + j * ldb;
        *b = alpha == T(0) ? T(0) : alpha * *b;
    }
}

// y = alpha*A*x + beta*y: one thread per row, coalesced down the columns.
template <typename T>
__global__ void gemv_n_kernel(int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                              T beta, T* y, int64_t incy) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) acc = fma((double)A[i + j * lda], (double)x[j * incx], acc);
    T* yi = y + i * incy;
    const T v = alpha * (T)acc;
    *yi = beta == T(0) ? v : v + beta * *yi;
}

// y = alpha*A^T*x + beta*y: one workgroup per output, tree reduction.
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                                                     const T* x, int64_t incx, T beta, T* y, int64_t incy) {
    __shared__ double red[256];
    const int64_t j = blockIdx.x;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < m; i += 256) acc = fma((double)A[i + j * lda], (double)x[i * incx], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        T* yj = y + j * incy;
        const T v = alpha * (T)red[0];
        *yj = beta == T(0) ? v : v + beta * *yj;
    }
}

template <typename T>
__global__ void strided_copy_kernel(const T* src, int64_t sinc, T* dst, int64_t dinc, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i * dinc] = src[i * sinc];
}

__global__ void fill_i64_kernel(int64_t* p, int64_t n, int64_t v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65535LL * 32) g = 65535LL * 32;
    return (unsigned)g;
}

}  // namespace

template <typename T>
rmh_status launch_trtri_blocks(const T* A, int64_t lda, int64_t n, int nb, bool upper, bool trans, bool unit,
                               T* Dinv, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    if (nb > kDiagMax) return fail(RMH_STATUS_INVALID_VALUE, "trtri_blocks: block larger than 64");
    const bool eff_upper = upper != trans;
    const int64_t nblk = (n + nb - 1) / nb;
    hipLaunchKernelGGL(trtri_blocks_kernel<T>, dim3((unsigned)nblk), dim3(64), 0, stream, A, lda, n, nb, eff_upper,
                       trans, unit, Dinv);
    return check_launch("trtri_blocks_kernel");
}

template <typename T>
rmh_status launch_trtri_strip(const T* A, int64_t lda, int64_t n, bool s_is_At, bool unit, T* D, int64_t ldd,
                              int64_t nbd, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    if (nbd % kDiagMax != 0) return fail(RMH_STATUS_INVALID_VALUE, "trtri_strip: block not a multiple of 64");
    const int64_t nblk = (n + kDiagMax - 1) / kDiagMax;
    hipLaunchKernelGGL(trtri_strip_kernel<T>, dim3((unsigned)nblk), dim3(64), 0, stream, A, lda, n, s_is_At, unit, D, ldd,
                       nbd);
    return check_launch("trtri_strip_kernel");
}

template <typename T>
rmh_status launch_scale(T* B, int64_t ldb, int64_t rows, int64_t cols, T alpha, hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(grid_for(rows * cols, 256)), dim3(256), 0, stream, B, ldb, rows, cols, alpha);
    return check_launch("scale_kernel");
}

template <typename T>
rmh_status launch_gemv(bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                       T beta, T* y, int64_t incy, hipStream_t stream) {
    if (!trans) {
        if (m <= 0) return RMH_STATUS_SUCCESS;
        hipLaunchKernelGGL(gemv_n_kernel<T>, dim3(grid_for(m, 256)), dim3(256), 0, stream, m, n, alpha, A, lda, x, incx,
                           beta, y, incy);
    } else {
        if (n <= 0) return RMH_STATUS_SUCCESS;
        hipLaunchKernelGGL(gemv_t_kernel<T>, dim3((unsigned)n), dim3(256), 0, stream, m, n, alpha, A, lda, x, incx,
                           beta, y, incy);
    }
    return check_launch("gemv_kernel");
}

template <typename T>
rmh_status launch_strided_copy(const T* src, int64_t sinc, T* dst, int64_t dinc, int64_t n, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL(strided_copy_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, stream, src, sinc, dst, dinc, n);
    return check_launch("strided_copy_kernel");
}

rmh_status launch_fill_i64(int64_t* p, int64_t n, int64_t value, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, n, value);
    return check_launch("fill_i64_kernel");
}

#define RMH_INST(T)                                                                                         \
    template rmh_status launch_trtri_blocks<T>(const T*, int64_t, int64_t, int, bool, bool, bool, T*,       \
                                               hipStream_t);                                               \
    template rmh_status launch_scale<T>(T*, int64_t, int64_t, int64_t, T, hipStream_t);                    \
    template rmh_status launch_trtri_strip<T>(const T*, int64_t, int64_t, bool, bool, T*, int64_t, int64_t,  \
                                              hipStream_t);                                               \
    template rmh_status launch_gemv<T>(bool, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T, \
                                       T*, int64_t, hipStream_t);                                          \
    template rmh_status launch_strided_copy<T>(const T*, int64_t, T*, int64_t, int64_t, hipStream_t);
RMH_INST(double)
RMH_INST(float)
#undef RMH_INST

}  // namespace rmh
