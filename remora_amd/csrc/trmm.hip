// Triangular products: x = op(T) x (trmv) and B = alpha op(T) B or
// B = alpha B op(T) (trmm), T the upper or lower triangle of A with an
// optional unit diagonal; both in place.  These are the blas_device::trmv /
// trmm entry points (include/remora/hip/cublas_backend.hpp:222-297) that
// kernels::trmv / kernels::trmm on hip_tag call (kernels/hip/trmv.hpp,
// kernels/hip/trmm.hpp).
//
// trmv reads the old x from a contiguous copy, then one pass writes op(T) x:
// without transpose one thread per row (coalesced down the columns), with
// transpose one workgroup per output (a column dot product).  trmm copies
// op(T)'s triangle into a zero-padded square (unit diagonal written as 1) and
// B into a scratch copy, and forms the product on the MFMA GEMM.
#include "rmh_context.hpp"

namespace rmh {
namespace {

// x_i = sum_{k in row i of T} T(i, k) t_k, T = lower (k <= i) or upper (k >= i)
template <typename T>
__global__ void trmv_n_kernel(int64_t n, bool upper, bool unit, const T* A, int64_t lda, const T* t, T* x,
                              int64_t incx) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k0 = upper ? i + 1 : 0, k1 = upper ? n : i;  // off-diagonal part
    double acc = unit ? (double)t[i] : (double)A[i + i * lda] * (double)t[i];
    for (int64_t k = k0; k < k1; ++k) acc = fma((double)A[i + k * lda], (double)t[k], acc);
    x[i * incx] = (T)acc;
}

// x_j = sum_{i in column j of T} T(i, j) t_i, T = lower (i >= j) or upper (i <= j)
template <typename T>
__global__ __launch_bounds__(256) void trmv_t_kernel(int64_t n, bool upper, bool unit, const T* A, int64_t lda,
                                                     const T* t, T* x, int64_t incx) {
    __shared__ double red[256];
    const int64_t j = blockIdx.x;
    const int64_t i0 = upper ? 0 : j + 1, i1 = upper ? j : n;  // off-diagonal part
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc = fma((double)A[i + j * lda], (double)t[i], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double d = unit ? (double)t[j] : (double)A[j + j * lda] * (double)t[j];
        x[j * incx] = (T)(red[0] + d);
    }
}

// Z (na x na, ld na) <- the triangle of A (upper or lower), zeros elsewhere,
// 1 on the diagonal when unit.
template <typename T>
__global__ void tri_pad_kernel(const T* A, int64_t lda, int64_t na, bool upper, bool unit, T* Z) {
    const int64_t total = na * na;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % na, j = e / na;
        T v = T(0);
        if (i == j) v = unit ? T(1) : A[i + j * lda];
        else if (upper ? i < j : i > j) v = A[i + j * lda];
        Z[e] = v;
    }
}

unsigned blocks_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65535LL * 32) g = 65535LL * 32;
    return (unsigned)g;
}

}  // namespace

template <typename T>
rmh_status trmv_impl(Context* h, bool upper, bool trans, bool unit, int64_t n, const T* A, int64_t lda, T* x,
                     int64_t incx, hipStream_t s) {
    if (n == 0) return RMH_STATUS_SUCCESS;
    if (n > (int64_t)65535 * 32 * 256) return fail(RMH_STATUS_INVALID_VALUE, "trmv: order too large");
    RMH_TRY(h->reserve((size_t)n * sizeof(T)));
    T* t = static_cast<T*>(h->ws);
    RMH_TRY(launch_strided_copy<T>(x, incx, t, 1, n, s));
    if (!trans)
        hipLaunchKernelGGL(trmv_n_kernel<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, upper, unit, A, lda, t, x,
                           incx);
    else
        hipLaunchKernelGGL(trmv_t_kernel<T>, dim3((unsigned)n), dim3(256), 0, s, n, upper, unit, A, lda, t, x, incx);
    return check_launch("trmv_kernel");
}

template <typename T>
rmh_status trmm_impl(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                     const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    const int64_t na = left ? m : n;
    const size_t z_el = ((size_t)na * na + 63) / 64 * 64;
    RMH_TRY(h->reserve((z_el + (size_t)m * n) * sizeof(T)));
    T* Z = static_cast<T*>(h->ws);
    T* Bc = Z + z_el;
    hipLaunchKernelGGL(tri_pad_kernel<T>, dim3(blocks_for(na * na, 256)), dim3(256), 0, s, A, lda, na, upper, unit, Z);
    RMH_TRY(check_launch("tri_pad_kernel"));
    RMH_TRY_HIP(hipMemcpy2DAsync(Bc, (size_t)m * sizeof(T), B, (size_t)ldb * sizeof(T), (size_t)m * sizeof(T),
                                 (size_t)n, hipMemcpyDeviceToDevice, s));
    GemmDesc<T> g;
    g.M = m; g.N = n; g.K = na;
    g.alpha = alpha; g.beta = T(0);
    if (left) {
        g.A = Z; g.lda = na; g.transA = trans;
        g.B = Bc; g.ldb = m;
    } else {
        g.A = Bc; g.lda = m;
        g.B = Z; g.ldb = na; g.transB = trans;
    }
    g.C = B; g.ldc = ldb;
    return launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), s);
}

template rmh_status trmv_impl<double>(Context*, bool, bool, bool, int64_t, const double*, int64_t, double*, int64_t,
                                      hipStream_t);
template rmh_status trmv_impl<float>(Context*, bool, bool, bool, int64_t, const float*, int64_t, float*, int64_t,
                                     hipStream_t);
template rmh_status trmm_impl<double>(Context*, bool, bool, bool, bool, int64_t, int64_t, double, const double*,
                                      int64_t, double*, int64_t, hipStream_t);
template rmh_status trmm_impl<float>(Context*, bool, bool, bool, bool, int64_t, int64_t, float, const float*, int64_t,
                                     float*, int64_t, hipStream_t);

}  // namespace rmh
