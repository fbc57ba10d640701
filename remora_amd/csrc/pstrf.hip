// Pivoted Cholesky factorisation P^T A P = L L^T of a symmetric positive
// semi-definite matrix, with the reference's algorithm
// (include/remora/kernels/default/pstrf.hpp:40-160): blocks of 20 columns;
// per column the pivot is the first largest running diagonal value, rows and
// then columns of the whole matrix are swapped, the column takes the block's
// deferred updates and is scaled, and the row right of the diagonal is
// cleared; after each block the whole trailing square takes -= L21 L21^T.
// The factorisation stops when the pivot drops below n^2 * eps * max|A_ii|
// and the trailing square is cleared.
//
// On the device: the 20 columns of a block are one workgroup's sequential
// loop (each column is O(n) work: pivot update, an argmax reduction, two
// swaps, a 20-term GEMV per row), the trailing update is the MFMA GEMM with
// K = 20, and a device flag set at termination makes every later launch
// return at once.  L(i,j) of the lower view is A[i + j*lda] (lower) or
// A[j + i*lda] (upper: the factor U = L^T, as the reference runs the lower
// code on trans(A)).
#include <cfloat>

#include "rmh_context.hpp"

namespace rmh {
namespace {

constexpr int kPsBlock = 20;     // block_size of the reference
constexpr int kPsThreads = 1024;

template <typename T>
struct PsView {
    T* A;
    int64_t rs, cs;  // L(i,j) = A[i*rs + j*cs]
    __device__ T& operator()(int64_t i, int64_t j) const { return A[i * rs + j * cs]; }
};

template <typename T>
constexpr double type_eps() {
    return sizeof(T) == 8 ? DBL_EPSILON : (double)FLT_EPSILON;
}

// P = identity, rank = n, done = 0, eps = n^2 * epsilon * max_diag with
// max_diag = max(A(0,0), |A(i,i)| for i >= 1) as the reference computes it
template <typename T>
__global__ __launch_bounds__(kPsThreads) void pstrf_init_kernel(PsView<T> L, int64_t n, int64_t* __restrict__ P,
                                                                int64_t* __restrict__ rank, int64_t* __restrict__ done,
                                                                double* __restrict__ eps) {
    __shared__ double red[kPsThreads];
    const int tid = threadIdx.x;
    double m = -INFINITY;
    for (int64_t i = tid; i < n; i += kPsThreads) {
        P[i] = i;
        const double d = (double)L(i, i);
        m = fmax(m, i == 0 ? d : fabs(d));
    }
    red[tid] = m;
    __syncthreads();
    for (int off = kPsThreads / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] = fmax(red[tid], red[tid + off]);
        __syncthreads();
    }
    if (tid == 0) {
        *rank = n;
        *done = 0;
        *eps = (double)n * (double)n * type_eps<T>() * red[0];
    }
}

// the columns k .. k+19 (fewer at the end) of the factorisation
template <typename T>
__global__ __launch_bounds__(kPsThreads) void pstrf_block_kernel(PsView<T> L, int64_t n, int64_t k, T* __restrict__ piv,
                                                                 int64_t* __restrict__ P, int64_t* __restrict__ rank,
                                                                 int64_t* __restrict__ done,
                                                                 const double* __restrict__ eps_p) {
    __shared__ T rv[kPsThreads];
    __shared__ int64_t ri[kPsThreads];
    if (*done != 0) return;
    const int tid = threadIdx.x;
    const double eps = *eps_p;
    const int cur = (int)min((int64_t)kPsBlock, n - k);
    for (int j = 0; j < cur; ++j) {
        const int64_t c = k + j;
        // running diagonal values: restarted at the block's first column
        if (j == 0) {
            for (int64_t r = k + tid; r < n; r += kPsThreads) piv[r] = L(r, r);
        } else {
            for (int64_t r = c + tid; r < n; r += kPsThreads) {
                const T x = L(r, c - 1);
                piv[r] -= x * x;
            }
        }
        __syncthreads();
        // first largest of piv[c..n) (std::max_element)
        T bv = T(0);
        int64_t bi = -1;
        for (int64_t r = c + tid; r < n; r += kPsThreads) {
            const T x = piv[r];
            if (bi < 0 || bv < x) {
                bv = x;
                bi = r;
            }
        }
        rv[tid] = bv;
        ri[tid] = bi;
        __syncthreads();
        for (int off = kPsThreads / 2; off > 0; off >>= 1) {
            if (tid < off) {
                const int64_t oi = ri[tid + off];
                if (oi >= 0) {
                    const int64_t mi = ri[tid];
                    const T ov = rv[tid + off], mv = rv[tid];
                    if (mi < 0 || mv < ov || (!(ov < mv) && oi < mi)) {
                        rv[tid] = ov;
                        ri[tid] = oi;
                    }
                }
            }
            __syncthreads();
        }
        const int64_t p = ri[0];
        __syncthreads();
        if (p != c) {
            if (tid == 0) P[c] = p;
            for (int64_t t = tid; t < n; t += kPsThreads) {
                const T x = L(c, t);
                L(c, t) = L(p, t);
                L(p, t) = x;
            }
            __syncthreads();
            for (int64_t t = tid; t < n; t += kPsThreads) {
                const T x = L(t, c);
                L(t, c) = L(t, p);
                L(t, p) = x;
            }
            if (tid == 0) {
                const T x = piv[c];
                piv[c] = piv[p];
                piv[p] = x;
            }
            __syncthreads();
        }
        const T pv = piv[c];
        if ((double)pv < eps) {
            if (tid == 0) {
                *rank = c;
                *done = 1;
            }
            return;
        }
        const T d = sqrt(pv);
        if (tid == 0) L(c, c) = d;
        // the block's earlier columns have not updated this one yet
        for (int64_t r = c + 1 + tid; r < n; r += kPsThreads) {
            T acc = T(0);
            for (int t = 0; t < j; ++t) acc += L(r, k + t) * L(c, k + t);
            L(r, c) = (L(r, c) - acc) / d;
        }
        for (int64_t t = c + 1 + tid; t < n; t += kPsThreads) L(c, t) = T(0);
        __syncthreads();
    }
}

// after a stop at rank r < n: the trailing square [r, n) x [r, n) is zero
template <typename T>
__global__ __launch_bounds__(256) void pstrf_clear_kernel(T* __restrict__ A, int64_t lda, int64_t n,
                                                         const int64_t* __restrict__ rank) {
    const int64_t r = *rank;
    if (r >= n) return;
    const int64_t w = n - r;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < w * w; e += (int64_t)gridDim.x * 256) {
        const int64_t i = r + e % w, j = r + e / w;
        A[i + j * lda] = T(0);
    }
}

}  // namespace

template <typename T>
rmh_status pstrf_impl(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* piv_out, int64_t* rank,
                      hipStream_t s) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    void* base = nullptr;
    RMH_TRY(h->reserve_slot(WS_PSTRF, al((size_t)std::max<int64_t>(n, 1) * sizeof(T)) + 512, &base));
    T* piv = static_cast<T*>(base);
    char* tail = static_cast<char*>(base) + al((size_t)std::max<int64_t>(n, 1) * sizeof(T));
    double* eps = reinterpret_cast<double*>(tail);
    int64_t* done = reinterpret_cast<int64_t*>(tail + 256);
    const PsView<T> L{A, upper ? lda : 1, upper ? 1 : lda};
    hipLaunchKernelGGL((pstrf_init_kernel<T>), dim3(1), dim3(kPsThreads), 0, s, L, n, piv_out, rank, done, eps);
    RMH_TRY(check_launch("pstrf_init_kernel"));
    if (n == 0) return RMH_STATUS_SUCCESS;
    for (int64_t k = 0; k < n; k += kPsBlock) {
        hipLaunchKernelGGL((pstrf_block_kernel<T>), dim3(1), dim3(kPsThreads), 0, s, L, n, k, piv, piv_out, rank, done,
                           eps);
        RMH_TRY(check_launch("pstrf_block_kernel"));
        const int64_t m2 = n - k - kPsBlock;
        if (m2 <= 0) continue;
        // trailing square -= L21 L21^T (both triangles, as the reference's gemm)
        GemmDesc<T> g;
        g.M = m2; g.N = m2; g.K = kPsBlock;
        g.alpha = T(-1); g.beta = T(1);
        if (!upper) {  // L21 = A[k+20:, k:k+20]
            g.A = A + (k + kPsBlock) + k * lda; g.transA = false;
            g.B = g.A; g.transB = true;
        } else {       // L21 = S^T, S = A[k:k+20, k+20:]
            g.A = A + k + (k + kPsBlock) * lda; g.transA = true;
            g.B = g.A; g.transB = false;
        }
        g.lda = lda; g.ldb = lda;
        g.C = A + (k + kPsBlock) + (k + kPsBlock) * lda; g.ldc = lda;
        g.abort_flag = done;
        RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), s));
    }
    const int64_t cells = n * n;
    const unsigned grid = (unsigned)std::min<int64_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL((pstrf_clear_kernel<T>), dim3(grid), dim3(256), 0, s, A, lda, n, rank);
    return check_launch("pstrf_clear_kernel");
}

template rmh_status pstrf_impl<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, int64_t*, hipStream_t);
template rmh_status pstrf_impl<float>(Context*, bool, int64_t, float*, int64_t, int64_t*, int64_t*, hipStream_t);

}  // namespace rmh
