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
// loop (each column is O(n) work: pivot update, an argmax reduction, the
// swap, a 20-term GEMV per row), the trailing update is the MFMA GEMM with
// K = 20, and a device flag set at termination makes every later launch
// return at once.  Only the factor's triangle is read and updated: the
// symmetric row-and-column swap is applied to its pieces (rows c and p left
// of c, the two diagonal entries, column c against row p between them, the
// columns below p), which gives the same triangle as the reference's full
// swaps of a symmetric matrix while touching a fraction of the strided
// elements; the reference's row clears and the other triangle are replaced
// by one final clear of that triangle.  The row swaps of the columns left of
// the block (finished columns, not read while the block runs) are applied
// after the block by a parallel kernel, in the same order.  L(i,j) of the lower view is A[i + j*lda] (lower) or
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

// (value, index) maximum with ties to the smaller index: the element
// std::max_element returns (index < 0: no candidate)
template <typename T>
__device__ __forceinline__ void argmax_combine(T& v, int& i, T ov, int oi) {
    if (oi < 0) return;
    if (i < 0 || v < ov || (!(ov < v) && oi < i)) {
        v = ov;
        i = oi;
    }
}

// the columns k .. k+19 (fewer at the end) of the factorisation; the running
// diagonal values live in LDS when lds_piv (n * sizeof(T) <= 64 KB), else in
// the global workspace
template <typename T>
__global__ __launch_bounds__(kPsThreads) void pstrf_block_kernel(PsView<T> L, int64_t n, int64_t k, T* __restrict__ gpiv,
                                                                 int64_t* __restrict__ P, int64_t* __restrict__ rank,
                                                                 int64_t* __restrict__ done,
                                                                 const double* __restrict__ eps_p, int lds_piv) {
    extern __shared__ unsigned char ps_dyn[];
    __shared__ T wv[kPsThreads / 64];
    __shared__ int wi[kPsThreads / 64];
    __shared__ int s_pivot;
    __shared__ T rowc[kPsBlock];  // the pivot row in the block's earlier columns
    if (*done != 0) return;
    T* piv = lds_piv ? reinterpret_cast<T*>(ps_dyn) : gpiv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double eps = *eps_p;
    const int cur = (int)min((int64_t)kPsBlock, n - k);
    // this thread's candidate for the next column's pivot, from the running
    // diagonal values it updated with the column it just computed
    T nv = T(0);
    int ni = -1;
    for (int j = 0; j < cur; ++j) {
        const int64_t c = k + j;
        // running diagonal values: restarted at the block's first column,
        // later columns' values were updated (and their candidates found)
        // where the previous column was computed
        T bv = nv;
        int bi = ni;
        if (j == 0) {
            for (int64_t r = k + tid; r < n; r += kPsThreads) {
                const T x = L(r, r);
                piv[r] = x;
                argmax_combine(bv, bi, x, (int)r);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(bv, off);
            const int oi = __shfl_xor(bi, off);
            argmax_combine(bv, bi, ov, oi);
        }
        if (lane == 0) {
            wv[wave] = bv;
            wi[wave] = bi;
        }
        __syncthreads();
        if (wave == 0) {
            T v = lane < kPsThreads / 64 ? wv[lane] : T(0);
            int i = lane < kPsThreads / 64 ? wi[lane] : -1;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const T ov = __shfl_xor(v, off);
                const int oi = __shfl_xor(i, off);
                argmax_combine(v, i, ov, oi);
            }
            if (lane == 0) s_pivot = i;
        }
        __syncthreads();
        const int64_t p = s_pivot;
        if (p != c) {  // p > c; the four pieces are disjoint
            if (tid == 0) {
                P[c] = p;
                T x = L(c, c);
                L(c, c) = L(p, p);
                L(p, p) = x;
                x = piv[c];
                piv[c] = piv[p];
                piv[p] = x;
            }
            // rows c and p in the block's own columns; the earlier blocks'
            // columns take the block's swaps afterwards (pstrf_row_swaps_kernel)
            for (int64_t t = k + tid; t < c; t += kPsThreads) {
                const T x = L(c, t), y = L(p, t);
                L(c, t) = y;
                L(p, t) = x;
                rowc[t - k] = y;
            }
            for (int64_t t = c + 1 + tid; t < p; t += kPsThreads) {
                const T x = L(t, c);
                L(t, c) = L(p, t);
                L(p, t) = x;
            }
            for (int64_t t = p + 1 + tid; t < n; t += kPsThreads) {
                const T x = L(t, c);
                L(t, c) = L(t, p);
                L(t, p) = x;
            }
        } else if (tid < j) {
            rowc[tid] = L(c, k + tid);
        }
        __syncthreads();
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
        // the block's earlier columns have not updated this one yet; the
        // running diagonal value of each row takes the new entry at once
        nv = T(0);
        ni = -1;
        for (int64_t r = c + 1 + tid; r < n; r += kPsThreads) {
            T acc = T(0);
            for (int t = 0; t < j; ++t) acc += L(r, k + t) * rowc[t];
            const T y = (L(r, c) - acc) / d;
            L(r, c) = y;
            const T x = piv[r] - y * y;
            piv[r] = x;
            argmax_combine(nv, ni, x, (int)r);
        }
        __syncthreads();
    }
}

// the row swaps of block k .. k+cur-1, in order, on the columns before k (not
// read while the block runs).  The block's transpositions are first composed
// into one permutation of the (at most 2 * 20) rows they touch -- wavefront
// 0 keeps the list one row per lane -- and then thread (q, y) of a workgroup
// moves entry q of column 4 * blockIdx.x + y: all loads of a column before a
// barrier, all stores after it, instead of one dependent swap after another.
constexpr int kSwapCols = 4;  // columns per workgroup (64 x 4 threads)
template <typename T>
__global__ __launch_bounds__(64 * kSwapCols) void pstrf_row_swaps_kernel(PsView<T> L, int64_t k, int cur,
                                                                        const int64_t* __restrict__ P) {
    constexpr int R = 2 * kPsBlock;
    __shared__ int64_t rows[R];  // the touched rows
    __shared__ int64_t src[R];   // the row whose entry ends up in rows[q]
    __shared__ int nrows;
    const int lane = threadIdx.x, y = threadIdx.y;
    if (y == 0) {
        int64_t myrow = -1;  // list slot `lane`: a touched row and where its entry comes from
        int myfrom = lane;
        int cnt = 0;
        for (int j = 0; j < cur; ++j) {
            const int64_t c = k + j, p = P[c];
            if (p == c) continue;
            unsigned long long m = __ballot(myrow == c);
            int a = cnt;
            if (m) a = __ffsll((long long)m) - 1;
            else { if (lane == cnt) myrow = c; ++cnt; }
            m = __ballot(myrow == p);
            int b = cnt;
            if (m) b = __ffsll((long long)m) - 1;
            else { if (lane == cnt) myrow = p; ++cnt; }
            const int fa = __shfl(myfrom, a), fb = __shfl(myfrom, b);
            if (lane == a) myfrom = fb;
            if (lane == b) myfrom = fa;
        }
        const int64_t from_row = __shfl(myrow, myfrom);
        if (lane < R) {
            rows[lane] = myrow;
            src[lane] = from_row;
        }
        if (lane == 0) nrows = cnt;
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * kSwapCols + y;
    const bool active = t < k && lane < nrows;
    T v = T(0);
    if (active) v = L(src[lane], t);
    __syncthreads();
    if (active) L(rows[lane], t) = v;
}

// the reference's zeros: the triangle opposite the factor, and after a stop
// at rank r < n the trailing square [r, n) x [r, n)
template <typename T>
__global__ __launch_bounds__(256) void pstrf_clear_kernel(T* __restrict__ A, int64_t lda, int64_t n, bool upper,
                                                         const int64_t* __restrict__ rank) {
    const int64_t r = *rank;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % n, j = e / n;
        const bool other = upper ? i > j : i < j;
        if (other || (i >= r && j >= r)) A[i + j * lda] = T(0);
    }
}

// row interchanges of a pivoted factorisation on an m x ncols matrix B,
// element (i,j) at B[i*rs + j*cs]: swap rows j and piv[j] for j = 0 .. m-1
// (swap_rows) or m-1 .. 0 (swap_rows_inverted); one thread per column
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(T* __restrict__ B, int64_t rs, int64_t cs, int64_t m,
                                                   int64_t ncols, const int64_t* __restrict__ piv, bool inverse) {
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols) return;
    for (int64_t q = 0; q < m; ++q) {
        const int64_t j = inverse ? m - 1 - q : q, p = piv[j];
        if (p == j || p < 0 || p >= m) continue;
        const T x = B[j * rs + col * cs];
        B[j * rs + col * cs] = B[p * rs + col * cs];
        B[p * rs + col * cs] = x;
    }
}

}  // namespace

template <typename T>
rmh_status laswp_impl(bool inverse, int64_t m, int64_t ncols, T* B, int64_t ldb, bool row_major, const int64_t* piv,
                      hipStream_t s) {
    if (m == 0 || ncols == 0) return RMH_STATUS_SUCCESS;
    const int64_t rs = row_major ? ldb : 1, cs = row_major ? 1 : ldb;
    hipLaunchKernelGGL((laswp_kernel<T>), dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s, B, rs, cs, m, ncols,
                       piv, inverse);
    return check_launch("laswp_kernel");
}
template rmh_status laswp_impl<double>(bool, int64_t, int64_t, double*, int64_t, bool, const int64_t*, hipStream_t);
template rmh_status laswp_impl<float>(bool, int64_t, int64_t, float*, int64_t, bool, const int64_t*, hipStream_t);

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
    const size_t piv_bytes = (size_t)n * sizeof(T);
    const int lds_piv = piv_bytes <= 64 * 1024 ? 1 : 0;
    for (int64_t k = 0; k < n; k += kPsBlock) {
        hipLaunchKernelGGL((pstrf_block_kernel<T>), dim3(1), dim3(kPsThreads), lds_piv ? piv_bytes : 0, s, L, n, k, piv,
                           piv_out, rank, done, eps, lds_piv);
        RMH_TRY(check_launch("pstrf_block_kernel"));
        if (k > 0) {
            const int cur = (int)std::min<int64_t>(kPsBlock, n - k);
            hipLaunchKernelGGL((pstrf_row_swaps_kernel<T>), dim3((unsigned)((k + kSwapCols - 1) / kSwapCols)),
                               dim3(64, kSwapCols), 0, s, L, k, cur, piv_out);
            RMH_TRY(check_launch("pstrf_row_swaps_kernel"));
        }
        const int64_t m2 = n - k - kPsBlock;
        if (m2 <= 0) continue;
        // trailing triangle -= L21 L21^T (the reference updates the whole square;
        // only the factor's triangle is read here)
        GemmDesc<T> g;
        g.part = upper ? Part::Upper : Part::Lower;
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
        RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, g.part), s));
    }
    const int64_t cells = n * n;
    const unsigned grid = (unsigned)std::min<int64_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL((pstrf_clear_kernel<T>), dim3(grid), dim3(256), 0, s, A, lda, n, upper, rank);
    return check_launch("pstrf_clear_kernel");
}

template rmh_status pstrf_impl<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, int64_t*, hipStream_t);
template rmh_status pstrf_impl<float>(Context*, bool, int64_t, float*, int64_t, int64_t*, int64_t*, hipStream_t);

}  // namespace rmh
