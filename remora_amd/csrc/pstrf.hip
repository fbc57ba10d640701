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
// after the block by a parallel kernel, in the same order, on a second
// stream beside the next block.  Each thread owns fixed rows of the block
// (k + tid + 1024 u) and keeps their running diagonal values; a column's
// loads are all issued before its stores.  L(i,j) of the lower view is
// A[i + j*lda] (lower) or A[j + i*lda] (upper: the factor U = L^T, as the
// reference runs the lower code on trans(A)).
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

// lane s's value (s uniform) to the whole wavefront
template <typename T>
__device__ __forceinline__ T ps_lane(T v, int s) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long b = __double_as_longlong((double)v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, s);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), s);
        return (T)__longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int((float)v), s));
    }
}

constexpr int kPsWaves = kPsThreads / 64;
// trailing rows per thread and pass (row k + tid + 1024 u): what the 128
// registers of a 1024-thread workgroup hold without spilling
template <typename T>
constexpr int ps_rows() { return sizeof(T) == 8 ? 2 : 4; }

// the (value, index) maximum of a wavefront, complete in lane 0
template <typename T>
__device__ __forceinline__ void ps_wave_max(T& v, int& i, int width) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        if (off >= width) continue;
        const T ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(i, off);
        argmax_combine(v, i, ov, oi);
    }
}

// the lower view with its storage order fixed at compile time: L(i,j) is
// A[i + j*ld] (lower) or A[j + i*ld] (upper)
template <typename T, bool UP>
struct PsViewS {
    T* A;
    int64_t ld;
    __device__ T& operator()(int64_t i, int64_t j) const { return UP ? A[j + i * ld] : A[i + j * ld]; }
    __device__ int64_t cs() const { return UP ? 1 : ld; }  // step between columns of a row
};

template <typename T, bool UP>
struct PsBlock {
    PsViewS<T, UP> L;
    int64_t n, k;
    T* gpiv;  // running diagonal values when they do not fit in LDS
    int64_t *P, *rank, *done;
    double eps;
};

// per-wavefront maxima (two buffers, by column parity) and, when they fit,
// the running diagonal values: LDS variables, so that every access is an
// LDS instruction
template <typename T>
__shared__ T ps_wv[2][kPsWaves];
__shared__ int ps_wi[2][kPsWaves];
extern __shared__ unsigned char ps_dyn[];

template <typename T, bool LP>
__device__ __forceinline__ T& ps_piv(const T* gpiv, int64_t i) {
    if constexpr (LP) return reinterpret_cast<T*>(ps_dyn)[i];
    else return const_cast<T*>(gpiv)[i];
}

// Column c = k + j of the block.  Every wavefront reduces the previous
// column's per-wavefront maxima to the pivot p (the first largest running
// diagonal value).  Then all reads come first:
//   * lanes s < j hold row p and row c of the block's earlier columns (the
//     new row c and, for the row that becomes p, its new entries);
//   * each thread, for its rows t > c, reads the entry that the symmetric
//     swap moves into column c (L(p,t) for t < p, L(t,p) for t > p, L(t,c)
//     itself at t = p or without a swap), the old L(t,c) it moves out, and
//     its j block-column entries (all j loads of a row issued together).
// The thread then forms L(t,c) = (raw - sum_s L(t,k+s) L(c,k+s)) / d in the
// order s = 0 .. j-1, updates its running value, and stores.  The block's
// rows c and p are exchanged after a barrier (every wavefront has read
// them), and a second barrier makes the column visible to the next one.
// Returns false when the factorisation stops at this column.
template <typename T, bool UP, bool LP>
__device__ __forceinline__ bool pstrf_column(const PsBlock<T, UP>& b, int j, int tid) {
    const PsViewS<T, UP>& L = b.L;
    const int64_t n = b.n, k = b.k, c = k + j;
    const int lane = tid & 63, wave = tid >> 6;
    T v = lane < kPsWaves ? ps_wv<T>[j & 1][lane] : T(0);
    int vi = lane < kPsWaves ? ps_wi[j & 1][lane] : -1;
    ps_wave_max(v, vi, kPsWaves);
    const int64_t p = __builtin_amdgcn_readlane(vi, 0);
    const T pv = ps_lane(v, 0);
    const bool sw = p != c;
    T rp = T(0), rc = T(0);  // lane s < j: L(p, k+s) and L(c, k+s) before the swap
    if (lane < j) {
        rp = L(p, k + lane);
        rc = L(c, k + lane);
    }
    if ((double)pv < b.eps) {
        // the swap of rows c and p in the block's columns still applies; the
        // rest of the swap lies in the square that is cleared
        if (wave == 0 && sw && lane < j) {
            L(c, k + lane) = rp;
            L(p, k + lane) = rc;
        }
        if (tid == 0) {
            b.P[c] = p;
            *b.rank = c;
            *b.done = 1;
        }
        return false;
    }
    const T d = sqrt(pv);
    T xd = T(0);
    if (tid == 0 && sw) xd = L(c, c);
    T nv = T(0);  // this thread's candidate for the next pivot
    int ni = -1;
    const int64_t cs = L.cs();
    for (int64_t q0 = k + tid; q0 < n; q0 += (int64_t)kPsThreads * ps_rows<T>()) {
        T oc[ps_rows<T>()], raw[ps_rows<T>()], acc[ps_rows<T>()];
#pragma unroll
        for (int u = 0; u < ps_rows<T>(); ++u) {
            const int64_t t = q0 + (int64_t)u * kPsThreads;
            const int64_t r = t < n ? t : n - 1;  // rows past the end read a valid row and store nothing
            oc[u] = L(r, c);
            raw[u] = oc[u];
            if (sw && r != p) raw[u] = r < p ? L(p, r) : L(r, p);
            acc[u] = T(0);
        }
        // one row's block entries in flight at a time (the register budget
        // of 1024 threads; at 4096 rows and j = 19 they are more than the
        // register file, and the reads are bound by the L2)
#pragma unroll
        for (int u = 0; u < ps_rows<T>(); ++u) {
            const int64_t t = q0 + (int64_t)u * kPsThreads;
            const int64_t r = t < n ? t : n - 1;
            const T* row = &L(sw && r == p ? c : r, k);
            T a[kPsBlock - 1];
#pragma unroll
            for (int s = 0; s < kPsBlock - 1; ++s) a[s] = s < j ? row[s * cs] : T(0);
#pragma unroll
            for (int s = 0; s < kPsBlock - 1; ++s) {
                if (s < j) acc[u] += a[s] * ps_lane(rp, s);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int u = 0; u < ps_rows<T>(); ++u) {
            const int64_t t = q0 + (int64_t)u * kPsThreads;
            if (!(t > c && t < n)) continue;
            const T y = (raw[u] - acc[u]) / d;
            if (sw && t != p) {
                if (t < p) L(p, t) = oc[u];
                else L(t, p) = oc[u];
            }
            L(t, c) = y;
            const T x = ps_piv<T, LP>(b.gpiv, t == p ? c : t) - y * y;
            ps_piv<T, LP>(b.gpiv, t) = x;
            argmax_combine(nv, ni, x, (int)t);
        }
    }
    ps_wave_max(nv, ni, 64);
    if (lane == 0) {
        ps_wv<T>[(j + 1) & 1][wave] = nv;
        ps_wi[(j + 1) & 1][wave] = ni;
    }
    __syncthreads();
    if (wave == 0 && sw && lane < j) {
        L(c, k + lane) = rp;
        L(p, k + lane) = rc;
    }
    if (tid == 0) {
        b.P[c] = p;
        if (sw) L(p, p) = xd;
        L(c, c) = d;
    }
    __syncthreads();
    return true;
}

// the columns k .. k+19 (fewer at the end) of the factorisation, one
// workgroup; the running diagonal values live in LDS when lds_piv
// (n * sizeof(T) <= 64 KB), else in the global workspace
template <typename T, bool UP, bool LP>
__global__ __launch_bounds__(kPsThreads) void pstrf_block_kernel(PsViewS<T, UP> L, int64_t n, int64_t k,
                                                                 T* __restrict__ gpiv, int64_t* __restrict__ P,
                                                                 int64_t* __restrict__ rank, int64_t* __restrict__ done,
                                                                 const double* __restrict__ eps_p) {
    if (*done != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const PsBlock<T, UP> b{L, n, k, gpiv, P, rank, done, *eps_p};
    // running diagonal values restart at the block's first column
    T bv = T(0);
    int bi = -1;
    for (int64_t r = k + tid; r < n; r += kPsThreads) {
        const T x = L(r, r);
        ps_piv<T, LP>(gpiv, r) = x;
        argmax_combine(bv, bi, x, (int)r);
    }
    ps_wave_max(bv, bi, 64);
    if (lane == 0) {
        ps_wv<T>[0][wave] = bv;
        ps_wi[0][wave] = bi;
    }
    __syncthreads();
    const int cur = (int)min((int64_t)kPsBlock, n - k);
    for (int j = 0; j < cur; ++j)
        if (!pstrf_column<T, UP, LP>(b, j, tid)) return;
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
    // The row swaps of a block's columns to the left run on the handle's side
    // stream: nothing later on the caller's stream reads or writes those
    // columns before the final clear, which waits for the side stream.
    hipStream_t side = nullptr;
    hipEvent_t ev_block = nullptr, ev_swaps = nullptr;
    if (n > kPsBlock) RMH_TRY(h->side_stream(&side, &ev_block, &ev_swaps));
    for (int64_t k = 0; k < n; k += kPsBlock) {
        const size_t dyn = lds_piv ? piv_bytes : 0;
#define PS_LAUNCH(UP, LP)                                                                                  \
    hipLaunchKernelGGL((pstrf_block_kernel<T, UP, LP>), dim3(1), dim3(kPsThreads), dyn, s, PsViewS<T, UP>{A, lda}, n, \
                       k, piv, piv_out, rank, done, eps)
        if (upper) {
            if (lds_piv) PS_LAUNCH(true, true);
            else PS_LAUNCH(true, false);
        } else {
            if (lds_piv) PS_LAUNCH(false, true);
            else PS_LAUNCH(false, false);
        }
#undef PS_LAUNCH
        RMH_TRY(check_launch("pstrf_block_kernel"));
        if (k > 0) {
            const int cur = (int)std::min<int64_t>(kPsBlock, n - k);
            RMH_TRY_HIP(hipEventRecord(ev_block, s));
            RMH_TRY_HIP(hipStreamWaitEvent(side, ev_block, 0));
            hipLaunchKernelGGL((pstrf_row_swaps_kernel<T>), dim3((unsigned)((k + kSwapCols - 1) / kSwapCols)),
                               dim3(64, kSwapCols), 0, side, L, k, cur, piv_out);
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
    if (side != nullptr) {
        RMH_TRY_HIP(hipEventRecord(ev_swaps, side));
        RMH_TRY_HIP(hipStreamWaitEvent(s, ev_swaps, 0));
    }
    const int64_t cells = n * n;
    const unsigned grid = (unsigned)std::min<int64_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL((pstrf_clear_kernel<T>), dim3(grid), dim3(256), 0, s, A, lda, n, upper, rank);
    return check_launch("pstrf_clear_kernel");
}

template rmh_status pstrf_impl<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, int64_t*, hipStream_t);
template rmh_status pstrf_impl<float>(Context*, bool, int64_t, float*, int64_t, int64_t*, int64_t*, hipStream_t);

}  // namespace rmh
