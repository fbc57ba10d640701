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
// return at once.
//
// The matrix is first made symmetric in both triangles (the other triangle
// is storage the caller does not read; it is cleared at the end, as the
// reference's row clears leave it).  A block then reads A without writing
// it: its row-and-column swaps are kept as a permutation, every raw column
// is a contiguous physical column, and the block's columns go to a panel
// buffer.  Two parallel kernels then put the panel and the (at most 20)
// moved trailing rows and columns in place, and the trailing update covers
// the whole square, as the reference's does.  The row swaps of the columns
// left of the block (finished columns, not read while the block runs) are
// applied after the block by a parallel kernel, in the same order, on a
// second stream beside the next block.  Each thread owns fixed rows of the
// block (k + tid + 1024 u) and keeps their running diagonal values; a
// column's loads are all issued before its stores.
//
// Both storage orders run this lower algorithm: an upper factor U = L^T (the
// reference runs the lower code on trans(A)) is the transpose of the lower
// result, copied into the upper triangle at the end.
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
// Trailing rows per thread and pass (row k + tid + 1024 u), R: 4 fit the
// 128 registers of a 1024-thread workgroup without spilling.  Each pass
// costs a dependent round trip for its swapped-in column, each row one for
// its panel entries (rows past the end skip theirs).  In FP64, blocks with
// more than 2048 trailing rows take four rows in one pass (n = 4096:
// 34.75 -> 34.27 ms), smaller ones two (n = 2000: 11.80 against 12.04 ms
// with four).
constexpr int kPsRowsMax = 4;

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

// ---- the block of 20 columns ----------------------------------------------
// The matrix is kept symmetric in both triangles while the factorisation
// runs, and a block reads it without writing it: the block's row-and-column
// swaps are a permutation perm (logical row -> physical row of A, identity
// at the block's start), so the column that the swap brings into position
// c is physical column perm(p) read at rows perm(t) -- contiguous, where the
// stored triangle alone would need row p, one strided line per entry.  The
// block's columns go to a panel buffer Y by logical row; after the block,
// pstrf_gather_kernel / pstrf_scatter_kernel move the (at most 20) trailing
// rows and columns the permutation touched into place and write the panel,
// and the trailing update runs on the whole square.  Same values, same
// order of operations as swapping in place.

template <typename T>
struct PsLazy {
    T* A;
    int64_t lda;
    int64_t n, k;
    T* Y;  // the block's columns: L(k + r, k + s) = Y[r + s * n]
    T* gpiv;  // running diagonal values (local row r = t - k) when not in LDS
    int* gperm;  // the permutation (local row -> physical row); always exported here
    int64_t *P, *rank, *done;
    int* info;  // [0] columns finished, [1] moved trailing rows (at most 20), [2..] their local indices
    const double* eps_p;  // the stopping threshold (set by pstrf_init_kernel)
    double eps;
};

// per-wavefront maxima (two buffers, by column parity) and, when they fit,
// the running diagonal values and the permutation: LDS variables, so that
// every access is an LDS instruction
template <typename T>
__shared__ T ps_wv[2][kPsWaves];
__shared__ int ps_wi[2][kPsWaves];
__shared__ int ps_cnt;
extern __shared__ unsigned char ps_dyn[];

template <typename T, bool LP>
__device__ __forceinline__ T& lz_piv(const PsLazy<T>& b, int64_t r) {
    if constexpr (LP) return reinterpret_cast<T*>(ps_dyn)[r];
    else return b.gpiv[r];
}
// PL: the block's panel also lives in LDS (m x 20, ld m) behind the running
// values and the permutation, for blocks whose panel fits (m <= ~940 rows in
// FP64): its reads are then LDS reads, the copy in Y is for the scatter
template <typename T>
__device__ __forceinline__ T* lz_pan(int64_t m) {
    return reinterpret_cast<T*>(ps_dyn + ((size_t)m * (sizeof(T) + sizeof(int)) + 15) / 16 * 16);
}
template <typename T, bool LP>
__device__ __forceinline__ int& lz_perm(const PsLazy<T>& b, int64_t m, int64_t r) {
    if constexpr (LP) return reinterpret_cast<int*>(reinterpret_cast<T*>(ps_dyn) + m)[r];
    else return b.gperm[r];
}

// Column c = k + j of the block (local index j).  Every wavefront reduces
// the previous column's per-wavefront maxima to the pivot p (the first
// largest running diagonal value).  Then all reads come first:
//   * lanes s < j hold rows p and c of the block's panel before the swap;
//   * each thread, for its rows t > c, reads the entry that the swap brings
//     into column c, A(perm(t'), perm(p)) with t' the row's position before
//     the swap, and its j panel entries (row c's for the row that becomes p).
// The thread then forms L(t,c) = (raw - sum_s L(t,k+s) L(c,k+s)) / d in the
// order s = 0 .. j-1 and updates its running value.  Rows c and p of the
// panel and the permutation are exchanged after a barrier (every wavefront
// has read them); a second barrier makes them visible to the next column.
// Returns false when the factorisation stops at this column.
template <typename T, bool LP, int R, bool PL>
__device__ __forceinline__ bool pstrf_column(const PsLazy<T>& b, int j, int tid) {
    const int64_t n = b.n, k = b.k, m = n - k, c = k + j;
    const int lane = tid & 63, wave = tid >> 6;
    T v = lane < kPsWaves ? ps_wv<T>[j & 1][lane] : T(0);
    int vi = lane < kPsWaves ? ps_wi[j & 1][lane] : -1;
    ps_wave_max(v, vi, kPsWaves);
    const int64_t p = __builtin_amdgcn_readlane(vi, 0);
    const T pv = ps_lane(v, 0);
    const bool sw = p != c;
    const int64_t pl = p - k;  // local index of the pivot row
    T* const Y = b.Y;
    T* const pan = PL ? lz_pan<T>(m) : nullptr;
    // panel entry (local row r, column s): read from LDS when it is there;
    // written to Y (and LDS)
    auto prd = [&](int64_t r, int s) -> T { return PL ? pan[r + (int64_t)s * m] : Y[r + (int64_t)s * n]; };
    auto pwr = [&](int64_t r, int s, T x) {
        Y[r + (int64_t)s * n] = x;
        if constexpr (PL) pan[r + (int64_t)s * m] = x;
    };
    T rp = T(0), rc = T(0);  // lane s < j: panel rows p and c before the swap
    if (lane < j) {
        rp = prd(pl, lane);
        rc = prd(j, lane);
    }
    if ((double)pv < b.eps) {
        // the swap of rows c and p in the block's columns still applies; the
        // rest of the swap lies in the square that is cleared
        if (wave == 0 && sw && lane < j) {
            pwr(j, lane, rp);
            pwr(pl, lane, rc);
        }
        if (tid == 0) {
            b.P[c] = p;
            *b.rank = c;
            *b.done = 1;
            b.info[0] = j;
        }
        return false;
    }
    const int phys_c = lz_perm<T, LP>(b, m, j), phys_p = lz_perm<T, LP>(b, m, pl);
    const T d = sqrt(pv);
    const T* const colp = b.A + (int64_t)phys_p * b.lda;  // becomes column c
    T nv = T(0);  // this thread's candidate for the next pivot
    int ni = -1;
    for (int64_t q0 = tid; q0 < m; q0 += (int64_t)kPsThreads * R) {
        T raw[R], acc[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t r = q0 + (int64_t)u * kPsThreads;
            const int64_t rl = r < m ? r : m - 1;  // rows past the end read a valid row and store nothing
            raw[u] = colp[sw && rl == pl ? phys_c : lz_perm<T, LP>(b, m, rl)];
            acc[u] = T(0);
        }
        // one row's panel entries in flight at a time (the register budget
        // of 1024 threads); rows past the end skip theirs (whole wavefronts
        // but one), each batch being a dependent round trip
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t r = q0 + (int64_t)u * kPsThreads;
            if (r >= m) continue;
            const int64_t rl = r;
            const int64_t ri = sw && rl == pl ? (int64_t)j : rl;
            T a[kPsBlock - 1];
#pragma unroll
            for (int s = 0; s < kPsBlock - 1; ++s) a[s] = s < j ? prd(ri, s) : T(0);
#pragma unroll
            for (int s = 0; s < kPsBlock - 1; ++s) {
                if (s < j) acc[u] += a[s] * ps_lane(rp, s);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t r = q0 + (int64_t)u * kPsThreads;
            if (!(r > j && r < m)) continue;
            const T y = (raw[u] - acc[u]) / d;
            pwr(r, j, y);
            const T x = lz_piv<T, LP>(b, sw && r == pl ? (int64_t)j : r) - y * y;
            lz_piv<T, LP>(b, r) = x;
            argmax_combine(nv, ni, x, (int)(k + r));
        }
    }
    ps_wave_max(nv, ni, 64);
    if (lane == 0) {
        ps_wv<T>[(j + 1) & 1][wave] = nv;
        ps_wi[(j + 1) & 1][wave] = ni;
    }
    __syncthreads();
    if (wave == 0 && sw && lane < j) {
        pwr(j, lane, rp);
        pwr(pl, lane, rc);
    }
    if (tid == 0) {
        b.P[c] = p;
        pwr(j, j, d);
        if (sw) {
            lz_perm<T, LP>(b, m, j) = phys_p;
            lz_perm<T, LP>(b, m, pl) = phys_c;
        }
    }
    __syncthreads();
    return true;
}

// the columns k .. k+19 (fewer at the end) of the factorisation, one
// workgroup; the running values and the permutation live in LDS when LP
template <typename T, bool LP, int R, bool PL>
__global__ __launch_bounds__(kPsThreads) void pstrf_block_kernel(PsLazy<T> b) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (*b.done != 0) {
        if (tid == 0) b.info[0] = 0;
        return;
    }
    const int64_t n = b.n, k = b.k, m = n - k;
    b.eps = *b.eps_p;
    // running diagonal values restart at the block's first column; the
    // permutation starts as the identity
    T bv = T(0);
    int bi = -1;
    for (int64_t r = tid; r < m; r += kPsThreads) {
        const T x = b.A[(k + r) + (k + r) * b.lda];
        lz_piv<T, LP>(b, r) = x;
        lz_perm<T, LP>(b, m, r) = (int)(k + r);
        argmax_combine(bv, bi, x, (int)(k + r));
    }
    ps_wave_max(bv, bi, 64);
    if (lane == 0) {
        ps_wv<T>[0][wave] = bv;
        ps_wi[0][wave] = bi;
    }
    if (tid == 0) ps_cnt = 0;
    __syncthreads();
    const int cur = (int)min((int64_t)kPsBlock, m);
    for (int j = 0; j < cur; ++j)
        if (!pstrf_column<T, LP, R, PL>(b, j, tid)) return;
    // the permutation of the trailing rows, and the rows it moves
    for (int64_t r = cur + tid; r < m; r += kPsThreads) {
        const int pr = lz_perm<T, LP>(b, m, r);
        if constexpr (LP) b.gperm[r] = pr;
        if (pr != (int)(k + r)) b.info[2 + atomicAdd(&ps_cnt, 1)] = (int)r;
    }
    __syncthreads();
    if (tid == 0) {
        b.info[0] = cur;
        b.info[1] = ps_cnt;
    }
}

// the trailing square's moved columns, gathered before anything is written:
// column q of tmp (rows cur .. m-1 of the block, local) is logical column
// info[2 + q], i.e. physical column perm(i) read at rows perm(r)
template <typename T>
__global__ __launch_bounds__(256) void pstrf_gather_kernel(const T* __restrict__ A, int64_t lda, int64_t n, int64_t k,
                                                          int cur, const int* __restrict__ perm,
                                                          const int* __restrict__ info, T* __restrict__ tmp) {
    if (info[0] != cur) return;  // the factorisation stopped: the trailing square is cleared
    const int q = blockIdx.y;
    if (q >= info[1]) return;
    const int64_t m = n - k, h = m - cur;
    const T* col = A + (int64_t)perm[info[2 + q]] * lda;
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < h; r += (int64_t)gridDim.x * 256)
        tmp[r + q * h] = col[perm[cur + r]];
}

// the block's panel into its columns (lower part), and the gathered columns
// into the trailing square as columns and as rows (the square stays
// symmetric); blockIdx.y < cur: panel column, else moved column y - cur
template <typename T>
__global__ __launch_bounds__(256) void pstrf_scatter_kernel(T* __restrict__ A, int64_t lda, int64_t n, int64_t k, int cur,
                                                           const T* __restrict__ Y, const int* __restrict__ info,
                                                           const T* __restrict__ tmp) {
    const int64_t m = n - k;
    const int y = blockIdx.y;
    if (y < cur) {
        if (y >= info[0]) return;  // columns past a stop
        for (int64_t r = y + (int64_t)blockIdx.x * 256 + threadIdx.x; r < m; r += (int64_t)gridDim.x * 256)
            A[(k + r) + (k + y) * lda] = Y[r + (int64_t)y * n];
        return;
    }
    if (info[0] != cur) return;
    const int q = y - cur;
    if (q >= info[1]) return;
    const int64_t h = m - cur, i = k + info[2 + q];
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < h; r += (int64_t)gridDim.x * 256) {
        const T x = tmp[r + q * h];
        A[(k + cur + r) + i * lda] = x;
        A[i + (k + cur + r) * lda] = x;
    }
}

// one triangle copied onto the other (32 x 32 tiles through LDS): from the
// lower (from_lower) or the upper triangle, strictly off the diagonal
template <typename T>
__global__ __launch_bounds__(256) void pstrf_mirror_kernel(T* __restrict__ A, int64_t lda, int64_t n, bool from_lower) {
    __shared__ T tile[32][33];
    const int64_t t = blockIdx.x;
    int64_t bi = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (bi * (bi + 1) / 2 > t) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    const int64_t bj = t - bi * (bi + 1) / 2;  // lower tile (bi, bj), bi >= bj
    const int64_t R0 = 32 * (from_lower ? bi : bj), C0 = 32 * (from_lower ? bj : bi);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t r = R0 + tx, c = C0 + yy;
        if (r < n && c < n) tile[yy][tx] = A[r + c * lda];
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int64_t sr = R0 + yy, sc = C0 + tx;  // source of destination (sc, sr)
        if (sr < n && sc < n && (from_lower ? sr > sc : sr < sc)) A[sc + sr * lda] = tile[tx][yy];
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
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t b_piv = al(nn * sizeof(T)), b_perm = al(nn * sizeof(int)), b_pan = al(nn * kPsBlock * sizeof(T));
    void* base = nullptr;
    RMH_TRY(h->reserve_slot(WS_PSTRF, b_piv + b_perm + 2 * b_pan + 1024, &base));
    char* w = static_cast<char*>(base);
    T* gpiv = reinterpret_cast<T*>(w);
    int* gperm = reinterpret_cast<int*>(w + b_piv);
    T* Y = reinterpret_cast<T*>(w + b_piv + b_perm);
    T* tmp = reinterpret_cast<T*>(w + b_piv + b_perm + b_pan);
    char* tail = w + b_piv + b_perm + 2 * b_pan;
    double* eps = reinterpret_cast<double*>(tail);
    int64_t* done = reinterpret_cast<int64_t*>(tail + 256);
    int* info = reinterpret_cast<int*>(tail + 512);
    // both storage orders run the lower algorithm on a symmetric working copy
    // (the other triangle is the caller's don't-care storage, cleared at the
    // end); an upper factor is the transpose of the lower one
    const PsView<T> L{A, 1, lda};
    hipLaunchKernelGGL((pstrf_init_kernel<T>), dim3(1), dim3(kPsThreads), 0, s, L, n, piv_out, rank, done, eps);
    RMH_TRY(check_launch("pstrf_init_kernel"));
    if (n == 0) return RMH_STATUS_SUCCESS;
    const int64_t nt = (n + 31) / 32;
    const unsigned mirror_tiles = (unsigned)(nt * (nt + 1) / 2);
    hipLaunchKernelGGL((pstrf_mirror_kernel<T>), dim3(mirror_tiles), dim3(256), 0, s, A, lda, n, !upper);
    RMH_TRY(check_launch("pstrf_mirror_kernel"));
    // The row swaps of a block's columns to the left run on the handle's side
    // stream: nothing later on the caller's stream reads or writes those
    // columns before the final clear, which waits for the side stream.
    hipStream_t side = nullptr;
    hipEvent_t ev_block = nullptr, ev_swaps = nullptr;
    if (n > kPsBlock) RMH_TRY(h->side_stream(&side, &ev_block, &ev_swaps));
    constexpr size_t kLdsBudget = 96 * 1024;        // running values and permutation
    constexpr size_t kLdsPanelBudget = 156 * 1024;  // with the panel as well
    for (int64_t k = 0; k < n; k += kPsBlock) {
        const int64_t m = n - k;
        const int cur = (int)std::min<int64_t>(kPsBlock, m);
        const bool lp = (size_t)m * (sizeof(T) + sizeof(int)) <= kLdsBudget;
        const PsLazy<T> b{A, lda, n, k, Y, gpiv, gperm, piv_out, rank, done, info, eps, 0.0};
        const size_t dyn = lp ? (size_t)m * (sizeof(T) + sizeof(int)) : 0;
        const size_t dyn_pan = (dyn + 15) / 16 * 16 + (size_t)m * kPsBlock * sizeof(T);
        const bool pl = lp && dyn_pan <= kLdsPanelBudget;
        const bool four = sizeof(T) == 4 || m > 2 * kPsThreads;
        if (pl && four)
            hipLaunchKernelGGL((pstrf_block_kernel<T, true, kPsRowsMax, true>), dim3(1), dim3(kPsThreads), dyn_pan, s, b);
        else if (pl)
            hipLaunchKernelGGL((pstrf_block_kernel<T, true, 2, true>), dim3(1), dim3(kPsThreads), dyn_pan, s, b);
        else if (lp && four)
            hipLaunchKernelGGL((pstrf_block_kernel<T, true, kPsRowsMax, false>), dim3(1), dim3(kPsThreads), dyn, s, b);
        else if (lp)
            hipLaunchKernelGGL((pstrf_block_kernel<T, true, 2, false>), dim3(1), dim3(kPsThreads), dyn, s, b);
        else if (four)
            hipLaunchKernelGGL((pstrf_block_kernel<T, false, kPsRowsMax, false>), dim3(1), dim3(kPsThreads), 0, s, b);
        else
            hipLaunchKernelGGL((pstrf_block_kernel<T, false, 2, false>), dim3(1), dim3(kPsThreads), 0, s, b);
        RMH_TRY(check_launch("pstrf_block_kernel"));
        const int64_t h2 = m - cur;
        if (h2 > 0) {
            const unsigned gx = (unsigned)std::min<int64_t>((h2 + 255) / 256, 64);
            hipLaunchKernelGGL((pstrf_gather_kernel<T>), dim3(gx, kPsBlock), dim3(256), 0, s, A, lda, n, k, cur, gperm,
                               info, tmp);
            RMH_TRY(check_launch("pstrf_gather_kernel"));
        }
        const unsigned sx = (unsigned)std::min<int64_t>((m + 255) / 256, 64);
        hipLaunchKernelGGL((pstrf_scatter_kernel<T>), dim3(sx, (unsigned)(cur + (h2 > 0 ? kPsBlock : 0))), dim3(256), 0,
                           s, A, lda, n, k, cur, Y, info, tmp);
        RMH_TRY(check_launch("pstrf_scatter_kernel"));
        if (k > 0) {
            RMH_TRY_HIP(hipEventRecord(ev_block, s));
            RMH_TRY_HIP(hipStreamWaitEvent(side, ev_block, 0));
            hipLaunchKernelGGL((pstrf_row_swaps_kernel<T>), dim3((unsigned)((k + kSwapCols - 1) / kSwapCols)),
                               dim3(64, kSwapCols), 0, side, L, k, cur, piv_out);
            RMH_TRY(check_launch("pstrf_row_swaps_kernel"));
        }
        if (h2 <= 0) continue;
        // trailing square -= L21 L21^T, both triangles (the reference updates
        // the whole square too), so that the next block reads columns only
        GemmDesc<T> g;
        g.part = Part::Full;
        g.M = h2; g.N = h2; g.K = cur;
        g.alpha = T(-1); g.beta = T(1);
        g.A = A + (k + cur) + k * lda; g.transA = false;
        g.B = g.A; g.transB = true;
        g.lda = lda; g.ldb = lda;
        g.C = A + (k + cur) + (k + cur) * lda; g.ldc = lda;
        g.abort_flag = done;
        RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, g.part), s));
    }
    if (side != nullptr) {
        RMH_TRY_HIP(hipEventRecord(ev_swaps, side));
        RMH_TRY_HIP(hipStreamWaitEvent(s, ev_swaps, 0));
    }
    if (upper) {
        hipLaunchKernelGGL((pstrf_mirror_kernel<T>), dim3(mirror_tiles), dim3(256), 0, s, A, lda, n, true);
        RMH_TRY(check_launch("pstrf_mirror_kernel"));
    }
    const int64_t cells = n * n;
    const unsigned grid = (unsigned)std::min<int64_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL((pstrf_clear_kernel<T>), dim3(grid), dim3(256), 0, s, A, lda, n, upper, rank);
    return check_launch("pstrf_clear_kernel");
}

template rmh_status pstrf_impl<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, int64_t*, hipStream_t);
template rmh_status pstrf_impl<float>(Context*, bool, int64_t, float*, int64_t, int64_t*, int64_t*, hipStream_t);

}  // namespace rmh
