// Rank-one update of a Cholesky factor: the lower factor L of A becomes the
// factor of alpha*A + beta*v*v^T in O(n^2), as
// cholesky_decomposition::update does (include/remora/decompositions.hpp:132-169).
//
// The reference sweeps the columns in order, carrying a vector temp that
// every column j updates below the diagonal:
//     temp_i  -= (w_j / Ljj) * a*L(i,j)            (w_j = temp_j, Ljj = a*L(j,j), a = sqrt(alpha))
//     L(i,j)   = a*L(i,j) * (nLjj / Ljj) + (nLjj*beta*w_j / gamma_j) * temp_i
// The value temp_j reaches at its own column is L(j,j) * p_j with L p = v
// (the recurrence is forward substitution), so the column scalars need only
// p, the diagonal and a prefix sum of beta*w_j^2/Ljj^2.  Each new entry is
// then a_j-scaled old entry plus c_j times the running row sum
// v_i - sum_{k<=j} g_k * a*L(i,k), a scan along each row.  On the device:
//   1. p = L^-1 v on the blocked TRSM;
//   2. every column's scalars (and the first column whose update would make
//      the matrix indefinite), the prefix sum taken over 256-column blocks;
//   3. the strictly lower 256 x 256 tiles write their rows' partial sums;
//   4. every lower tile seeds its rows' running sums from the partial sums of
//      the tiles to its left and rewrites its entries in place.
// Steps 3 and 4 stream the triangle (read, read + write) with one thread per
// row walking 16 columns at a time.  When some column fails, step 2 records
// it in info and steps 3-4 leave L untouched (the reference throws at that
// column after rewriting the ones before it).
#include "rmh_context.hpp"

namespace rmh {
namespace {

constexpr int kRows = 256;  // rows (threads) per tile
constexpr int kCols = 256;  // columns per tile
constexpr int kStage = 16;  // columns staged per step
constexpr int kScalarThreads = 1024;

// tile t of a lower-triangular grid, row-major order: (r, c) with c <= r
__device__ __forceinline__ void tri_tile(int64_t t, int64_t& r, int64_t& c) {
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    r = i;
    c = t - i * (i + 1) / 2;
}

// Column scalars of the sweep, with the reference's formulas; beta' is the
// exclusive prefix sum 1 + sum_{k<j} beta*w_k^2/Ljj_k^2, taken in three
// launches over 256-column blocks (terms and block sums, a scan of the block
// sums, the in-block scan and the scalars).  flag_j = 1 when gamma_j == 0
// (the reference only rescales that column by a).
constexpr int kScanBlock = 256;

template <typename T>
struct ColumnTerm {
    double Ljj, dj, wj, swj2;
    __device__ ColumnTerm(const T* L, int64_t ldl, const T* p, int64_t j, double a, double beta) {
        const double d = (double)L[j + j * ldl];
        Ljj = a * d;
        dj = Ljj * Ljj;
        wj = d * (double)p[j];
        swj2 = beta * wj * wj;
    }
};

template <typename T>
__global__ __launch_bounds__(kScanBlock) void update_terms_kernel(const T* __restrict__ L, int64_t ldl, int64_t n,
                                                                  const T* __restrict__ p, double alpha, double beta,
                                                                  double* __restrict__ term, double* __restrict__ bsum) {
    __shared__ double red[kScanBlock];
    const int t = threadIdx.x;
    const int64_t j = (int64_t)blockIdx.x * kScanBlock + t;
    double x = 0.0;
    if (j < n) {
        const ColumnTerm<T> c(L, ldl, p, j, sqrt(alpha), beta);
        x = c.swj2 / c.dj;
        term[j] = x;
    }
    red[t] = x;
    __syncthreads();
    for (int off = kScanBlock / 2; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    if (t == 0) bsum[blockIdx.x] = red[0];
}

// exclusive scan of the nb block sums in place (one workgroup), and the
// first-failure marker reset
__global__ __launch_bounds__(kScalarThreads) void update_block_scan_kernel(double* __restrict__ bsum, int64_t nb,
                                                                           unsigned long long* __restrict__ first_bad) {
    __shared__ double part[kScalarThreads];
    const int t = threadIdx.x;
    const int64_t chunk = (nb + kScalarThreads - 1) / kScalarThreads;
    const int64_t b0 = min(nb, (int64_t)t * chunk), b1 = min(nb, b0 + chunk);
    double sum = 0.0;
    for (int64_t b = b0; b < b1; ++b) sum += bsum[b];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < kScalarThreads; off <<= 1) {
        const double add = t >= off ? part[t - off] : 0.0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    double run = t > 0 ? part[t - 1] : 0.0;
    for (int64_t b = b0; b < b1; ++b) {
        const double x = bsum[b];
        bsum[b] = run;
        run += x;
    }
    if (t == 0) *first_bad = ~0ull;
}

template <typename T>
__global__ __launch_bounds__(kScanBlock) void update_scalars_kernel(const T* __restrict__ L, int64_t ldl, int64_t n,
                                                                    const T* __restrict__ p, double alpha, double beta,
                                                                    const double* __restrict__ term,
                                                                    const double* __restrict__ bpre,
                                                                    double* __restrict__ g, double* __restrict__ r,
                                                                    double* __restrict__ c, double* __restrict__ dnew,
                                                                    int* __restrict__ flag,
                                                                    unsigned long long* __restrict__ first_bad) {
    __shared__ double sc[kScanBlock];
    const int t = threadIdx.x;
    const int64_t j = (int64_t)blockIdx.x * kScanBlock + t;
    const double x0 = j < n ? term[j] : 0.0;
    sc[t] = x0;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {  // inclusive scan
        const double add = t >= off ? sc[t - off] : 0.0;
        __syncthreads();
        sc[t] += add;
        __syncthreads();
    }
    if (j >= n) return;
    const double bp = 1.0 + bpre[blockIdx.x] + (sc[t] - x0);
    const ColumnTerm<T> cj(L, ldl, p, j, sqrt(alpha), beta);
    const double gamma = cj.dj * bp + cj.swj2;
    const double x = cj.dj + cj.swj2 / bp;
    if (x <= 0.0) atomicMin(first_bad, (unsigned long long)j);
    const double nLjj = sqrt(x);
    g[j] = cj.wj / cj.Ljj;
    r[j] = nLjj / cj.Ljj;
    c[j] = nLjj * beta * cj.wj / gamma;
    dnew[j] = nLjj;
    flag[j] = gamma == 0.0 ? 1 : 0;
}

__global__ void update_info_kernel(const unsigned long long* __restrict__ first_bad, int64_t* __restrict__ info) {
    const unsigned long long f = *first_bad;
    *info = f == ~0ull ? 0 : (int64_t)f + 1;
}

// One 256 x 256 tile of the triangle.  L(i,k) is at L[i + k*ldl] (lower
// storage) or L[k + i*ldl] (upper storage: the factor is U = L^T).
// APPLY == false: strictly lower tiles write S[cb*n + i] = sum over the tile's
// columns of g_k * a*L(i,k).  APPLY == true: the running sum of row i starts at
// v_i - sum_{c < cb} S[c*n + i] and the tile's entries are rewritten.
template <typename T, bool UPPER, bool APPLY>
__global__ __launch_bounds__(kRows) void update_tiles_kernel(T* __restrict__ L, int64_t ldl, int64_t n,
                                                            const T* __restrict__ v, const double* __restrict__ g,
                                                            const double* __restrict__ r, const double* __restrict__ c,
                                                            const double* __restrict__ dnew, const int* __restrict__ flag,
                                                            double a, double* __restrict__ S,
                                                            const int64_t* __restrict__ info) {
    __shared__ T stage[UPPER ? kStage : 1][kRows + 1];
    if (*info != 0) return;
    int64_t rb, cb;
    tri_tile(blockIdx.x, rb, cb);
    if (!APPLY) ++rb;  // strictly lower tiles only
    const int tid = threadIdx.x;
    const int64_t i0 = rb * kRows, k0 = cb * kCols;
    const int64_t i = i0 + tid;
    const bool row_ok = i < n;
    const int64_t kend = min(k0 + kCols, n);
    double q = 0.0;
    if (APPLY && row_ok) {
        q = (double)v[i];
        for (int64_t cc = 0; cc < cb; ++cc) q -= S[cc * n + i];
    }
    // upper storage staging: thread -> column u_kc of the stage and rows
    // u_r0 + kRowStep*s, so that kStage consecutive lanes read kStage
    // consecutive elements of one stored column
    constexpr int kRowStep = kRows / kStage;  // rows one pass of the workgroup covers
    const int u_kc = tid % kStage, u_r0 = tid / kStage;

    for (int64_t kk = k0; kk < kend; kk += kStage) {
        const int kw = (int)min((int64_t)kStage, kend - kk);
        T x[kStage];
        if constexpr (!UPPER) {
#pragma unroll
            for (int s = 0; s < kStage; ++s) {
                const int64_t k = kk + s;
                x[s] = (row_ok && s < kw && k <= i) ? L[i + k * ldl] : T(0);
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kRows / kRowStep; ++s) {
                const int rr = u_r0 + s * kRowStep;
                const int64_t ii = i0 + rr, k = kk + u_kc;
                stage[u_kc][rr] = (ii < n && u_kc < kw && k <= ii) ? L[k + ii * ldl] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kStage; ++s) x[s] = stage[s][tid];
        }
#pragma unroll
        for (int s = 0; s < kStage; ++s) {
            const int64_t k = kk + s;
            if (s < kw && k < i) {
                const double ls = (double)x[s] * a;
                q -= g[k] * ls;
                if (APPLY) x[s] = (T)(flag[k] ? ls : ls * r[k] + c[k] * q);
            } else if (APPLY && s < kw && k == i) {
                x[s] = (T)dnew[k];
            }
        }
        if (!APPLY) continue;
        if constexpr (!UPPER) {
#pragma unroll
            for (int s = 0; s < kStage; ++s) {
                const int64_t k = kk + s;
                if (row_ok && s < kw && k <= i) L[i + k * ldl] = x[s];
            }
        } else {
#pragma unroll
            for (int s = 0; s < kStage; ++s) stage[s][tid] = x[s];
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kRows / kRowStep; ++s) {
                const int rr = u_r0 + s * kRowStep;
                const int64_t ii = i0 + rr, k = kk + u_kc;
                if (ii < n && u_kc < kw && k <= ii) L[k + ii * ldl] = stage[u_kc][rr];
            }
        }
    }
    if (!APPLY && row_ok) S[cb * n + i] = -q;
}

__global__ void set_info_kernel(int64_t* info, int64_t value) { *info = value; }

}  // namespace

template <typename T>
rmh_status chol_update_impl(Context* h, bool upper, int64_t n, T alpha, T beta, T* L, int64_t ldl, const T* v,
                            int64_t incv, int64_t* info, hipStream_t s) {
    if (n == 0 || beta == T(0)) {
        // beta == 0: the reference scales the whole stored matrix by sqrt(alpha)
        if (n > 0) RMH_TRY(launch_scale<T>(L, ldl, n, n, (T)std::sqrt(alpha), s));
        hipLaunchKernelGGL(set_info_kernel, dim3(1), dim3(1), 0, s, info, (int64_t)0);
        return check_launch("set_info_kernel");
    }
    const int64_t nblk = (n + kRows - 1) / kRows;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t vb = al((size_t)n * sizeof(T)), db = al((size_t)n * sizeof(double)), fb = al((size_t)n * sizeof(int));
    const size_t sb = al((size_t)std::max<int64_t>(nblk - 1, 1) * n * sizeof(double));
    const int64_t nsb = (n + kScanBlock - 1) / kScanBlock;
    const size_t bb = al((size_t)nsb * sizeof(double)) + 256;
    void* base = nullptr;
    RMH_TRY(h->reserve_slot(WS_UPDATE, 2 * vb + 5 * db + fb + sb + bb, &base));
    char* w = static_cast<char*>(base);
    T* vin = reinterpret_cast<T*>(w); w += vb;
    T* p = reinterpret_cast<T*>(w); w += vb;
    double* g = reinterpret_cast<double*>(w); w += db;
    double* r = reinterpret_cast<double*>(w); w += db;
    double* c = reinterpret_cast<double*>(w); w += db;
    double* dnew = reinterpret_cast<double*>(w); w += db;
    int* flag = reinterpret_cast<int*>(w); w += fb;
    double* S = reinterpret_cast<double*>(w); w += sb;
    double* term = reinterpret_cast<double*>(w); w += db;
    double* bsum = reinterpret_cast<double*>(w); w += al((size_t)nsb * sizeof(double));
    unsigned long long* first_bad = reinterpret_cast<unsigned long long*>(w);

    RMH_TRY(launch_strided_copy<T>(v, incv, vin, 1, n, s));
    RMH_TRY(launch_strided_copy<T>(v, incv, p, 1, n, s));
    // p = L^-1 v (lower storage) or U^-T v (upper storage)
    RMH_TRY(trsm_impl<T>(h, true, upper, upper, false, n, 1, T(1), L, ldl, p, n, s));
    hipLaunchKernelGGL((update_terms_kernel<T>), dim3((unsigned)nsb), dim3(kScanBlock), 0, s, L, ldl, n, p,
                       (double)alpha, (double)beta, term, bsum);
    RMH_TRY(check_launch("update_terms_kernel"));
    hipLaunchKernelGGL(update_block_scan_kernel, dim3(1), dim3(kScalarThreads), 0, s, bsum, nsb, first_bad);
    RMH_TRY(check_launch("update_block_scan_kernel"));
    hipLaunchKernelGGL((update_scalars_kernel<T>), dim3((unsigned)nsb), dim3(kScanBlock), 0, s, L, ldl, n, p,
                       (double)alpha, (double)beta, term, bsum, g, r, c, dnew, flag, first_bad);
    RMH_TRY(check_launch("update_scalars_kernel"));
    hipLaunchKernelGGL(update_info_kernel, dim3(1), dim3(1), 0, s, first_bad, info);
    RMH_TRY(check_launch("update_info_kernel"));
    const double a = std::sqrt((double)alpha);
    if (nblk > 1) {
        const unsigned tiles = (unsigned)((nblk - 1) * nblk / 2);
        if (upper)
            hipLaunchKernelGGL((update_tiles_kernel<T, true, false>), dim3(tiles), dim3(kRows), 0, s, L, ldl, n, vin, g,
                               r, c, dnew, flag, a, S, info);
        else
            hipLaunchKernelGGL((update_tiles_kernel<T, false, false>), dim3(tiles), dim3(kRows), 0, s, L, ldl, n, vin,
                               g, r, c, dnew, flag, a, S, info);
        RMH_TRY(check_launch("update_tiles_kernel(sums)"));
    }
    const unsigned tiles = (unsigned)(nblk * (nblk + 1) / 2);
    if (upper)
        hipLaunchKernelGGL((update_tiles_kernel<T, true, true>), dim3(tiles), dim3(kRows), 0, s, L, ldl, n, vin, g, r,
                           c, dnew, flag, a, S, info);
    else
        hipLaunchKernelGGL((update_tiles_kernel<T, false, true>), dim3(tiles), dim3(kRows), 0, s, L, ldl, n, vin, g, r,
                           c, dnew, flag, a, S, info);
    return check_launch("update_tiles_kernel(apply)");
}

template rmh_status chol_update_impl<double>(Context*, bool, int64_t, double, double, double*, int64_t, const double*,
                                             int64_t, int64_t*, hipStream_t);
template rmh_status chol_update_impl<float>(Context*, bool, int64_t, float, float, float*, int64_t, const float*,
                                            int64_t, int64_t*, hipStream_t);

}  // namespace rmh
