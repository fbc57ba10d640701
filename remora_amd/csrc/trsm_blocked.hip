// Blocked triangular solve for tall systems with few right-hand sides
// (the two solves of potrs, the refinement corrections).
//
// A recursive solve splits into ever smaller GEMMs whose output is only
// nrhs columns wide: they cannot fill 256 CUs, and hundreds of dependent
// launches dominate.  Here the triangular matrix is cut into nbd-row blocks
// (at n = 16384: 1024 rows in FP64, 2048 in FP32) and the inverses of its diagonal blocks are formed
// first, hierarchically and in parallel: the 64-block inverses in one
// launch, then for s = 64, 128, ... the 2s-block inverse from its halves,
//     inv([P 0; Q R]) = [inv(P) 0; -inv(R) Q inv(P)  inv(R)],
// as two strided-batched GEMMs per level over all 2s-blocks of the matrix.
// The solve is then nbd-blocked substitution with two kernels per block:
//     X_b = inv(T_bb) B_b               (split-K GEMM: kb x nrhs x kb)
//     B_rest -= T_rest,b X_b            (tall GEMM: up to n x nrhs x nbd)
// Using explicit inverses of diagonal blocks (as MAGMA's trsm does) costs
// accuracy on the order of cond(T_bb) u; the Cholesky factors this serves
// have cond(L_bb) <= sqrt(cond(A_bb)).
#include <algorithm>

#include "rmh_context.hpp"

namespace rmh {
namespace {

template <typename T>
struct SView {
    const T* A;
    int64_t lda;
    bool s_is_At;  // S = A^T (upper storage) or S = A (lower storage)
    // S[r0:, c0:] as a GEMM operand (origin pointer, transpose flag)
    void block(int64_t r0, int64_t c0, const T** p, bool* tr) const {
        if (!s_is_At) { *p = A + r0 + c0 * lda; *tr = false; }
        else { *p = A + c0 + r0 * lda; *tr = true; }
    }
};

// Block size of the substitution.  Each block costs four dependent launches
// (two split-K GEMMs and their reductions), so fewer, larger blocks shorten
// the latency-bound sweep, while the inverses grow as nbd^2 per row.  FP32
// (the refinement solves of dsposv, three or four per call) takes 2048-row
// blocks above n = 8192: 1.0-1.4 ms less per mixed solve at n = 16384; FP64
// potrs, built and swept once, gains nothing from them.  Below that, 1024-row
// blocks from n = 2049 on and 512 from 1025 on (FP64 potrs with 64
// right-hand sides: 0.90 -> 0.53 ms at n = 4096, 1.22 -> 0.96 ms at 8192,
// 0.54 -> 0.41 ms at 2048 against 256 / 512 / 256-row blocks).
template <typename T>
int64_t choose_nbd(int64_t na) {
    if (na > 8192) return sizeof(T) == 4 ? 2048 : 1024;
    if (na > 2048) return 1024;
    if (na > 1024) return 512;
    return 128;
}

// D (na x nbd strip, ld na) <- inverses of the nbd diagonal blocks of S.
template <typename T>
rmh_status build_block_inverses(const SView<T>& S, int64_t na, bool unit, int64_t nbd, T* D, T* W, hipStream_t s) {
    RMH_TRY(launch_scale<T>(D, na, na, nbd, T(0), s));
    RMH_TRY(launch_trtri_strip<T>(S.A, S.lda, na, S.s_is_At, unit, D, na, nbd, s));
    for (int64_t h = 64; 2 * h <= nbd; h *= 2) {
        const int64_t ny = nbd / (2 * h);
        const int64_t full = na / (2 * h);  // pairs with both halves complete
        if (full > 0) {
            // a full round of 128 x 128 tiles over the batch takes them (the
            // top level at nbd = 2048); both tile shapes give the same sums
            const Tile lt = (h / 128) * (h / 128) * full >= 512 ? Tile::T128x128 : Tile::T64x64;
            // W_p = Q_p inv(P_p)   (Q_p = S[r0+h : r0+2h, r0 : r0+h], r0 = 2h p)
            GemmDesc<T> g1;
            g1.M = h; g1.N = h; g1.K = h; g1.alpha = T(1); g1.beta = T(0);
            S.block(h, 0, &g1.A, &g1.transA);
            g1.lda = S.lda;
            g1.B = D; g1.ldb = na; g1.transB = false;
            g1.C = W; g1.ldc = h;
            g1.batch = full; g1.batch_ny = ny;
            g1.sAy = 2 * h * (1 + S.lda); g1.sAz = ny * g1.sAy;
            g1.sBy = 2 * h * (1 + na);    g1.sBz = ny * 2 * h;
            g1.sCy = h * h;               g1.sCz = ny * h * h;
            RMH_TRY(launch_gemm<T>(g1, lt, s));
            // X_p = -inv(R_p) W_p  into the strip below inv(P_p)
            GemmDesc<T> g2;
            g2.M = h; g2.N = h; g2.K = h; g2.alpha = T(-1); g2.beta = T(0);
            g2.A = D + h + h * na; g2.lda = na; g2.transA = false;
            g2.B = W; g2.ldb = h; g2.transB = false;
            g2.C = D + h; g2.ldc = na;
            g2.batch = full; g2.batch_ny = ny;
            g2.sAy = 2 * h * (1 + na); g2.sAz = ny * 2 * h;
            g2.sBy = h * h;            g2.sBz = ny * h * h;
            g2.sCy = 2 * h * (1 + na); g2.sCz = ny * 2 * h;
            RMH_TRY(launch_gemm<T>(g2, lt, s));
        }
        const int64_t r0 = full * 2 * h;
        if (na - r0 > h) {  // the last pair has a short lower half
            const int64_t sc = na - r0 - h, o = r0 % nbd;
            GemmDesc<T> g1;
            g1.M = sc; g1.N = h; g1.K = h; g1.alpha = T(1); g1.beta = T(0);
            S.block(r0 + h, r0, &g1.A, &g1.transA);
            g1.lda = S.lda;
            g1.B = D + r0 + o * na; g1.ldb = na;
            g1.C = W; g1.ldc = h;
            RMH_TRY(launch_gemm<T>(g1, Tile::T64x64, s));
            GemmDesc<T> g2;
            g2.M = sc; g2.N = h; g2.K = sc; g2.alpha = T(-1); g2.beta = T(0);
            g2.A = D + (r0 + h) + (o + h) * na; g2.lda = na;
            g2.B = W; g2.ldb = h;
            g2.C = D + (r0 + h) + o * na; g2.ldc = na;
            RMH_TRY(launch_gemm<T>(g2, Tile::T64x64, s));
        }
    }
    return RMH_STATUS_SUCCESS;
}

int split_count(int64_t M, int64_t N, int64_t K) {
    const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
    int64_t sp = (256 + tiles - 1) / tiles;
    sp = std::min<int64_t>(sp, std::max<int64_t>(1, K / 32));
    return (int)std::max<int64_t>(2, std::min<int64_t>(sp, 64));
}

// splits of the tall update (M x N x K, M large): enough 64 x 64 tiles to
// give every compute unit two workgroups
int update_splits(int64_t M, int64_t N, int64_t K) {
    const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
    if (tiles >= 512) return 1;
    int64_t sp = (512 + tiles - 1) / tiles;
    sp = std::min<int64_t>(sp, std::max<int64_t>(1, K / 128));
    return (int)std::max<int64_t>(1, std::min<int64_t>(sp, 8));
}

constexpr int kMaxInvSplit = 64, kMaxUpdSplit = 8;

// One right-hand side: y[0:M] -= op(A) x with op(A) M x K, A column-major
// (TR == false) or its transpose (TR == true: op(A)(i,k) = A[k + i*lda]).  A
// workgroup takes 256 rows and one depth chunk, a thread one row; the
// transposed operand is staged through LDS 16 depths at a time so that 16
// consecutive lanes read 16 consecutive elements.  With several chunks the
// partial sums go to P and an ordered reduction subtracts them.  The MFMA
// GEMM would spend 63 of every 64 columns of its tiles on padding here.
constexpr int kGvRows = 256, kGvStage = 16;

template <typename T, bool TR>
__global__ __launch_bounds__(kGvRows) void gemv_update_kernel(int64_t M, int64_t K, int64_t kc,
                                                              const T* __restrict__ A, int64_t lda,
                                                              const T* __restrict__ x, T* __restrict__ y,
                                                              T* __restrict__ P) {
    __shared__ T stage[TR ? kGvStage : 1][kGvRows + 1];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * kGvRows, i = i0 + tid;
    const int64_t k0 = (int64_t)blockIdx.y * kc, k1 = min(K, k0 + kc);
    const bool ok = i < M;
    T acc = T(0);
    for (int64_t kk = k0; kk < k1; kk += kGvStage) {
        const int kw = (int)min((int64_t)kGvStage, k1 - kk);
        T a[kGvStage];
        if constexpr (!TR) {
#pragma unroll
            for (int s = 0; s < kGvStage; ++s) a[s] = (ok && s < kw) ? A[i + (kk + s) * lda] : T(0);
        } else {
            __syncthreads();
            const int kq = tid % kGvStage, r0 = tid / kGvStage;
#pragma unroll
            for (int s = 0; s < kGvRows / (kGvRows / kGvStage); ++s) {
                const int rr = r0 + s * (kGvRows / kGvStage);
                stage[kq][rr] = (i0 + rr < M && kq < kw) ? A[(kk + kq) + (i0 + rr) * lda] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kGvStage; ++s) a[s] = stage[s][tid];
        }
#pragma unroll
        for (int s = 0; s < kGvStage; ++s)
            if (s < kw) acc += a[s] * x[min(kk + s, K - 1)];
    }
    if (!ok) return;
    if (gridDim.y == 1) y[i] -= acc;
    else P[(int64_t)blockIdx.y * M + i] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_update_reduce_kernel(int64_t M, int splits, const T* __restrict__ P,
                                                                 T* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= M) return;
    T s = T(0);
    for (int sp = 0; sp < splits; ++sp) s += P[(int64_t)sp * M + i];
    y[i] -= s;
}

// y -= op(A) x (see gemv_update_kernel); P holds p_el elements
template <typename T>
rmh_status launch_gemv_update(int64_t M, int64_t K, const T* A, int64_t lda, bool trans, const T* x, T* y, T* P,
                              size_t p_el, hipStream_t s) {
    if (M <= 0 || K <= 0) return RMH_STATUS_SUCCESS;
    const int64_t row_tiles = (M + kGvRows - 1) / kGvRows;
    // about four workgroups per compute unit, within the partial workspace
    int64_t splits = (1024 + row_tiles - 1) / row_tiles;
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, K / 64));
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, (int64_t)(p_el / (size_t)M)));
    int64_t kc = (K + splits - 1) / splits;
    kc = (kc + kGvStage - 1) / kGvStage * kGvStage;
    splits = (K + kc - 1) / kc;
    if (splits > 1 && (size_t)splits * (size_t)M > p_el) return fail(RMH_STATUS_INVALID_VALUE, "trsv: split workspace");
    const dim3 grid((unsigned)row_tiles, (unsigned)splits);
    if (trans)
        hipLaunchKernelGGL((gemv_update_kernel<T, true>), grid, dim3(kGvRows), 0, s, M, K, kc, A, lda, x, y, P);
    else
        hipLaunchKernelGGL((gemv_update_kernel<T, false>), grid, dim3(kGvRows), 0, s, M, K, kc, A, lda, x, y, P);
    RMH_TRY(check_launch("gemv_update_kernel"));
    if (splits == 1) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL(gemv_update_reduce_kernel<T>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, M,
                       (int)splits, P, y);
    return check_launch("gemv_update_reduce_kernel");
}

template <typename T>
size_t partial_elems(int64_t na, int64_t other, int64_t nbd) {
    return std::max((size_t)kMaxInvSplit * nbd * other, (size_t)kMaxUpdSplit * na * other);
}

}  // namespace

bool trsm_blocked_applies(int64_t na, int64_t other) { return na >= 512 && other <= 1024 && other * 4 <= na; }

template <typename T>
size_t trsm_inverse_workspace(int64_t na, int64_t other) {
    const int64_t nbd = choose_nbd<T>(na);
    const size_t d_el = (size_t)na * nbd, w_el = (size_t)na * nbd / 2 + 64 * 64;
    const size_t p_el = partial_elems<T>(na, other, nbd);
    return ((d_el + 63) / 64 * 64 + (w_el + 63) / 64 * 64 + p_el) * sizeof(T) + 3 * 256;
}

template <typename T>
rmh_status diag_block_inverses(const T* A, int64_t lda, int64_t na, bool upper, bool unit, int64_t nbd, T* D, T* W,
                               hipStream_t s) {
    if (nbd < 64 || (nbd & (nbd - 1)) != 0) return fail(RMH_STATUS_INVALID_VALUE, "block inverses: order not 2^k >= 64");
    const SView<T> S{A, lda, upper};
    return build_block_inverses<T>(S, na, unit, nbd, D, W, s);
}

template <typename T>
rmh_status trsm_prepare(const T* A, int64_t lda, int64_t na, bool upper, bool unit, void* ws, TrsmInverses<T>* inv,
                        hipStream_t s) {
    const int64_t nbd = choose_nbd<T>(na);
    const size_t d_el = (size_t)na * nbd, w_el = (size_t)na * nbd / 2 + 64 * 64;
    T* D = static_cast<T*>(ws);
    T* W = D + (d_el + 63) / 64 * 64;
    T* P = W + (w_el + 63) / 64 * 64;
    inv->A = A; inv->lda = lda; inv->na = na; inv->nbd = nbd; inv->upper = upper; inv->unit = unit;
    inv->D = D; inv->P = P;
    const SView<T> S{A, lda, upper};
    return build_block_inverses<T>(S, na, unit, nbd, D, W, s);
}

template <typename T>
rmh_status trsm_apply(const TrsmInverses<T>& inv, bool left, bool trans, int64_t m, int64_t n, T alpha, T* B,
                      int64_t ldb, hipStream_t s) {
    const int64_t na = left ? m : n, other = left ? n : m;
    if (na != inv.na) return fail(RMH_STATUS_INVALID_VALUE, "trsm: prepared inverses of another order");
    const int64_t nbd = inv.nbd;
    const SView<T> S{inv.A, inv.lda, inv.upper};
    const bool eff_upper = inv.upper != trans;  // op(A) = S^T
    const T* D = inv.D;
    T* P = inv.P;
    const size_t p_el = partial_elems<T>(na, other, nbd);
    const int64_t lda = inv.lda;

    if (alpha != T(1)) RMH_TRY(launch_scale<T>(B, ldb, m, n, alpha, s));
    if (alpha == T(0)) return RMH_STATUS_SUCCESS;

    // op(A)[r0:r1, c0:c1] as an operand: S block, or a transposed S block
    auto opA = [&](int64_t r0, int64_t c0, const T** p, bool* tr) {
        if (!eff_upper) S.block(r0, c0, p, tr);
        else { S.block(c0, r0, p, tr); *tr = !*tr; }
    };
    const int64_t nblk = (na + nbd - 1) / nbd;
    // left & lower or right & upper: front to back
    const bool forward = left ? !eff_upper : eff_upper;
    for (int64_t q = 0; q < nblk; ++q) {
        const int64_t b = forward ? q : nblk - 1 - q;
        const int64_t r0 = b * nbd, kb = std::min(nbd, na - r0);
        const T* Db = D + r0;  // inv(S_bb), ld na; inv(op(A)_bb) = eff_upper ? Db^T : Db
        GemmDesc<T> g;
        g.alpha = T(1); g.beta = T(0);
        if (left) {  // X_b = inv(op(A)_bb) B_b
            g.M = kb; g.N = other; g.K = kb;
            g.A = Db; g.lda = na; g.transA = eff_upper;
            g.B = B + r0; g.ldb = ldb;
            g.C = B + r0; g.ldc = ldb;
        } else {     // X_b = B_b inv(op(A)_bb)
            g.M = other; g.N = kb; g.K = kb;
            g.A = B + r0 * ldb; g.lda = ldb;
            g.B = Db; g.ldb = na; g.transB = eff_upper;
            g.C = B + r0 * ldb; g.ldc = ldb;
        }
        // C aliases an input: the split-K path reads everything before its
        // reduction writes C (a single m-tile cannot race either)
        g.splits = split_count(g.M, g.N, g.K);
        g.split_ws = P;
        if ((size_t)g.splits * g.M * g.N > p_el) return fail(RMH_STATUS_INVALID_VALUE, "trsm: split workspace");
        RMH_TRY(launch_gemm<T>(g, Tile::T64x64, s));

        // update the blocks not yet solved
        int64_t u0, u1;  // index range of the rows (left) / columns (right) to update
        if (forward) { u0 = r0 + kb; u1 = na; }
        else { u0 = 0; u1 = r0; }
        if (u1 <= u0) continue;
        GemmDesc<T> u;
        u.alpha = T(-1); u.beta = T(1);
        if (left && other == 1) {  // one right-hand side: a split GEMV
            const T* ap = nullptr;
            bool tr = false;
            opA(u0, r0, &ap, &tr);
            RMH_TRY(launch_gemv_update<T>(u1 - u0, kb, ap, lda, tr, B + r0, B + u0, P, p_el, s));
            continue;
        }
        if (left) {  // B[u0:u1] -= op(A)[u0:u1, r0:r0+kb] X_b
            u.M = u1 - u0; u.N = other; u.K = kb;
            opA(u0, r0, &u.A, &u.transA);
            u.lda = lda;
            u.B = B + r0; u.ldb = ldb;
            u.C = B + u0; u.ldc = ldb;
        } else {     // B[:, u0:u1] -= X_b op(A)[r0:r0+kb, u0:u1]
            u.M = other; u.N = u1 - u0; u.K = kb;
            u.A = B + r0 * ldb; u.lda = ldb;
            opA(r0, u0, &u.B, &u.transB);
            u.ldb = lda;
            u.C = B + u0 * ldb; u.ldc = ldb;
        }
        const Tile ut = pick_tile(u.M, u.N, Part::Full);
        if (ut == Tile::T64x64) {  // C is disjoint from both operands
            u.splits = update_splits(u.M, u.N, u.K);
            if (u.splits > 1) {
                u.split_ws = P;
                if ((size_t)u.splits * u.M * u.N > p_el) return fail(RMH_STATUS_INVALID_VALUE, "trsm: split workspace");
            }
        }
        RMH_TRY(launch_gemm<T>(u, ut, s));
    }
    return RMH_STATUS_SUCCESS;
}

template <typename T>
rmh_status trsm_blocked(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                        const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
    const int64_t na = left ? m : n, other = left ? n : m;
    RMH_TRY(h->reserve(trsm_inverse_workspace<T>(na, other)));
    if (alpha == T(0)) return launch_scale<T>(B, ldb, m, n, T(0), s);
    TrsmInverses<T> inv;
    RMH_TRY(trsm_prepare<T>(A, lda, na, upper, unit, h->ws, &inv, s));
    return trsm_apply<T>(inv, left, trans, m, n, alpha, B, ldb, s);
}

#define RMH_TRSM_INST(T)                                                                                         \
    template size_t trsm_inverse_workspace<T>(int64_t, int64_t);                                                  \
    template rmh_status trsm_prepare<T>(const T*, int64_t, int64_t, bool, bool, void*, TrsmInverses<T>*,          \
                                        hipStream_t);                                                             \
    template rmh_status trsm_apply<T>(const TrsmInverses<T>&, bool, bool, int64_t, int64_t, T, T*, int64_t,       \
                                      hipStream_t);
RMH_TRSM_INST(double)
RMH_TRSM_INST(float)
#undef RMH_TRSM_INST

template rmh_status trsm_blocked<double>(Context*, bool, bool, bool, bool, int64_t, int64_t, double, const double*,
                                         int64_t, double*, int64_t, hipStream_t);
template rmh_status trsm_blocked<float>(Context*, bool, bool, bool, bool, int64_t, int64_t, float, const float*,
                                        int64_t, float*, int64_t, hipStream_t);

template rmh_status diag_block_inverses<double>(const double*, int64_t, int64_t, bool, bool, int64_t, double*, double*,
                                                hipStream_t);
template rmh_status diag_block_inverses<float>(const float*, int64_t, int64_t, bool, bool, int64_t, float*, float*,
                                               hipStream_t);

}  // namespace rmh
