// Batched Cholesky factorisation and solve of many independent SPD systems
// of uneven order (n_i <= 512) in one persistent launch.
//
// No counterpart in the reference (a new form beside the FP64 drop-in); per
// matrix it runs the same blocked algorithm as rmh_?potrf / rmh_?potrs.
//
// Launch: a persistent grid (CUs x workgroups per CU).  Each workgroup takes
// the next work item from a device counter (in the order a caller-supplied
// permutation gives, e.g. largest first), factors that matrix and solves its
// right-hand sides, and repeats.  The last workgroup to leave resets the
// counters, so consecutive launches (also replays of a captured graph) need
// no memset.
//
// Per matrix (one workgroup, 256 threads), for each NB-wide block column k:
//   1. the diagonal block is staged in LDS (padded with the identity) and
//      factored by wavefront 0 (wave_potrf);
//   2. the rows below are solved against it, one row per thread, and kept in
//      LDS (P) as the operand of
//   3. the trailing update A22 -= P P^T on the lower triangle, 16 x 16 MFMA
//      tiles spread over the four waves, each wave prefetching its next tile.
// In FP64 the forward substitution with L runs inside the factorisation, on
// each block column while its factor is in LDS (diagonal block solved per
// right-hand side, rows below updated one row per thread); the backward
// substitution with L^T follows (rows above updated one row per thread).  In
// FP32 both sweeps follow the factorisation, on 32-row blocks.  A solve with
// a given factor runs the sweeps the same way.  The factorisation and the
// small solves are FP64; the trailing update is on MFMA in T.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "rmh_context.hpp"
#include "wave_chol.hpp"

namespace rmh {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxN = 512;    // largest order of a batched matrix
constexpr int kMaxRhs = 16;   // largest number of right-hand sides

template <typename T>
struct BArgs {
    int64_t batch;
    const int64_t* n;
    T* const* A;
    const int64_t* lda;
    int nrhs;
    T* const* B;
    const int64_t* ldb;
    int64_t* info;
    const int64_t* order;
    unsigned* counters;  // [0] next work item, [1] workgroups done
    int factor;          // 1: factor A in place; 0: A already holds the factor
};

template <typename T>
struct MfmaT;
template <>
struct MfmaT<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane l, item r holds element (row, col) of the 16 x 16 result
    static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};
template <>
struct MfmaT<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// (ti, tj), ti >= tj, of lower-triangle tile t in row order
__device__ __forceinline__ void tri_tile(int t, int& ti, int& tj) {
    int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    ti = i;
    tj = t - i * (i + 1) / 2;
}

template <typename T, int NB>
struct Smem {
    static constexpr int LD = NB + 1;
    double L[NB * LD];        // diagonal block / its factor
    double rdiag[NB];
    T P[kMaxN * LD];          // solved rows below the diagonal block
    T P2[NB == 16 ? kMaxN * LD : 1];  // NB 16: the previous block's rows (deferred update)
    double X[NB * (kMaxRhs + 1)];  // solution block of the solve
    int bad;
    int64_t item;
};

// The workgroup's LDS.  The per-system routines reach it through this
// variable, so the accesses stay LDS instructions also where they are not
// inlined into the persistent loop.
template <typename T, int NB>
__shared__ Smem<T, NB> g_bsm;

template <int NB, int LD, int XL>
__device__ __forceinline__ void diag_solve(const double* L, double* X, int nrhs, bool forward, int tid);

// Block rows of the substitution sweeps of the solve (see solve_one_impl),
// and whether the forward sweep runs inside the factorisation instead, on
// each block column while its rows are in LDS (FP64: 12.5 -> 11.0 ms when
// introduced; FP32 is faster with both sweeps on 32-row blocks after it).
template <typename T, int NB>
constexpr int batched_solve_block() { return sizeof(T) == 8 ? NB : 32; }
template <typename T>
constexpr bool batched_fused_forward() { return sizeof(T) == 8; }

// Factor the n x n matrix A (column-major, ld) in place.  Returns 0 or the
// LAPACK info (k+1 of the first failing pivot).  Uniform across the block.
// With nrhs > 0 it also runs the forward substitution L Y = B in place in B
// (n x nrhs, ld ldb), block column by block column while the column's
// factor is still in LDS, in the order and arithmetic of solve_one's
// forward sweep.
template <typename T, int NB>
__device__ __forceinline__ int64_t factor_one_impl(T* A, int n, int64_t lda, T* B, int nrhs, int64_t ldb) {
    Smem<T, NB>& s = g_bsm<T, NB>;
    const int tid = threadIdx.x;
    typedef MfmaT<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int LD = Smem<T, NB>::LD;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    // NB 16 defers every other block's trailing update: block k's rows stay
    // in LDS, block k+16's column takes their update alone, and after block
    // k+16 is factored one rank-32 pass updates the rest with both -- the
    // trailing triangle is read and written half as often.
    int cur = 0;           // buffer holding this block's rows (0: P, 1: P2)
    bool pending = false;  // the other buffer holds the previous block's rows, not yet applied
    auto Pbuf = [&](int b, int idx) -> T& { return (NB == 16 && b) ? s.P2[idx] : s.P[idx]; };
    for (int k = 0; k < n; k += NB) {
        const int kb = min(NB, n - k), m = n - k - kb;
        if (NB == 16 && pending) {
            // this block column (rows k.., its kb columns) takes the previous
            // block's update: tiles (ti, 0), K = 16.  The previous block's row
            // r is matrix row k + r.
            const int prv = cur ^ 1;
            const int rows = n - k, tm = (rows + 15) / 16;
            T* Ak = A + k + (int64_t)k * lda;
            for (int t = wave; t < tm; t += 4) {
                T c4[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * t + li, j = M::row(lane, r);
                    c4[r] = (i < rows && j < kb && i >= j) ? Ak[i + (int64_t)j * lda] : T(0);
                }
                acc_t acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const T a = Pbuf(prv, (16 * t + li) * LD + 4 * q + lk);
                    const T b = Pbuf(prv, li * LD + 4 * q + lk);
                    acc = M::op(b, a, acc);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * t + li, j = M::row(lane, r);
                    if (i < rows && j < kb && i >= j) Ak[i + (int64_t)j * lda] = c4[r] - acc[r];
                }
            }
            __syncthreads();
        }
        for (int e = tid; e < NB * NB; e += kThreads) {
            const int i = e % NB, c = e / NB;
            s.L[i * LD + c] = (i < kb && c <= i) ? (double)A[(k + i) + (int64_t)(k + c) * lda] : (i == c ? 1.0 : 0.0);
        }
        __syncthreads();
        if (tid < NB) {
            const int bad = wave_potrf<NB, LD>(s.L, s.rdiag, tid);
            if (tid == 0) s.bad = bad;
        }
        __syncthreads();
        if (s.bad >= 0) return (int64_t)k + s.bad + 1;
        for (int e = tid; e < kb * kb; e += kThreads) {
            const int i = e % kb, c = e / kb;
            if (c <= i) A[(k + i) + (int64_t)(k + c) * lda] = (T)s.L[i * LD + c];
        }
        // rows below: x L^T = a, one row per thread; padded columns come out 0
        for (int r = tid; r < m; r += kThreads) {
            const T* row = A + (k + kb + r) + (int64_t)k * lda;
            double x[NB];
            // unconditional loads from clamped columns; the padding is zeroed after
#pragma unroll
            for (int c = 0; c < NB; ++c) x[c] = (double)row[(int64_t)min(c, kb - 1) * lda];
#pragma unroll
            for (int c = 0; c < NB; ++c) x[c] = c < kb ? x[c] : 0.0;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                double sum[4] = {x[c], 0.0, 0.0, 0.0};
#pragma unroll
                for (int t = 0; t < c; ++t) sum[t & 3] = fma(-x[t], s.L[c * LD + t], sum[t & 3]);
                x[c] = ((sum[0] + sum[1]) + (sum[2] + sum[3])) * s.rdiag[c];
                __builtin_amdgcn_sched_barrier(0);  // bounds the LDS loads in flight (registers)
            }
            // the store addresses are formed afresh: kept from the loads, the
            // NB 64-bit addresses would stay live across the solve
            int rr = r;
            asm volatile("" : "+v"(rr));
            T* wrow = A + (k + kb + rr) + (int64_t)k * lda;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < kb) wrow[(int64_t)c * lda] = (T)x[c];
                Pbuf(cur, r * LD + c) = (T)x[c];
            }
        }
        __syncthreads();
        if (nrhs > 0) {
            // forward substitution of this block: Y_k = L_kk^-1 B_k, then
            // B_rest -= L_rest,k Y_k with L_rest,k the rows in P
            constexpr int XL = kMaxRhs + 1;
            for (int e = tid; e < NB * nrhs; e += kThreads) {
                const int i = e % NB, j = e / NB;
                s.X[i * XL + j] = (i < kb) ? (double)B[(k + i) + (int64_t)j * ldb] : 0.0;
            }
            __syncthreads();
            diag_solve<NB, LD, XL>(s.L, s.X, nrhs, true, tid);
            __syncthreads();
            for (int e = tid; e < kb * nrhs; e += kThreads) {
                const int i = e % kb, j = e / kb;
                B[(k + i) + (int64_t)j * ldb] = (T)s.X[i * XL + j];
            }
            for (int r = tid; r < m; r += kThreads) {
                for (int j = 0; j < nrhs; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) acc = fma((double)Pbuf(cur, r * LD + c), s.X[c * XL + j], acc);
                    T* b = B + (k + kb + r) + (int64_t)j * ldb;
                    *b = (T)((double)*b - acc);
                }
            }
        }
        if (NB == 16 && !pending && m > 0) {
            // defer: the next block column takes this block's update first
            pending = true;
            cur ^= 1;
            __syncthreads();
            continue;
        }
        // (the MFMAs take the column tile first, so a lane's items are four
        // columns of one row and 16 lanes cover 16 consecutive rows of a
        // column: C moves in 128-byte column runs)
        // trailing update of the lower triangle of A22 (m x m): each wave
        // takes groups of G tiles and issues all their C loads before the
        // MFMAs, so G tiles' worth of loads are in flight at once; with a
        // deferred previous block its rows from 16 on (matrix rows k + 16 ..)
        // add a second rank-16 term
        const bool two = NB == 16 && pending;
        const int prv = cur ^ 1;
        constexpr int G = 8;
        const int tm = (m + 15) / 16, ntiles = tm * (tm + 1) / 2;
        T* A22 = A + (k + kb) + (int64_t)(k + kb) * lda;
        for (int t0 = wave * G; t0 < ntiles; t0 += 4 * G) {
            T c[G][4];
            int ti[G], tj[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                tri_tile(min(t0 + g, ntiles - 1), ti[g], tj[g]);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * ti[g] + li, j = 16 * tj[g] + M::row(lane, r);
                    c[g][r] = (t0 + g < ntiles && i < m && j < m && i >= j) ? A22[i + (int64_t)j * lda] : T(0);
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (t0 + g >= ntiles) break;
                acc_t acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int q = 0; q < NB / 4; ++q) {
                    const T a = Pbuf(cur, (16 * ti[g] + li) * LD + 4 * q + lk);
                    const T b = Pbuf(cur, (16 * tj[g] + li) * LD + 4 * q + lk);
                    acc = M::op(b, a, acc);
                }
                if (two) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const T a = Pbuf(prv, (16 + 16 * ti[g] + li) * LD + 4 * q + lk);
                        const T b = Pbuf(prv, (16 + 16 * tj[g] + li) * LD + 4 * q + lk);
                        acc = M::op(b, a, acc);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * ti[g] + li, j = 16 * tj[g] + M::row(lane, r);
                    if (i < m && j < m && i >= j) A22[i + (int64_t)j * lda] = c[g][r] - acc[r];
                }
            }
        }
        pending = false;
        __syncthreads();
    }
    return 0;
}

// Solve L_kk Y = X (forward) or L_kk^T Y = X (backward) in place in X (LDS,
// NB x nrhs, ld XL) with L_kk in LDS (identity-padded).  Thread t handles row
// t % NB of right-hand side t / NB; the solved entry of each step moves to
// the other rows of its right-hand side by a lane shuffle (the NB rows of a
// right-hand side sit in one wavefront), so a step costs a shuffle and an FMA.
template <int NB, int LD, int XL>
__device__ __forceinline__ void diag_solve(const double* L, double* X, int nrhs, bool forward, int tid) {
    constexpr int R = kThreads / NB;  // right-hand sides per pass
    const int lane = tid & 63, row = tid % NB, base = (lane / NB) * NB;
    for (int j0 = 0; j0 < nrhs; j0 += R) {
        const int j = j0 + tid / NB;
        const bool ok = j < nrhs;
        double v = ok ? X[row * XL + j] : 0.0;
        if (forward) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (row == c) v = v / L[c * LD + c];
                const double y = __shfl(v, base + c, 64);
                if (row > c) v = fma(-L[row * LD + c], y, v);
            }
        } else {
#pragma unroll
            for (int c = NB - 1; c >= 0; --c) {
                if (row == c) v = v / L[c * LD + c];
                const double x = __shfl(v, base + c, 64);
                if (row < c) v = fma(-L[c * LD + row], x, v);
            }
        }
        if (ok) X[row * XL + j] = v;
    }
}

// Solve L L^T X = B in place (B column-major n x nrhs, ld ldb); with
// forward_done B already holds Y, L Y = B, and only L^T X = Y is solved.
template <typename T, int NB>
__device__ __forceinline__ void solve_one_impl(const T* A, int n, int64_t lda, T* B, int nrhs, int64_t ldb,
                                               bool forward_done) {
    Smem<T, NB>& s = g_bsm<T, NB>;
    const int tid = threadIdx.x;
    constexpr int XL = kMaxRhs + 1;
    // the right-hand sides stay in LDS for both sweeps, in the P buffer the
    // factorisation no longer needs (n x (nrhs + 1) <= kMaxN x (NB + 1)
    // elements), with the same rounding to T after every update as in
    // global memory
    T* Y = s.P;
    const int YL = nrhs + 1;
    for (int e = tid; e < n * nrhs; e += kThreads) {
        const int i = e % n, j = e / n;
        Y[i * YL + j] = B[i + (int64_t)j * ldb];
    }
    __syncthreads();
    // Both sweeps run on SB-row blocks.  A block step costs two memory round
    // trips and three barriers whatever its width, so FP32 takes 32 rows
    // (half the steps of NB = 16: 7.55 -> 6.83 ms on the 4096-system
    // benchmark, with the forward sweep here rather than inside the
    // factorisation); FP64 keeps the factorisation's NB (32 rows measured
    // 9.1-9.4 against 8.5 ms there: the out-of-line routine pays for the
    // wider row registers).  The block and its right-hand sides live in the
    // factorisation's L and X when SB = NB, in its deferred-update buffer P2
    // (free once the factor is done) otherwise.
    constexpr int SB = batched_solve_block<T, NB>(), SLD = SB + 1;
    static_assert(SB == NB || sizeof(s.P2) >= sizeof(double) * SB * (SLD + XL), "solve buffers");
    double* const SL = SB == NB ? s.L : reinterpret_cast<double*>(s.P2);
    double* const SX = SB == NB ? s.X : reinterpret_cast<double*>(s.P2) + SB * SLD;
    // diagonal block of SB rows at k (identity-padded) into SL, its
    // right-hand sides into SX
    auto load_diag = [&](int k, int kb) {
        for (int e = tid; e < SB * SB; e += kThreads) {
            const int i = e % SB, c = e / SB;
            SL[i * SLD + c] = (i < kb && c <= i) ? (double)A[(k + i) + (int64_t)(k + c) * lda] : (i == c ? 1.0 : 0.0);
        }
        for (int e = tid; e < SB * nrhs; e += kThreads) {
            const int i = e % SB, j = e / SB;
            SX[i * XL + j] = (i < kb) ? (double)Y[(k + i) * YL + j] : 0.0;
        }
    };
    const int nblk = (n + SB - 1) / SB;
    // forward: L y = b
    for (int q = 0; q < nblk && !forward_done; ++q) {
        const int k = q * SB, kb = min(SB, n - k), m = n - k - kb;
        load_diag(k, kb);
        __syncthreads();
        diag_solve<SB, SLD, XL>(SL, SX, nrhs, true, tid);  // L_kk y = b
        __syncthreads();
        for (int e = tid; e < kb * nrhs; e += kThreads) {
            const int i = e % kb, j = e / kb;
            Y[(k + i) * YL + j] = (T)SX[i * XL + j];
        }
        for (int r = tid; r < m; r += kThreads) {  // b_r -= L[r, blk] y_blk
            const T* lrow = A + (k + kb + r) + (int64_t)k * lda;
            double l[SB];
#pragma unroll
            for (int c = 0; c < SB; ++c) l[c] = (double)lrow[(int64_t)min(c, kb - 1) * lda];
#pragma unroll
            for (int c = 0; c < SB; ++c) l[c] = c < kb ? l[c] : 0.0;
            for (int j = 0; j < nrhs; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < SB; ++c) acc = fma(l[c], SX[c * XL + j], acc);
                T* y = &Y[(k + kb + r) * YL + j];
                *y = (T)((double)*y - acc);
            }
        }
        __syncthreads();
    }
    // backward: L^T x = y, last block first
    for (int q = nblk - 1; q >= 0; --q) {
        const int k = q * SB, kb = min(SB, n - k);
        load_diag(k, kb);
        __syncthreads();
        diag_solve<SB, SLD, XL>(SL, SX, nrhs, false, tid);  // L_kk^T x = y
        __syncthreads();
        for (int e = tid; e < kb * nrhs; e += kThreads) {
            const int i = e % kb, j = e / kb;
            Y[(k + i) * YL + j] = (T)SX[i * XL + j];
        }
        for (int r = tid; r < k; r += kThreads) {  // y_r -= L[blk, r]^T x_blk
            const T* lcol = A + k + (int64_t)r * lda;  // L(k + c, r), contiguous in c
            double l[SB];
#pragma unroll
            for (int c = 0; c < SB; ++c) l[c] = (double)lcol[min(c, kb - 1)];
#pragma unroll
            for (int c = 0; c < SB; ++c) l[c] = c < kb ? l[c] : 0.0;
            for (int j = 0; j < nrhs; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < SB; ++c) acc = fma(l[c], SX[c * XL + j], acc);
                T* y = &Y[r * YL + j];
                *y = (T)((double)*y - acc);
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * nrhs; e += kThreads) {
        const int i = e % n, j = e / n;
        B[i + (int64_t)j * ldb] = Y[i * YL + j];
    }
    __syncthreads();
}

// Out-of-line copies: each routine then gets its own register allocation
// and the persistent loop keeps almost nothing live across the calls.
template <typename T, int NB>
__device__ __attribute__((noinline)) int64_t factor_one(T* A, int n, int64_t lda, T* B, int nrhs, int64_t ldb) {
    return factor_one_impl<T, NB>(A, n, lda, B, nrhs, ldb);
}
template <typename T, int NB>
__device__ __attribute__((noinline)) void solve_one(const T* A, int n, int64_t lda, T* B, int nrhs, int64_t ldb,
                                                    bool forward_done) {
    solve_one_impl<T, NB>(A, n, lda, B, nrhs, ldb, forward_done);
}

// The persistent loop.  FP64 calls the out-of-line routines (about 390
// registers, one workgroup per CU).  FP32 inlines them into a kernel held
// to 256 registers, two workgroups per CU: a few spills, but 10.5 against
// 13.3 ms on the 4096-system benchmark.  FP64 cannot run two per CU with
// the deferred update (its two panel buffers take 139 KB of LDS); the
// inlined NB 16 kernel without it (74 KB, 115 registers spilled) took
// 11.1 ms against 8.5 ms for the deferred one per CU.
template <typename T, int NB, bool INL>
__device__ __forceinline__ void batched_posv_body(const BArgs<T>& a) {
    Smem<T, NB>& s = g_bsm<T, NB>;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s.item = (int64_t)atomicAdd(&a.counters[0], 1u);
        __syncthreads();
        const int64_t w = s.item;
        __syncthreads();
        if (w >= a.batch) break;
        const int64_t mat = a.order != nullptr ? a.order[w] : w;
        const int64_t n64 = a.n[mat];
        int64_t info = 0;
        if (n64 < 0 || n64 > kMaxN) {
            info = -1;
        } else if (n64 > 0) {
            const int n = (int)n64;
            if (a.factor) {
                // the forward substitution runs inside the factorisation
                constexpr bool fuse = batched_fused_forward<T>();
                T* const b = a.nrhs > 0 && fuse ? a.B[mat] : nullptr;
                const int nr = fuse ? a.nrhs : 0;
                const int64_t ldb = a.nrhs > 0 ? a.ldb[mat] : 1;
                if constexpr (INL) info = factor_one_impl<T, NB>(a.A[mat], n, a.lda[mat], b, nr, ldb);
                else info = factor_one<T, NB>(a.A[mat], n, a.lda[mat], b, nr, ldb);
            }
            if (info == 0 && a.nrhs > 0) {
                const bool fwd = a.factor != 0 && batched_fused_forward<T>();
                if constexpr (INL) solve_one_impl<T, NB>(a.A[mat], n, a.lda[mat], a.B[mat], a.nrhs, a.ldb[mat], fwd);
                else solve_one<T, NB>(a.A[mat], n, a.lda[mat], a.B[mat], a.nrhs, a.ldb[mat], fwd);
            }
        }
        if (tid == 0 && a.info != nullptr) a.info[mat] = info;
    }
    // the last workgroup out resets the counters for the next launch
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&a.counters[1], 1u) == gridDim.x - 1) {
            atomicExch(&a.counters[0], 0u);
            atomicExch(&a.counters[1], 0u);
        }
    }
}

template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void batched_posv_kernel(BArgs<T> a) {
    batched_posv_body<T, NB, false>(a);
}
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void batched_posv_kernel_inl(BArgs<T> a) {
    batched_posv_body<T, NB, true>(a);
}
// the kernel of a precision and block size
template <typename T, int NB>
constexpr auto batched_kernel() {
    if constexpr (sizeof(T) == 4) return batched_posv_kernel_inl<T, NB>;
    else return batched_posv_kernel<T, NB>;
}

template <typename T, int NB>
rmh_status launch_batched(const BArgs<T>& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((batched_kernel<T, NB>()), dim3((unsigned)grid), dim3(kThreads), 0, s, a);
    return check_launch("batched_posv_kernel");
}

template <typename T>
int max_wg_per_cu(int nb) {
    int blocks = 0;
    hipError_t e = nb == 16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batched_kernel<T, 16>(), kThreads, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batched_kernel<T, 32>(), kThreads, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return std::max(1, blocks);
}

// synthetic SPD matrices for tuning: a_ij = u(i, j) in [0, 1) symmetric,
// a_ii = n + 1 (diagonally dominant)
template <typename T>
__global__ void synth_spd_kernel(T* const* A, const int64_t* n, const int64_t* lda, int64_t batch) {
    for (int64_t mat = blockIdx.y; mat < batch; mat += gridDim.y) {
    const int64_t nn = n[mat];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn * nn; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % nn, j = e / nn;
        const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        const uint64_t key = (uint64_t)(mat * 1000003 + lo * 7919 + hi);
        uint64_t z = key + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        A[mat][i + j * lda[mat]] = (T)(i == j ? (double)nn + 1.0 : u);
    }
    }
}

}  // namespace

rmh_status Context::batch_counters(unsigned** c) {
    if (batch_ctr == nullptr) {
        RMH_TRY_HIP(hipMalloc(&batch_ctr, 256));
        RMH_TRY_HIP(hipMemset(batch_ctr, 0, 256));  // once, outside any stream capture
        RMH_TRY_HIP(hipDeviceSynchronize());
    }
    *c = static_cast<unsigned*>(batch_ctr);
    return RMH_STATUS_SUCCESS;
}

template <typename T>
rmh_status batched_posv_impl(Context* h, int64_t batch, int max_n, const int64_t* n, T* const* A, const int64_t* lda,
                             int64_t nrhs, T* const* B, const int64_t* ldb, int64_t* info, const int64_t* order,
                             const rmh_batch_config* cfg, bool factor, hipStream_t s) {
    if (batch == 0) return RMH_STATUS_SUCCESS;
    if (max_n < 0 || max_n > kMaxN) return fail(RMH_STATUS_INVALID_VALUE, "batched: max_n must be in [0, 512]");
    if (nrhs < 0 || nrhs > kMaxRhs) return fail(RMH_STATUS_INVALID_VALUE, "batched: nrhs must be in [0, 16]");
    if (nrhs > 0 && (B == nullptr || ldb == nullptr))
        return fail(RMH_STATUS_INVALID_VALUE, "batched: right-hand sides missing");
    rmh_batch_config c = cfg != nullptr ? *cfg : h->batch_cfg[sizeof(T) == 8 ? 1 : 0];
    if (c.nb != 16 && c.nb != 32) return fail(RMH_STATUS_INVALID_VALUE, "batched: block size must be 16 or 32");
    const int occ = max_wg_per_cu<T>(c.nb);
    const int per_cu = c.wg_per_cu > 0 ? std::min(c.wg_per_cu, occ) : occ;
    int cus = 0;
    RMH_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    const int grid = (int)std::min<int64_t>(batch, (int64_t)cus * per_cu);
    unsigned* ctr = nullptr;
    RMH_TRY(h->batch_counters(&ctr));
    BArgs<T> a{batch, n, A, lda, (int)nrhs, B, ldb, info, order, ctr, factor ? 1 : 0};
    return c.nb == 16 ? launch_batched<T, 16>(a, grid, s) : launch_batched<T, 32>(a, grid, s);
}

// Time each candidate configuration on synthetic matrices of the caller's
// sizes; keep the fastest in the handle.
template <typename T>
rmh_status batched_autotune_impl(Context* h, int64_t batch, const int64_t* h_n, int64_t nrhs, rmh_batch_config* best,
                                 double* best_ms) {
    if (batch <= 0) return fail(RMH_STATUS_INVALID_VALUE, "autotune: empty batch");
    std::vector<int64_t> off(batch + 1, 0), ld(batch), boff(batch + 1, 0);
    for (int64_t i = 0; i < batch; ++i) {
        if (h_n[i] < 0 || h_n[i] > kMaxN) return fail(RMH_STATUS_INVALID_VALUE, "autotune: size out of [0, 512]");
        ld[i] = std::max<int64_t>(1, h_n[i]);
        off[i + 1] = off[i] + ld[i] * std::max<int64_t>(1, h_n[i]);
        boff[i + 1] = boff[i] + ld[i] * std::max<int64_t>(1, nrhs);
    }
    // device layout: sizes | lds | order | info | A pointers | B pointers | A0 | A | B
    const size_t el = sizeof(T);
    size_t bytes = 0;
    auto take = [&](size_t b) { size_t o = bytes; bytes += (b + 255) / 256 * 256; return o; };
    const size_t o_n = take(batch * 8), o_ld = take(batch * 8), o_ord = take(batch * 8), o_info = take(batch * 8),
                 o_pa = take(batch * 8), o_pb = take(batch * 8), o_a0 = take(off[batch] * el), o_a = take(off[batch] * el),
                 o_b = take(boff[batch] * el);
    char* base = nullptr;
    RMH_TRY_HIP(hipMalloc(&base, bytes));
    struct Free { char* p; ~Free() { (void)hipFree(p); } } guard{base};
    std::vector<int64_t> order(batch);
    for (int64_t i = 0; i < batch; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return h_n[x] > h_n[y]; });
    std::vector<T*> pa(batch), pa0(batch), pb(batch);
    for (int64_t i = 0; i < batch; ++i) {
        pa[i] = reinterpret_cast<T*>(base + o_a) + off[i];
        pa0[i] = reinterpret_cast<T*>(base + o_a0) + off[i];
        pb[i] = reinterpret_cast<T*>(base + o_b) + boff[i];
    }
    RMH_TRY_HIP(hipMemcpy(base + o_n, h_n, batch * 8, hipMemcpyHostToDevice));
    RMH_TRY_HIP(hipMemcpy(base + o_ld, ld.data(), batch * 8, hipMemcpyHostToDevice));
    RMH_TRY_HIP(hipMemcpy(base + o_ord, order.data(), batch * 8, hipMemcpyHostToDevice));
    RMH_TRY_HIP(hipMemcpy(base + o_pa, pa0.data(), batch * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL((synth_spd_kernel<T>), dim3(64, (unsigned)std::min<int64_t>(batch, 65535)), dim3(256), 0, nullptr,
                       reinterpret_cast<T* const*>(base + o_pa), reinterpret_cast<const int64_t*>(base + o_n),
                       reinterpret_cast<const int64_t*>(base + o_ld), batch);
    RMH_TRY(check_launch("synth_spd_kernel"));
    RMH_TRY_HIP(hipMemcpy(base + o_pa, pa.data(), batch * 8, hipMemcpyHostToDevice));
    RMH_TRY_HIP(hipMemcpy(base + o_pb, pb.data(), batch * 8, hipMemcpyHostToDevice));
    hipStream_t s = nullptr;
    hipEvent_t e0, e1;
    RMH_TRY_HIP(hipEventCreate(&e0));
    RMH_TRY_HIP(hipEventCreate(&e1));
    double best_t = 1e30;
    rmh_batch_config best_c{32, 0};
    for (int nb : {16, 32}) {
        const int occ = max_wg_per_cu<T>(nb);
        for (int wpc = 1; wpc <= occ; wpc *= 2) {
            for (int use_order = 0; use_order < 2; ++use_order) {
                rmh_batch_config c{nb, wpc};
                float tot = 0.f;
                const int reps = 3;
                for (int r = 0; r < reps + 1; ++r) {
                    RMH_TRY_HIP(hipMemcpyAsync(base + o_a, base + o_a0, off[batch] * el, hipMemcpyDeviceToDevice, s));
                    RMH_TRY_HIP(hipMemsetAsync(base + o_b, 0, boff[batch] * el, s));
                    RMH_TRY_HIP(hipEventRecord(e0, s));
                    RMH_TRY(batched_posv_impl<T>(h, batch, kMaxN, reinterpret_cast<const int64_t*>(base + o_n),
                                                 reinterpret_cast<T* const*>(base + o_pa),
                                                 reinterpret_cast<const int64_t*>(base + o_ld), nrhs,
                                                 reinterpret_cast<T* const*>(base + o_pb),
                                                 reinterpret_cast<const int64_t*>(base + o_ld),
                                                 reinterpret_cast<int64_t*>(base + o_info),
                                                 use_order ? reinterpret_cast<const int64_t*>(base + o_ord) : nullptr, &c,
                                                 true, s));
                    RMH_TRY_HIP(hipEventRecord(e1, s));
                    RMH_TRY_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    RMH_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) tot += ms;
                }
                const double t = tot / reps;
                if (t < best_t) {
                    best_t = t;
                    best_c = c;
                    best_c.lpt_order = use_order;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->batch_cfg[sizeof(T) == 8 ? 1 : 0] = best_c;
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    return RMH_STATUS_SUCCESS;
}

template rmh_status batched_posv_impl<double>(Context*, int64_t, int, const int64_t*, double* const*, const int64_t*,
                                              int64_t, double* const*, const int64_t*, int64_t*, const int64_t*,
                                              const rmh_batch_config*, bool, hipStream_t);
template rmh_status batched_posv_impl<float>(Context*, int64_t, int, const int64_t*, float* const*, const int64_t*,
                                             int64_t, float* const*, const int64_t*, int64_t*, const int64_t*,
                                             const rmh_batch_config*, bool, hipStream_t);
template rmh_status batched_autotune_impl<double>(Context*, int64_t, const int64_t*, int64_t, rmh_batch_config*,
                                                  double*);
template rmh_status batched_autotune_impl<float>(Context*, int64_t, const int64_t*, int64_t, rmh_batch_config*,
                                                 double*);

}  // namespace rmh
