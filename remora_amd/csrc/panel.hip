// Fused left-looking Cholesky panel step.  For one block column of width
// kb (<= 64) inside a panel it
//   1. applies the updates of the panel's earlier columns (FP64 MFMA),
//   2. factors the kb x kb diagonal block in one wavefront, and
//   3. solves the rows below it against the factor
// -- one launch per block column.  Replaces the reference HIP backend's
// per-16-column potrf_block_kernel + cuBLAS trsm pair
// (include/remora/kernels/hip/potrf.hpp:41-80, 113).
//
// View: V(i, c) = P[i*rs + c*cs], V(0, K) the diagonal element of the block
// column; columns [0, K) are the panel's already factored columns, columns
// [K, K+kb) the block column; rows [0, rows) run from the diagonal down.
//
// Each workgroup owns 64 rows below the diagonal block, 16 per wavefront.
// The launches form a serial chain, so the kernel is built for latency:
//  * Update.  Each wavefront loads its MFMA operands straight from global
//    memory (L2) into registers -- its own 16 rows R and the diagonal rows
//    D = V(0:kb, 0:K), 16 depths per chunk, several chunks in flight -- and
//    accumulates own -= R D^T and its share of the lower blocks of
//    diag -= D D^T on 16x16x4 FP64 MFMA tiles, with no LDS staging and no
//    barriers in the K loop.
//  * Factor and solve.  For kb <= 32 every factoring wavefront holds the
//    updated diagonal block in lanes 0..kb-1 (lane i row i) and 64-kb own
//    rows in the other lanes, and runs the right-looking elimination of the
//    block on all of them (wave_chol_rows): the block lanes end with the
//    factor L, the row lanes with x, x L^T = a.  Every wavefront and
//    workgroup forms the same factor, so none waits for another.  (64-wide
//    blocks run this twice, on their two 32-wide halves, with the rank-32
//    update of the second half between.)  Workgroup 0 writes the factor to
//    a side buffer rather than over the block, which the other workgroups
//    may still be reading; the driver copies all diagonal factors into
//    place once the factorisation is done.
// The arithmetic is FP64 whatever the storage type (the mixed-precision
// factorisation keeps its diagonal blocks in FP64).
#include <type_traits>

#include "rmh_internal.hpp"
#include "wave_chol.hpp"

namespace rmh {
namespace {

// Threads per workgroup: blocks up to 32 wide run eight wavefronts, two per
// SIMD, the wavefronts w and w+4 taking alternate K chunks of the same rows
// (one f64 MFMA chain per SIMD leaves the matrix unit half idle); 64-wide
// blocks need the registers of four.  The narrow form (four wavefronts, one
// K group) fits beside a trailing-update workgroup whose registers fill
// more than half of each SIMD, so it can run while that update runs.
constexpr int panel_threads(int nb, bool narrow = false) { return nb <= 32 && !narrow ? 512 : 256; }
constexpr int kRW = 64;  // rows per workgroup
constexpr int kKc = 16;  // K chunk: four 16x16x4 MFMA steps

typedef double dx4 __attribute__((ext_vector_type(4)));

// d-th block (bi >= bj) of the lower block triangle, row by row
__device__ __forceinline__ void lower_block(int d, int& bi, int& bj) {
    bi = 0;
    while (d > bi) { d -= bi + 1; ++bi; }
    bj = d;
}

template <typename T, int NB, int THREADS>
__device__ __forceinline__ void panel_body(T* P, int64_t rs, int64_t cs, int64_t rows, int K, int kb, int64_t* info,
                                           int64_t info_base, T* dfac, int64_t ldf, __bf16* sph, __bf16* spl,
                                           int64_t sp_kstride) {
    constexpr int kThreads = THREADS;
    constexpr int KG = kThreads / 256;        // wavefronts sharing a row block (K groups)
    constexpr int NCB = NB / 16;              // 16-column blocks
    constexpr int NDB = NCB * (NCB + 1) / 2;  // lower blocks of the diagonal block
    constexpr int DPW = (NDB + 3) / 4;        // diagonal blocks per wave
    constexpr int LD = NB + 1;
    constexpr int CD = (NB * NB + kThreads - 1) / kThreads, CR = kRW * NB / kThreads;  // block-column elements per thread
    constexpr int DEP = 2;                    // K chunks in flight per wavefront (deeper only grows the code)
    static_assert(NB % 16 == 0 && CR * kThreads == kRW * NB, "unsupported block width");

    __shared__ double L[NB * LD];  // updated diagonal block, then its factor
    __shared__ double X[kRW * LD]; // updated own rows
    __shared__ __align__(16) double col[4 * NB];  // pivot-column buffers (two per factoring wavefront)
    __shared__ int failed;

    // the panel chain is the factorisation's critical path: where a panel
    // wavefront shares a SIMD with a trailing-update wavefront, it issues first
    if constexpr (THREADS == 256 && NB <= 32) __builtin_amdgcn_s_setprio(3);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave & 3, kg = wave >> 2;  // row block, K group (chunks kg, kg + KG, ...)
    const int64_t r0 = kb + (int64_t)blockIdx.x * kRW;  // first own row
    const int nown = (int)max((int64_t)0, min((int64_t)kRW, rows - r0));
    const int nch = (K + kKc - 1) / kKc;

    // K-loop operands in the MFMA lane layout: lane l holds row l % 16 at
    // depth l / 16 of a 4-deep step.  Loads keep the raw values (converting
    // or masking them at load time would wait for them there).  Rows outside
    // the matrix read row 0 and only produce values that are never stored
    // (own rows past nown, columns of X past kb, entries of L outside the
    // kb x kb block, which is not updated); depths past K read depth K-1 and
    // are zeroed in A when the chunk is computed.
    const int li = lane & 15, lk = lane >> 4;
    const bool a_ok = 16 * wr + li < nown;
    const T* pa = P + (a_ok ? r0 + 16 * wr + li : 0) * rs;
    // This wavefront's diagonal blocks of D D^T (its share of the lower
    // block triangle) and the order in which it holds D's column blocks:
    // for two column blocks, waves 1 and 2 (blocks (1,0) and (1,1)) hold
    // block 1 first, so every diagonal product reads fixed registers -- x
    // from slot 0, y from slot 0 or 1 (diag mode dm 1 or 2; 0: none; 3:
    // wider blocks select per product).
    int dbi[DPW], dbj[DPW];
#pragma unroll
    for (int q = 0; q < DPW; ++q) {
        dbi[q] = dbj[q] = 0;
        if (wr + 4 * q < NDB) lower_block(wr + 4 * q, dbi[q], dbj[q]);
    }
    int pblk[NCB];
#pragma unroll
    for (int b = 0; b < NCB; ++b) pblk[b] = b;
    int dm = 3;
    if (NCB <= 2) {
        if (NCB == 2 && (wr == 1 || wr == 2)) { pblk[0] = 1; pblk[NCB - 1] = 0; }
        dm = wr >= NDB ? 0 : (wr == 1 ? 2 : 1);
    }
    const T* pb[NCB];
#pragma unroll
    for (int b = 0; b < NCB; ++b) pb[b] = P + (int64_t)(16 * pblk[b] + li < kb ? 16 * pblk[b] + li : 0) * rs;
    // Chunk loads are unconditional (chunks past the end read depth kmax and
    // are never computed), so the compiler counts the loads in flight
    // exactly and waits only for the ones it needs.
    const int kmax = max(K - 1, 0);
    T av[DEP][4], bv[DEP][4][NCB];
    auto load_chunk = [&](int ch, T(&a)[4], T(&bm)[4][NCB]) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int64_t off = (int64_t)min(ch * kKc + 4 * s4 + lk, kmax) * cs;
            a[s4] = pa[off];
#pragma unroll
            for (int b = 0; b < NCB; ++b) bm[s4][b] = pb[b][off];
        }
    };
#pragma unroll
    for (int s = 0; s < DEP - 1; ++s) load_chunk(kg + KG * s, av[s], bv[s]);

    // block column: own rows and the diagonal block, loaded after the first
    // K chunks from always-valid (clamped) addresses; the raw values go to
    // LDS after the K loop, masked, the diagonal block padded with the
    // identity
    T cur_r[CR], cur_d[CD];
#pragma unroll
    for (int q = 0; q < CR; ++q) {
        const int e = tid + q * kThreads, i = e % kRW, c = e / kRW;
        cur_r[q] = P[(i < nown ? r0 + i : 0) * rs + (int64_t)(K + (c < kb ? c : 0)) * cs];
    }
#pragma unroll
    for (int q = 0; q < CD; ++q) {
        const int e = min(tid + q * kThreads, NB * NB - 1), i = e % NB, c = e / NB;
        const bool ok = i < kb && c <= i;
        cur_d[q] = P[(ok ? i : 0) * rs + (int64_t)(K + (ok ? c : 0)) * cs];
    }

    if (tid == 0) failed = -1;

    // 1. left-looking update: wave w owns rows [16w, 16w+16) of X (all
    // column blocks) and diagonal blocks w, w+4, ...
    dx4 acc[NCB], dacc[DPW];
#pragma unroll
    for (int b = 0; b < NCB; ++b) acc[b] = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < DPW; ++b) dacc[b] = dx4{0.0, 0.0, 0.0, 0.0};
    // One chunk: MASK zeroes the depths past K (last chunk only), DM is the
    // diagonal mode above.  Operands go straight from the load registers to
    // the MFMAs.
    auto compute = [&](auto mask_c, auto dm_c, int ch, const T(&a)[4], const T(&bm)[4][NCB]) {
        constexpr bool MASK = decltype(mask_c)::value;
        constexpr int DM = decltype(dm_c)::value;
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const bool k_ok = !MASK || ch * kKc + 4 * s4 + lk < K;
            const double av4 = k_ok ? (double)a[s4] : 0.0;
            double bv4[NCB];
#pragma unroll
            for (int b = 0; b < NCB; ++b) bv4[b] = (double)bm[s4][b];
#pragma unroll
            for (int b = 0; b < NCB; ++b) acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av4, bv4[b], acc[b], 0, 0, 0);
            if constexpr (DM == 1 || DM == 2) {
                const double x = k_ok ? bv4[0] : 0.0;  // the zeroed depths must not enter this product either
                dacc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, bv4[DM == 2 ? NCB - 1 : 0], dacc[0], 0, 0, 0);
            } else if constexpr (DM == 3) {
#pragma unroll
                for (int q = 0; q < DPW; ++q) {
                    if (wr + 4 * q < NDB) {
                        double x = bv4[0], y = bv4[0];
#pragma unroll
                        for (int b = 1; b < NCB; ++b) {
                            x = (dbi[q] == b) ? bv4[b] : x;
                            y = (dbj[q] == b) ? bv4[b] : y;
                        }
                        x = k_ok ? x : 0.0;
                        dacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, dacc[q], 0, 0, 0);
                    }
                }
            }
        }
    };
    // the t-th chunk of this K group (ch = kg + KG t) sits in register slot
    // t % DEP; the slot refilled at each step is the one computed at the
    // step before
    auto kloop = [&](auto dm_c) {
        for (int t0 = 0; kg + KG * t0 < nch; t0 += DEP) {
#pragma unroll
            for (int s = 0; s < DEP; ++s) {
                const int ch = kg + KG * (t0 + s);
                if (ch >= nch) break;
                load_chunk(ch + KG * (DEP - 1), av[(s + DEP - 1) % DEP], bv[(s + DEP - 1) % DEP]);
                if (ch * kKc + kKc <= K) compute(std::false_type{}, dm_c, ch, av[s], bv[s]);
                else compute(std::true_type{}, dm_c, ch, av[s], bv[s]);
            }
        }
    };
    if (NCB > 2) kloop(std::integral_constant<int, 3>{});
    else if (dm == 1) kloop(std::integral_constant<int, 1>{});
    else if (dm == 2) kloop(std::integral_constant<int, 2>{});
    else kloop(std::integral_constant<int, 0>{});
#pragma unroll
    for (int q = 0; q < CR; ++q) {
        const int e = tid + q * kThreads, i = e % kRW, c = e / kRW;
        X[i * LD + c] = (i < nown && c < kb) ? (double)cur_r[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CD; ++q) {
        const int e = tid + q * kThreads, i = e % NB, c = e / NB;
        if (e < NB * NB) L[i * LD + c] = (i < kb && c <= i) ? (double)cur_d[q] : (i == c ? 1.0 : 0.0);
    }
    __syncthreads();  // the block column is in X and L
    // subtract, one K group after the other: lane l, item r of a 16x16 tile
    // is element (4r + l/16, l%16)
#pragma unroll
    for (int g = 0; g < KG; ++g) {
        if (K > 0 && kg == g) {
#pragma unroll
            for (int b = 0; b < NCB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(16 * wr + 4 * r + lk) * LD + 16 * pblk[b] + li] -= acc[b][r];
#pragma unroll
            for (int q = 0; q < DPW; ++q) {
                if (wr + 4 * q < NDB) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = 16 * dbi[q] + 4 * r + lk, c = 16 * dbj[q] + li;
                        if (c <= i && i < kb) L[i * LD + c] -= dacc[q][r];  // the identity padding stays
                    }
                }
            }
        }
        __syncthreads();
    }

    if constexpr (NB <= 32) {
        // 2+3. wavefronts 0.. hold the diagonal block in lanes 0..NB-1 and
        // 64-NB own rows in the other lanes, and factor the block while
        // solving the rows along (wave_chol_rows).  Every such wavefront
        // forms the same factor, so none waits for another.
        constexpr int OPW = 64 - NB;  // own rows per factoring wavefront
        constexpr int NFW = (kRW + OPW - 1) / OPW;
        if (wave >= NFW) return;
        const int orow = wave * OPW + lane - NB;  // own row of a lane >= NB
        const bool is_diag = lane < NB, is_row = !is_diag && orow < kRW;
        const double* src = is_diag ? &L[lane * LD] : &X[(is_row ? orow : 0) * LD];
        double r[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) r[c] = src[c];
        const int bad = wave_chol_rows<NB>(r, &col[wave * 2 * NB], lane);
        if (bad >= 0) {
            // keep the first failure: later launches run on garbage but must not overwrite it
            if (blockIdx.x == 0 && tid == 0)
                atomicCAS(reinterpret_cast<unsigned long long*>(info), 0ull, (unsigned long long)(info_base + bad + 1));
            return;
        }
        if (is_row && orow < nown) {
            const int64_t i = r0 + orow;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < kb) P[i * rs + (int64_t)(K + c) * cs] = (T)r[c];
            if constexpr (sizeof(T) == 4) {
                // the stored FP32 values split into BF16 hi + lo for the
                // split-BF16x3 trailing update: panel column k of row i at
                // (k / 32) * sp_kstride + i * 32 + k % 32, written 8 at a
                // time (K is a multiple of NB, so the NB columns are one
                // contiguous run; padded columns hold zeros)
                if (sph != nullptr) {
                    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
                    const int64_t e0 = (int64_t)(K >> 5) * sp_kstride + i * 32 + (K & 31);
#pragma unroll
                    for (int c8 = 0; c8 < NB; c8 += 8) {
                        bf16x8 hv, lv;
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            float v = (float)r[c8 + q];
                            // split the stored FP32 value: without the empty
                            // asm the compiler merges the two conversions
                            // into one rounding of the FP64 value to BF16,
                            // which differs on ties
                            __asm__("" : "+v"(v));
                            const __bf16 h = (__bf16)v;
                            hv[q] = h;
                            lv[q] = (__bf16)(v - (float)h);
                        }
                        *reinterpret_cast<bf16x8*>(sph + e0 + c8) = hv;
                        *reinterpret_cast<bf16x8*>(spl + e0 + c8) = lv;
                    }
                }
            }
        }
        // the factor of the diagonal block goes to the side buffer: every
        // workgroup of this launch still reads the block's original values
        if (blockIdx.x == 0 && wave == 0 && lane < kb) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= lane) dfac[lane + c * ldf] = (T)r[c];
        }
    } else {
        // 64-wide block column as two 32-wide halves, each factored with the
        // rows along (wave_chol_rows): half 1 solves the lower half of the
        // diagonal block and the own rows, then both take the rank-32
        // update of half 1, then half 2 solves the own rows.  The whole
        // 64 x 64 lower block goes to the side buffer.
        constexpr int H = NB / 2;
        double* cbuf = &col[wave * 2 * H];
        // half 1: wave 0 holds rows 0..63 of the block (pivot rows 0..31 and
        // the rows 32..63 below them), waves 1 and 2 the pivot rows and 32
        // own rows each
        {
            const int orow = (wave - 1) * H + lane - H;
            const bool own = wave >= 1 && wave < 3 && lane >= H;
            const double* src = own ? &X[orow * LD] : &L[lane * LD];
            double r[H];
#pragma unroll
            for (int c = 0; c < H; ++c) r[c] = src[c];
            __syncthreads();  // every wavefront holds its rows before wave 0 writes the block back
            if (wave < 3) {
                const int bad = wave_chol_rows<H>(r, cbuf, lane);
                if (wave == 0 && lane == 0) failed = bad;
                // wave 0 writes the first half of the block (its pivot rows
                // and the rows below them), waves 1 and 2 their own rows
                double* dst = (wave == 0) ? &L[lane * LD] : (own ? &X[orow * LD] : nullptr);
                if (dst != nullptr) {
#pragma unroll
                    for (int c = 0; c < H; ++c) dst[c] = r[c];
                }
            }
        }
        __syncthreads();
        if (failed >= 0) {
            if (blockIdx.x == 0 && tid == 0)
                atomicCAS(reinterpret_cast<unsigned long long*>(info), 0ull, (unsigned long long)(info_base + failed + 1));
            return;
        }
        // rank-32 update of the second half: rows 32..63 of the block (lower
        // part) and the own rows, a[:, 32 + c] -= x[:, 0:32] . L[32 + c, 0:32]
        for (int e = tid; e < (kRW + H) * H; e += kThreads) {
            const int i = e % (kRW + H), c = e / (kRW + H);
            double* row = i < H ? &L[(H + i) * LD] : &X[(i - H) * LD];
            if (i < H && c > i) continue;  // upper part of the second diagonal block
            const double* lr = &L[(H + c) * LD];
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int t = 0; t < H; t += 2) {
                s0 = fma(row[t], lr[t], s0);
                s1 = fma(row[t + 1], lr[t + 1], s1);
            }
            row[H + c] -= s0 + s1;
        }
        __syncthreads();
        // half 2: waves 0 and 1 hold the second diagonal block (lanes 0..31)
        // and 32 own rows each
        if (wave < 2) {
            const int orow = wave * H + lane - H;
            const double* src = lane < H ? &L[(H + lane) * LD + H] : &X[orow * LD + H];
            double r[H];
#pragma unroll
            for (int c = 0; c < H; ++c) r[c] = src[c];
            const int bad = wave_chol_rows<H>(r, cbuf, lane);
            if (bad >= 0) {
                if (blockIdx.x == 0 && tid == 0)
                    atomicCAS(reinterpret_cast<unsigned long long*>(info), 0ull,
                              (unsigned long long)(info_base + H + bad + 1));
                return;
            }
            if (lane >= H && orow < nown) {
                const int64_t i = r0 + orow;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const double v = c < H ? X[orow * LD + c] : r[c - H];
                    if (c < kb) P[i * rs + (int64_t)(K + c) * cs] = (T)v;
                }
            }
            // the factor of the diagonal block goes to the side buffer: every
            // workgroup of this launch still reads the block's original values
            if (blockIdx.x == 0 && wave == 0) {
                if (lane < H) {
                    const int i = H + lane;
                    if (i < kb) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) {
                            const double v = c < H ? L[i * LD + c] : r[c - H];
                            if (c <= i) dfac[i + c * ldf] = (T)v;
                        }
                    }
                }
            }
            if (blockIdx.x == 0 && wave == 1 && lane < H && lane < kb) {
#pragma unroll
                for (int c = 0; c < H; ++c)
                    if (c <= lane) dfac[lane + c * ldf] = (T)L[lane * LD + c];
            }
        }
    }
}

template <typename T, int NB, int THREADS>
__global__ __launch_bounds__(THREADS) void potrf_panel_kernel(T* P, int64_t rs, int64_t cs, int64_t rows, int K,
                                                               int kb, int64_t* info, int64_t info_base, T* dfac,
                                                               int64_t ldf, __bf16* sph, __bf16* spl,
                                                               int64_t sp_kstride) {
    panel_body<T, NB, THREADS>(P, rs, cs, rows, K, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
}

// The narrow form held to 168 registers (three wavefronts' worth per SIMD),
// so that one workgroup fits in what one trailing-update workgroup leaves
// when it retires.
template <typename T, int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void potrf_panel_kernel_narrow(
    T* P, int64_t rs, int64_t cs, int64_t rows, int K, int kb, int64_t* info, int64_t info_base, T* dfac, int64_t ldf,
    __bf16* sph, __bf16* spl, int64_t sp_kstride) {
    panel_body<T, NB, 256>(P, rs, cs, rows, K, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
}

}  // namespace

// Copy the diagonal-block factors from the side buffer into the matrix:
// element (i, c) of the block starting at j0 = (r / nbi) * nbi, r = j0 + i,
// lives at dfac[r + c * ldf].  Blocks at or after a failed pivot (info != 0)
// keep their original values.
template <typename T>
__global__ __launch_bounds__(256) void scatter_diag_kernel(T* A, int64_t rs, int64_t cs, int64_t n, int nbi,
                                                           const T* __restrict__ dfac, int64_t ldf,
                                                           const int64_t* __restrict__ info) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int64_t j0 = (r / nbi) * nbi;
    const int64_t failed = *info;
    if (failed != 0 && j0 + nbi >= failed) return;
    const int i = (int)(r - j0);
    for (int c = 0; c <= i; ++c) A[r * rs + (j0 + c) * cs] = dfac[r + c * ldf];
}

template <typename T>
rmh_status launch_potrf_panel(T* P, int64_t rs, int64_t cs, int64_t rows, int64_t K, int kb, int64_t* info,
                              int64_t info_base, T* dfac, int64_t ldf, hipStream_t stream, __bf16* sph, __bf16* spl,
                              int64_t sp_kstride, bool narrow) {
    if (kb <= 0) return RMH_STATUS_SUCCESS;
    if (rows < kb) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: fewer rows than columns");
    if (K < 0 || K > (int64_t)1 << 30) return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: bad panel offset");
    const int64_t below = rows - kb;
    const unsigned grid = (unsigned)std::max<int64_t>(1, (below + kRW - 1) / kRW);
    const int k = (int)K;
    if (sph != nullptr && kb > 32)
        return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: split output needs blocks of at most 32 columns");
    if (kb <= 16 && narrow)
        hipLaunchKernelGGL((potrf_panel_kernel_narrow<T, 16>), dim3(grid), dim3(panel_threads(16, true)), 0, stream, P,
                           rs, cs, rows, k, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
    else if (kb <= 16)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 16, panel_threads(16)>), dim3(grid), dim3(panel_threads(16)), 0,
                           stream, P, rs, cs, rows, k, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
    else if (kb <= 32 && narrow)
        hipLaunchKernelGGL((potrf_panel_kernel_narrow<T, 32>), dim3(grid), dim3(panel_threads(32, true)), 0, stream, P,
                           rs, cs, rows, k, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
    else if (kb <= 32)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 32, panel_threads(32)>), dim3(grid), dim3(panel_threads(32)), 0,
                           stream, P, rs, cs, rows, k, kb, info, info_base, dfac, ldf, sph, spl, sp_kstride);
    else if (kb <= 64)
        hipLaunchKernelGGL((potrf_panel_kernel<T, 64, panel_threads(64)>), dim3(grid), dim3(panel_threads(64)), 0,
                           stream, P, rs, cs, rows, k, kb, info, info_base, dfac, ldf, nullptr, nullptr, (int64_t)0);
    else
        return fail(RMH_STATUS_INVALID_VALUE, "potrf_panel: block wider than 64");
    return check_launch("potrf_panel_kernel");
}

template <typename T>
rmh_status launch_scatter_diag(T* A, int64_t rs, int64_t cs, int64_t n, int nbi, const T* dfac, int64_t ldf,
                               const int64_t* info, hipStream_t stream) {
    if (n <= 0) return RMH_STATUS_SUCCESS;
    hipLaunchKernelGGL((scatter_diag_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, A, rs, cs,
                       n, nbi, dfac, ldf, info);
    return check_launch("scatter_diag_kernel");
}

template rmh_status launch_potrf_panel<double>(double*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t,
                                               double*, int64_t, hipStream_t, __bf16*, __bf16*, int64_t, bool);
template rmh_status launch_potrf_panel<float>(float*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t,
                                              float*, int64_t, hipStream_t, __bf16*, __bf16*, int64_t, bool);
template rmh_status launch_scatter_diag<double>(double*, int64_t, int64_t, int64_t, int, const double*, int64_t,
                                                const int64_t*, hipStream_t);
template rmh_status launch_scatter_diag<float>(float*, int64_t, int64_t, int64_t, int, const float*, int64_t,
                                               const int64_t*, hipStream_t);

}  // namespace rmh
