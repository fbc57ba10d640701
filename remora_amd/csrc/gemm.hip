// MFMA-tiled GEMM / SYRK for gfx950 (CDNA4).
//
// One kernel template serves C = alpha*op(A)*op(B) + beta*C on a full
// rectangle and the triangular / trapezoidal SYRK updates of the Cholesky
// factorisation (the trailing update A22 -= P P^T and the in-panel update).
// It replaces the cuBLAS gemm/syrk calls of the reference's HIP backend
// (include/remora/hip/cublas_backend.hpp:121-166, 389-431) that
// include/remora/kernels/hip/{gemm,syrk}.hpp forward to.
//
// Design (64-wide wavefronts, 4 waves per workgroup):
//  * the workgroup tile BM x BN is split 2 x 2 over the waves; each wave
//    owns (BM/2) x (BN/2) of C as a grid of 16x16 MFMA accumulators
//    (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32);
//  * op(A) and op(B) tiles of depth BK are staged in LDS k-major
//    ([k][m] and [k][n], rows padded by 16 elements so the 64 lanes of an
//    MFMA operand fetch -- 16 consecutive m at 4 consecutive k -- hit
//    distinct banks), double-buffered with a register prefetch of the next
//    tile so global latency overlaps the MFMAs of the current one;
//  * the MFMA is issued as D = op(B)^T op(A)^T so that a lane's four
//    accumulator entries are four columns of C and the 16 lanes of a row
//    group write 16 consecutive rows: each store instruction writes four
//    contiguous 16-element column segments of the column-major C (the f64
//    and f32 instructions interleave D's rows differently; see Mfma);
//  * workgroup ids are remapped so each XCD (8 per MI355X, with private L2)
//    receives a contiguous range of tiles, and tiles are visited in groups of
//    8 tile-rows so neighbouring workgroups share op(A) and op(B) panels;
//  * Part::Lower / Part::Upper launch only the tiles that touch the wanted
//    trapezoid and mask the diagonal-crossing ones at the store.
//
// In-place use: when C aliases op(A) (or op(B)) and the launch has a single
// tile column (BN >= N) (or a single tile row), every workgroup reads all of
// its own rows (columns) before it writes them and no other workgroup reads
// them.  The Cholesky panel solve uses that (A21 <- A21 * W^T).
#include <atomic>

#include "rmh_internal.hpp"

namespace rmh {
namespace {

template <typename T>
struct Mfma;
// Result layout of a 16x16 accumulator (rows i, columns j of D):
//   f32 16x16x4: lane l, item r holds D[4*(l/16) + r][l%16]
//   f64 16x16x4: lane l, item r holds D[4*r + l/16][l%16]
// (operands: A[i=l%16][k=l/16], B[k=l/16][j=l%16] for both).
template <>
struct Mfma<double> {
    static constexpr bool kRowInterleaved = true;
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<float> {
    static constexpr bool kRowInterleaved = false;
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename T>
struct KParams {
    int64_t M, N, K;
    T alpha, beta;
    const T* A;
    int64_t lda;
    const T* B;
    int64_t ldb;
    T* C;
    int64_t ldc;
    int64_t tiles_m, tiles_n, ntiles;
    const int64_t* abort_flag;
    // strided batch: blockIdx.z = b -> (y = b % batch_ny, z = b / batch_ny),
    // operand origins advance by y*s?y + z*s?z elements
    int64_t batch_ny;
    int64_t sAy, sAz, sBy, sBz, sCy, sCz;
    // split-K: blockIdx.y owns K range [y*k_chunk, (y+1)*k_chunk) and writes
    // its partial product to split_ws + y*M*N (ld M); a reduction combines them
    int64_t k_chunk;
    T* split_ws;
    // triangular operands: a tile whose rows start at m0 only accumulates
    // k >= m0 (op(A)(i, k) = 0 for k < i)
    int k_from_row;
    // every element offset inside a staged tile fits 32 bits (in bytes):
    // full stages load from a wave-uniform base plus fixed per-lane offsets
    int fast32;
    // quarter tiles (PART 1 / 2, sub_tbig > 0): workgroup w computes quarter
    // w % 4 of tile sub_t0 + w / 4 of the grouped trapezoid order on tiles of
    // twice this kernel's size, sub_tbig of them along the short side
    int64_t sub_t0, sub_tbig;
};

// word 3 of a raw (untyped) buffer descriptor on gfx9-family parts
constexpr int kBufWord3 = 0x00020000;
// a wave-uniform pointer, moved to scalar registers (a buffer resource must
// be scalar; the compiler cannot always prove the uniformity itself)
template <typename T>
__device__ __forceinline__ const T* uniform_ptr(const T* p) {
    const uint64_t u = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return reinterpret_cast<const T*>(((uint64_t)hi << 32) | lo);
}
// voff: per-lane byte offset (VGPR); soff: wave-uniform byte offset (SGPR)
template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

constexpr int kThreads = 256;
constexpr int kPad = 2;  // LDS row padding (elements): 16 gave heavy bank conflicts on the operand reads
constexpr int kXcd = 8;
constexpr int kGroupM = 8;
constexpr int64_t kMaxLd = (int64_t)1 << 23;

// Lower trapezoid of tiles: tile-row i holds min(i+1, tn) tiles.
__device__ __forceinline__ void lower_trapezoid_tile(int64_t t, int64_t tn, int64_t& ti, int64_t& tj) {
    const int64_t tri = tn * (tn + 1) / 2;
    if (t < tri) {
        int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > t) --i;
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        ti = i;
        tj = t - i * (i + 1) / 2;
    } else {
        const int64_t r = t - tri;
        ti = tn + r / tn;
        tj = r % tn;
    }
}

// The same tiles in groups of kGroupT tile rows, each group column by
// column: tile rows [i0, i0 + g) take the tiles (i0 .. i0+g-1, j) for
// j = 0 .. i0-1, then their g x g diagonal block column by column.  An XCD
// works through a contiguous range of this order, so its 64 concurrent
// tiles form a block of about 8 x 8 tiles and share 8 row and 8 column
// blocks of the operands (row order shares one row block and reads 64
// column blocks).  The rectangular part below a triangle stays row by row
// (rows of tn tiles).
constexpr int kGroupT = 8;
__device__ __forceinline__ void lower_trapezoid_tile_grouped(int64_t t, int64_t tn, int64_t& ti, int64_t& tj) {
    const int64_t tri = tn * (tn + 1) / 2;
    if (t >= tri) {
        const int64_t r = t - tri;
        ti = tn + r / tn;
        tj = r % tn;
        return;
    }
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);  // row of t in row order
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    const int64_t i0 = (i / kGroupT) * kGroupT;  // first row of its group
    const int64_t g = min((int64_t)kGroupT, tn - i0);
    int64_t r = t - i0 * (i0 + 1) / 2;
    if (r < i0 * g) {
        tj = r / g;
        ti = i0 + r % g;
        return;
    }
    r -= i0 * g;
    int64_t jj = 0;
    while (r >= g - jj) {
        r -= g - jj;
        ++jj;
    }
    tj = i0 + jj;
    ti = i0 + jj + r;
}

// SYM != 0: op(A) is a symmetric matrix of which only the lower (SYM == 1)
// or upper (SYM == 2) triangle is stored (requires !TA).  Tiles wholly in
// the stored triangle load directly, tiles wholly in the other one load the
// mirrored element with the coalesced transposed mapping, and the few tiles
// that cross the diagonal select per element.
template <typename T, int BM, int BN, int BK, bool TA, bool TB, int PART, int SYM = 0>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void gemm_kernel(KParams<T> p) {
    typedef Mfma<T> MF;
    typedef typename MF::acc_t acc_t;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA_S = BM + kPad, LDB_S = BN + kPad;
    constexpr int A_PER = BM * BK / kThreads;
    constexpr int B_PER = BN * BK / kThreads;
    static_assert(A_PER * kThreads == BM * BK && B_PER * kThreads == BN * BK, "tile/thread mismatch");

    __shared__ T As[2][BK * LDA_S];
    __shared__ T Bs[2][BK * LDB_S];

    if (p.abort_flag != nullptr && *p.abort_flag != 0) return;
    if (gridDim.z > 1) {
        const int64_t b = blockIdx.z, by = b % p.batch_ny, bz = b / p.batch_ny;
        p.A += by * p.sAy + bz * p.sAz;
        p.B += by * p.sBy + bz * p.sBz;
        p.C += by * p.sCy + bz * p.sCz;
    }
    int64_t kb0 = 0;  // SYM: absolute depth of this split (A is addressed absolutely)
    if (gridDim.y > 1) {
        const int64_t k0s = (int64_t)blockIdx.y * p.k_chunk;
        if (SYM == 0) p.A += TA ? k0s : k0s * p.lda;
        else kb0 = k0s;
        p.B += TB ? k0s * p.ldb : k0s;
        p.K = min(p.K, k0s + p.k_chunk) - k0s;
        p.C = p.split_ws + (int64_t)blockIdx.y * p.M * p.N;
        p.ldc = p.M;
        p.alpha = T(1);
        p.beta = T(0);
    }

    // ---- tile selection: XCD-contiguous ranges, then grouped order -------
    // Workgroups go to the XCDs in turn; each XCD takes a contiguous range
    // of tiles so that neighbouring tiles share its L2.  With k_from_row the
    // depth shrinks row by row, so contiguous ranges would give the first
    // XCD several times the work of the last: those tiles stay in row order,
    // dealt out to the XCDs in turn.
    const int64_t G = gridDim.x;
    int64_t pid = blockIdx.x;
    if (!p.k_from_row) {
        const int64_t full = (G / kXcd) * kXcd;
        if (pid < full) pid = (pid % kXcd) * (G / kXcd) + pid / kXcd;
    }
    int64_t tm, tn;
    if (PART == 0) {
        const int64_t width = (int64_t)kGroupM * p.tiles_n;
        const int64_t group = pid / width;
        const int64_t first = group * kGroupM;
        const int64_t gsz = min((int64_t)kGroupM, p.tiles_m - first);
        const int64_t r = pid % width;
        tm = first + r % gsz;
        tn = r / gsz;
    } else if (p.sub_tbig > 0) {
        int64_t bi, bj;  // the big tile in (long side, short side) order
        lower_trapezoid_tile_grouped(p.sub_t0 + pid / 4, p.sub_tbig, bi, bj);
        const int64_t qi = 2 * bi + ((pid >> 1) & 1), qj = 2 * bj + (pid & 1);
        tm = PART == 1 ? qi : qj;
        tn = PART == 1 ? qj : qi;
    } else if (PART == 1) {
        if (p.k_from_row) lower_trapezoid_tile(pid, p.tiles_n, tm, tn);
        else lower_trapezoid_tile_grouped(pid, p.tiles_n, tm, tn);
    } else {
        if (p.k_from_row) lower_trapezoid_tile(pid, p.tiles_m, tn, tm);
        else lower_trapezoid_tile_grouped(pid, p.tiles_m, tn, tm);
    }
    // the tile origin is uniform: keep it in scalar registers so that the
    // operand bases of the K loop are scalar too
    const int64_t m0 = __builtin_amdgcn_readfirstlane((int)(tm * BM));
    const int64_t n0 = __builtin_amdgcn_readfirstlane((int)(tn * BN));
    if (PART != 0 && p.sub_tbig > 0) {
        // quarters past the matrix edge or wholly outside the trapezoid
        // (the whole workgroup leaves together)
        if (m0 >= p.M || n0 >= p.N) return;
        if (PART == 1 ? m0 + BM <= n0 : n0 + BN <= m0) return;
    }
    if (p.k_from_row) {
        const int64_t kb = min(p.K, m0);  // m0 is a multiple of BK
        p.A += TA ? kb : kb * p.lda;
        p.B += TB ? kb * p.ldb : kb;
        p.K -= kb;
    }

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t M = p.M, N = p.N, K = p.K;

    // ---- global -> register staging -------------------------------------
    // Per-thread coordinates inside a tile are fixed; the tile origin moves
    // with k0 and is wave-uniform, so loads are a scalar base plus a 32-bit
    // lane offset.  Out-of-range rows/columns/depths are clamped to a valid
    // address and zeroed after the load.
    constexpr int AK_STEP = TA ? 0 : kThreads / BM;   // !TA: k advances by this per r
    constexpr int AM_STEP = TA ? kThreads / BK : 0;   //  TA: m advances by this per r
    constexpr int BK_STEP = TB ? kThreads / BN : 0;   //  TB: k advances per r
    constexpr int BN_STEP = TB ? 0 : kThreads / BK;   // !TB: n advances per r
    const int a_m0 = TA ? tid / BK : tid % BM;
    const int a_k0 = TA ? tid % BK : tid / BM;
    const int b_n0 = TB ? tid % BN : tid / BK;
    const int b_k0 = TB ? tid / BN : tid % BK;
    const int lda32 = (int)p.lda, ldb32 = (int)p.ldb;

    static_assert(SYM == 0 || !TA, "symmetric A is read without transpose");
    // mirrored mapping (used by SYM tiles that lie in the unstored triangle)
    constexpr int MK_STEP = 0, MM_STEP = kThreads / BK;
    const int m_m0 = tid / BK, m_k0 = tid % BK;

    T ra[A_PER], rb[B_PER];
    int a_mode = 0;     // how ra[] of the staged tile was loaded: 0 direct, 1 mirrored
    int st_kmax = 0;    // last valid k inside the staged tile
    const int mmax = (int)min((int64_t)BM, M - m0) - 1;  // last valid row / column of this tile
    const int nmax = (int)min((int64_t)BN, N - n0) - 1;
    // The MFMA and VALU pipes of a SIMD do not overlap, so a stage's VALU
    // work adds to its MFMA time.  Full stages (all BK depths inside K, and
    // 32-bit offsets) therefore load from a wave-uniform base plus per-lane
    // byte offsets computed once per tile (rows and columns clamped), and
    // interior tiles store them to LDS unmasked.  FP64 keeps one per-lane
    // offset per operand and adds r times a wave-uniform step in the scalar
    // offset -- exact where the clamped row or column of element r is linear
    // in r (always along the depth, and across the tile for interior tiles):
    // 14 fewer registers and DGEMM 66.6 -> 68.2 TFLOP/s.  FP32, whose
    // trailing update and solves measured slower that way, keeps an offset
    // per element.
    constexpr bool kStep = sizeof(T) == 8;
    const int cm0 = min(a_m0, mmax), cn0 = min(b_n0, nmax);
    const uint32_t a_vo = (uint32_t)(TA ? a_k0 + cm0 * lda32 : cm0 + a_k0 * lda32) * (uint32_t)sizeof(T);
    const uint32_t b_vo = (uint32_t)(TB ? cn0 + b_k0 * ldb32 : b_k0 + cn0 * ldb32) * (uint32_t)sizeof(T);
    const uint32_t a_step = (uint32_t)((TA ? AM_STEP : AK_STEP) * lda32) * (uint32_t)sizeof(T);
    const uint32_t b_step = (uint32_t)((TB ? BK_STEP : BN_STEP) * ldb32) * (uint32_t)sizeof(T);
    uint32_t aoff[kStep ? 1 : A_PER], boff[kStep ? 1 : B_PER];
    if constexpr (!kStep && SYM == 0) {
#pragma unroll
        for (int r = 0; r < A_PER; ++r) {
            const int cm = min(a_m0 + r * AM_STEP, mmax), ck = a_k0 + r * AK_STEP;
            aoff[r] = (uint32_t)(TA ? ck + cm * lda32 : cm + ck * lda32) * (uint32_t)sizeof(T);
        }
#pragma unroll
        for (int r = 0; r < B_PER; ++r) {
            const int cn = min(b_n0 + r * BN_STEP, nmax), ck = b_k0 + r * BK_STEP;
            boff[r] = (uint32_t)(TB ? cn + ck * ldb32 : ck + cn * ldb32) * (uint32_t)sizeof(T);
        }
    }
    const bool interior = mmax == BM - 1 && nmax == BN - 1;
    const bool a_lin = !kStep || !TA || mmax == BM - 1, b_lin = !kStep || TB || nmax == BN - 1;
    // Loads go to clamped (always valid) addresses and are not inspected
    // here: zeroing the out-of-range elements happens at the LDS store, after
    // the MFMAs of the current stage, so the loads stay in flight meanwhile.
    auto load_tiles = [&](int64_t k0) {
        const int kmax = (int)min((int64_t)BK, K - k0) - 1;  // >= 0
        st_kmax = kmax;
        if (SYM == 0 && p.fast32 && kmax == BK - 1 && a_lin && b_lin) {
            // buffer loads: the stage's tile origin in the (scalar) resource,
            // the lane offset in the VGPR operand, r steps in the SGPR one
            const T* ab = uniform_ptr(TA ? p.A + k0 + m0 * p.lda : p.A + m0 + k0 * p.lda);
            const T* bb = uniform_ptr(TB ? p.B + n0 + k0 * p.ldb : p.B + k0 + n0 * p.ldb);
            const __amdgpu_buffer_rsrc_t ar = __builtin_amdgcn_make_buffer_rsrc((void*)ab, (short)0, 0x7fffffff, kBufWord3);
            const __amdgpu_buffer_rsrc_t br = __builtin_amdgcn_make_buffer_rsrc((void*)bb, (short)0, 0x7fffffff, kBufWord3);
            a_mode = 0;
#pragma unroll
            for (int r = 0; r < A_PER; ++r)
                ra[r] = kStep ? buffer_load<T>(ar, a_vo, (uint32_t)r * a_step) : buffer_load<T>(ar, aoff[kStep ? 0 : r], 0u);
#pragma unroll
            for (int r = 0; r < B_PER; ++r)
                rb[r] = kStep ? buffer_load<T>(br, b_vo, (uint32_t)r * b_step) : buffer_load<T>(br, boff[kStep ? 0 : r], 0u);
            return;
        }
        const int64_t ka = k0 + kb0;  // absolute depth of A (differs from k0 only for SYM splits)
        const T* abase = TA ? p.A + ka + m0 * p.lda : p.A + m0 + ka * p.lda;
        const T* bbase = TB ? p.B + n0 + k0 * p.ldb : p.B + k0 + n0 * p.ldb;
        // SYM: stored element (m,k) is direct when (m >= k) for lower storage
        bool all_direct = true, all_mirror = false;
        if (SYM == 1) {
            all_direct = m0 >= ka + BK - 1;
            all_mirror = m0 + BM - 1 < ka;
        } else if (SYM == 2) {
            all_direct = m0 + BM - 1 <= ka;
            all_mirror = m0 > ka + BK - 1;
        }
        const T* mbase = p.A + ka + m0 * p.lda;  // mirrored element (m,k) at A[k + m*lda]
        if (all_direct) {
            a_mode = 0;
#pragma unroll
            for (int r = 0; r < A_PER; ++r) {
                const int cm = min(a_m0 + r * AM_STEP, mmax), ck = min(a_k0 + r * AK_STEP, kmax);
                ra[r] = TA ? abase[ck + cm * lda32] : abase[cm + ck * lda32];
            }
        } else if (all_mirror) {
            a_mode = 1;
#pragma unroll
            for (int r = 0; r < A_PER; ++r) {
                const int cm = min(m_m0 + r * MM_STEP, mmax), ck = min(m_k0 + r * MK_STEP, kmax);
                ra[r] = mbase[ck + cm * lda32];
            }
        } else {
            a_mode = 0;
            const int64_t dk = ka - m0;
#pragma unroll
            for (int r = 0; r < A_PER; ++r) {
                const int cm = min(a_m0 + r * AM_STEP, mmax), ck = min(a_k0 + r * AK_STEP, kmax);
                const bool direct = SYM == 1 ? (cm - ck >= dk) : (cm - ck <= dk);
                ra[r] = direct ? abase[cm + ck * lda32] : mbase[ck + cm * lda32];
            }
        }
#pragma unroll
        for (int r = 0; r < B_PER; ++r) {
            const int cn = min(b_n0 + r * BN_STEP, nmax), ck = min(b_k0 + r * BK_STEP, kmax);
            rb[r] = TB ? bbase[cn + ck * ldb32] : bbase[ck + cn * ldb32];
        }
    };
    auto store_tiles = [&](int buf) {
        const int kmax = st_kmax;
        if (SYM == 0 && interior && kmax == BK - 1) {
#pragma unroll
            for (int r = 0; r < A_PER; ++r) As[buf][(a_k0 + r * AK_STEP) * LDA_S + a_m0 + r * AM_STEP] = ra[r];
#pragma unroll
            for (int r = 0; r < B_PER; ++r) Bs[buf][(b_k0 + r * BK_STEP) * LDB_S + b_n0 + r * BN_STEP] = rb[r];
            return;
        }
        if (SYM != 0 && a_mode == 1) {
#pragma unroll
            for (int r = 0; r < A_PER; ++r) {
                const int mm = m_m0 + r * MM_STEP, kk = m_k0 + r * MK_STEP;
                As[buf][kk * LDA_S + mm] = (mm <= mmax && kk <= kmax) ? ra[r] : T(0);
            }
        } else {
#pragma unroll
            for (int r = 0; r < A_PER; ++r) {
                const int mm = a_m0 + r * AM_STEP, kk = a_k0 + r * AK_STEP;
                As[buf][kk * LDA_S + mm] = (mm <= mmax && kk <= kmax) ? ra[r] : T(0);
            }
        }
#pragma unroll
        for (int r = 0; r < B_PER; ++r) {
            const int nn = b_n0 + r * BN_STEP, kk = b_k0 + r * BK_STEP;
            Bs[buf][kk * LDB_S + nn] = (nn <= nmax && kk <= kmax) ? rb[r] : T(0);
        }
    };

    const int lr = lane & 15;   // operand row inside a 16-block
    const int lk = lane >> 4;   // operand k inside a 4-deep step
    // Accumulator item (i, j, r) is element (row_of(i), col_of(j, r)) of the
    // tile: D = C^T, so a lane's items are 4 columns of C and the 16 lanes of
    // a group hold 16 consecutive rows.
    auto row_of = [&](int i) { return wm * WM + i * 16 + lr; };
    auto col_of = [&](int j, int r) { return wn * WN + j * 16 + (MF::kRowInterleaved ? 4 * r + lk : 4 * lk + r); };
    T* cbase = p.C + m0 + n0 * p.ldc;
    const int ldc32 = (int)p.ldc;

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int64_t nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
    }
    __syncthreads();

    for (int64_t kt = 0; kt < nk; ++kt) {
        const int buf = (int)(kt & 1);
        if (kt + 1 < nk) load_tiles((kt + 1) * BK);
        const T* as = As[buf];
        const T* bs = Bs[buf];
#pragma unroll
        for (int k4 = 0; k4 < BK / 4; ++k4) {
            const int kr = k4 * 4 + lk;
            T a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = as[kr * LDA_S + wm * WM + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = bs[kr * LDB_S + wn * WN + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MF::op(b[j], a[i], acc[i][j]);
        }
        if (kt + 1 < nk) store_tiles(buf ^ 1);
        __syncthreads();
    }

    // ---- epilogue: C = alpha*acc + beta*C.  The C values of a batch of
    // accumulator rows are all loaded (clamped addresses, no branches) before
    // any is stored, so a batch costs one memory round trip: one batch in
    // FP32, two in FP64 (the registers the K loop used hold them).
    const T alpha = p.alpha, beta = p.beta;
    const bool whole = m0 + BM <= M && n0 + BN <= N &&
                       (PART == 0 || (PART == 1 ? m0 >= n0 + BN - 1 : m0 + BM - 1 <= n0));
    constexpr int kBatches = sizeof(T) == 8 && FM % 2 == 0 ? 2 : 1;
    constexpr int FMB = FM / kBatches;
#pragma unroll
    for (int bt = 0; bt < kBatches; ++bt) {
        T out[FMB][FN][4];
        if (beta != T(0)) {
#pragma unroll
            for (int ii = 0; ii < FMB; ++ii) {
                const int cm = min(row_of(bt * FMB + ii), mmax);
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) out[ii][j][r] = cbase[cm + min(col_of(j, r), nmax) * ldc32];
            }
#pragma unroll
            for (int ii = 0; ii < FMB; ++ii)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) out[ii][j][r] = alpha * acc[bt * FMB + ii][j][r] + beta * out[ii][j][r];
        } else {
#pragma unroll
            for (int ii = 0; ii < FMB; ++ii)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) out[ii][j][r] = alpha * acc[bt * FMB + ii][j][r];
        }
        if (whole) {
#pragma unroll
            for (int ii = 0; ii < FMB; ++ii)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) cbase[row_of(bt * FMB + ii) + col_of(j, r) * ldc32] = out[ii][j][r];
        } else {
#pragma unroll
            for (int ii = 0; ii < FMB; ++ii) {
                const int mm = row_of(bt * FMB + ii);
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int nn = col_of(j, r);
                        const int64_t gm = m0 + mm, gn = n0 + nn;
                        bool ok = mm <= mmax && nn <= nmax;
                        if (PART == 1) ok = ok && (gm >= gn);
                        if (PART == 2) ok = ok && (gm <= gn);
                        if (ok) cbase[mm + nn * ldc32] = out[ii][j][r];
                    }
            }
        }
    }
}

// ---- split-BF16x3 variant of the FP32 trailing update ------------------
// C = alpha * A * B^T + beta * C on the lower trapezoid (A: M x K, B: N x K,
// both column-major), every FP32 operand split as x = hi + lo with hi, lo
// BF16 (lo the BF16 rounding of x - hi), and the product formed as
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
// about 16 significant bits per product, at three BF16 MFMAs per FP32-sized
// block instead of eight FP32 MFMAs.  The split happens once per element, when
// the tile is staged into LDS as [row][k] (k contiguous, so a lane's eight
// depth values are one 16-byte read).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// SPLIT: the operands come split already (launch_potrf_panel's split
// output, SplitPanel): the stage loads BF16 hi and lo directly, 16 depths of
// one row per thread as two 16-byte vectors each, with no conversion.
template <bool SPLIT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void gemm_bf16x3_kernel(KParams<float> p,
                                                                                                SplitPanel sp) {
    constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;  // LDS row: 32 depth values + pad
    constexpr int WM = 64, WN = 64, FM = 4, FN = 4;
    __shared__ __bf16 Ah[2][BM * LDK], Al[2][BM * LDK];
    __shared__ __bf16 Bh[2][BN * LDK], Bl[2][BN * LDK];

    const int64_t G = gridDim.x;
    int64_t pid = blockIdx.x;
    {
        const int64_t full = (G / kXcd) * kXcd;
        if (pid < full) pid = (pid % kXcd) * (G / kXcd) + pid / kXcd;
    }
    int64_t tm, tn;
    lower_trapezoid_tile_grouped(pid, p.tiles_n, tm, tn);
    const int64_t m0 = __builtin_amdgcn_readfirstlane((int)(tm * BM));
    const int64_t n0 = __builtin_amdgcn_readfirstlane((int)(tn * BN));
    const int64_t M = p.M, N = p.N, K = p.K;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lk = lane >> 4;

    // staging: thread t owns row t % 128 and depths 16*(t/128) .. +15 of both tiles
    const int srow = tid % 128, sk0 = 16 * (tid / 128);
    const int mmax = (int)min((int64_t)BM, M - m0) - 1, nmax = (int)min((int64_t)BN, N - n0) - 1;
    const int lda32 = (int)p.lda, ldb32 = (int)p.ldb;
    float ra[16], rb[16];
    bf16x8 sa[4], sb[4];  // SPLIT: hi(0..7), hi(8..15), lo(0..7), lo(8..15) of A's and B's row
    int st_kmax = 0;
    // full stages load through buffer descriptors: one fixed lane offset,
    // the depth step in the scalar offset (no per-element address VALU work)
    const bool buf_ok = (int64_t)(BM + BK * p.lda) * 4 < ((int64_t)1 << 30) && (int64_t)(BN + BK * p.ldb) * 4 < ((int64_t)1 << 30);
    const uint32_t a_vo = (uint32_t)(min(srow, mmax) + sk0 * lda32) * 4u, b_vo = (uint32_t)(min(srow, nmax) + sk0 * ldb32) * 4u;
    auto load_tiles = [&](int64_t k0) {
        const int kmax = (int)min((int64_t)BK, K - k0) - 1;
        st_kmax = kmax;
        if constexpr (SPLIT) {
            // stage origin: k-block k0 / 32, this thread's row and depths sk0..sk0+15
            const int64_t kb = (k0 >> 5) * sp.kstride + sk0;
            const int64_t ea = kb + (sp.a_row0 + m0 + min(srow, mmax)) * 32;
            const int64_t eb = kb + (sp.b_row0 + n0 + min(srow, nmax)) * 32;
            sa[0] = *reinterpret_cast<const bf16x8*>(sp.hi + ea);
            sa[1] = *reinterpret_cast<const bf16x8*>(sp.hi + ea + 8);
            sa[2] = *reinterpret_cast<const bf16x8*>(sp.lo + ea);
            sa[3] = *reinterpret_cast<const bf16x8*>(sp.lo + ea + 8);
            sb[0] = *reinterpret_cast<const bf16x8*>(sp.hi + eb);
            sb[1] = *reinterpret_cast<const bf16x8*>(sp.hi + eb + 8);
            sb[2] = *reinterpret_cast<const bf16x8*>(sp.lo + eb);
            sb[3] = *reinterpret_cast<const bf16x8*>(sp.lo + eb + 8);
            return;
        }
        if (buf_ok && kmax == BK - 1) {
            const __amdgpu_buffer_rsrc_t ar =
                __builtin_amdgcn_make_buffer_rsrc((void*)uniform_ptr(p.A + m0 + k0 * p.lda), (short)0, 0x7fffffff, kBufWord3);
            const __amdgpu_buffer_rsrc_t br =
                __builtin_amdgcn_make_buffer_rsrc((void*)uniform_ptr(p.B + n0 + k0 * p.ldb), (short)0, 0x7fffffff, kBufWord3);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                ra[r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ar, (int)a_vo, r * lda32 * 4, 0));
                rb[r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(br, (int)b_vo, r * ldb32 * 4, 0));
            }
            return;
        }
        const float* abase = p.A + m0 + k0 * p.lda;
        const float* bbase = p.B + n0 + k0 * p.ldb;
        const int cm = min(srow, mmax), cn = min(srow, nmax);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int ck = min(sk0 + r, kmax);
            ra[r] = abase[cm + ck * lda32];
            rb[r] = bbase[cn + ck * ldb32];
        }
    };
    auto split_store = [&](const float (&v)[16], bool row_ok, int kmax, __bf16* hi, __bf16* lo) {
        bf16x8 h0, h1, l0, l1;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float x = (row_ok && sk0 + r <= kmax) ? v[r] : 0.0f;
            const __bf16 xh = (__bf16)x;
            const __bf16 xl = (__bf16)(x - (float)xh);
            if (r < 8) { h0[r] = xh; l0[r] = xl; }
            else { h1[r - 8] = xh; l1[r - 8] = xl; }
        }
        *reinterpret_cast<bf16x8*>(hi) = h0;
        *reinterpret_cast<bf16x8*>(hi + 8) = h1;
        *reinterpret_cast<bf16x8*>(lo) = l0;
        *reinterpret_cast<bf16x8*>(lo + 8) = l1;
    };
    // SPLIT: rows past the operand and depths past K are zeroed here
    auto masked_store = [&](bf16x8 (&v)[4], bool row_ok, int kmax, __bf16* hi, __bf16* lo) {
        if (!row_ok || sk0 + 15 > kmax) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool k0_ok = row_ok && sk0 + e <= kmax, k1_ok = row_ok && sk0 + 8 + e <= kmax;
                v[0][e] = k0_ok ? v[0][e] : (__bf16)0.0f;
                v[1][e] = k1_ok ? v[1][e] : (__bf16)0.0f;
                v[2][e] = k0_ok ? v[2][e] : (__bf16)0.0f;
                v[3][e] = k1_ok ? v[3][e] : (__bf16)0.0f;
            }
        }
        *reinterpret_cast<bf16x8*>(hi) = v[0];
        *reinterpret_cast<bf16x8*>(hi + 8) = v[1];
        *reinterpret_cast<bf16x8*>(lo) = v[2];
        *reinterpret_cast<bf16x8*>(lo + 8) = v[3];
    };
    auto store_tiles = [&](int buf) {
        const int kmax = st_kmax;
        if constexpr (SPLIT) {
            masked_store(sa, srow <= mmax, kmax, &Ah[buf][srow * LDK + sk0], &Al[buf][srow * LDK + sk0]);
            masked_store(sb, srow <= nmax, kmax, &Bh[buf][srow * LDK + sk0], &Bl[buf][srow * LDK + sk0]);
            return;
        }
        split_store(ra, srow <= mmax, kmax, &Ah[buf][srow * LDK + sk0], &Al[buf][srow * LDK + sk0]);
        split_store(rb, srow <= nmax, kmax, &Bh[buf][srow * LDK + sk0], &Bl[buf][srow * LDK + sk0]);
    };

    // item (i, j, r) of the accumulators is tile element (row_of(i), col_of(j, r))
    auto row_of = [&](int i) { return wm * WM + i * 16 + lr; };
    auto col_of = [&](int j, int r) { return wn * WN + j * 16 + 4 * lk + r; };
    float* cbase = p.C + m0 + n0 * p.ldc;
    const int ldc32 = (int)p.ldc;
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load_tiles(0);
        store_tiles(0);
    }
    __syncthreads();
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int buf = (int)(kt & 1);
        if (kt + 1 < nk) load_tiles((kt + 1) * BK);
        bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int row = wm * WM + i * 16 + lr;
            ah[i] = *reinterpret_cast<const bf16x8*>(&Ah[buf][row * LDK + 8 * lk]);
            al[i] = *reinterpret_cast<const bf16x8*>(&Al[buf][row * LDK + 8 * lk]);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int row = wn * WN + j * 16 + lr;
            bh[j] = *reinterpret_cast<const bf16x8*>(&Bh[buf][row * LDK + 8 * lk]);
            bl[j] = *reinterpret_cast<const bf16x8*>(&Bl[buf][row * LDK + 8 * lk]);
        }
        // D = B-block x A-block^T: lane items are 4 columns (n) of one row (m) of C
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
            }
        if (kt + 1 < nk) store_tiles(buf ^ 1);
        __syncthreads();
    }

    // epilogue: every C value is loaded (clamped addresses, no branches)
    // before any is stored, so the tile costs one memory round trip
    const float alpha = p.alpha, beta = p.beta;
    const bool whole = m0 + BM <= M && n0 + BN <= N && m0 >= n0 + BN - 1;  // inside the lower trapezoid
    float out[FM][FN][4];
    if (beta != 0.f) {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int cm = min(row_of(i), mmax);
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[i][j][r] = cbase[cm + min(col_of(j, r), nmax) * ldc32];
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[i][j][r] = alpha * acc[i][j][r] + beta * out[i][j][r];
    } else {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[i][j][r] = alpha * acc[i][j][r];
    }
    if (whole) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) cbase[row_of(i) + col_of(j, r) * ldc32] = out[i][j][r];
    } else {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int mm = row_of(i);
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int nn = col_of(j, r);
                    if (mm <= mmax && nn <= nmax && m0 + mm >= n0 + nn) cbase[mm + nn * ldc32] = out[i][j][r];
                }
        }
    }
}

// C = alpha * sum_s ws[s] + beta * C over the split-K partials.
template <typename T>
__global__ void split_reduce_kernel(const T* __restrict__ ws, int splits, int64_t M, int64_t N, T alpha, T beta, T* C,
                                    int64_t ldc) {
    const int64_t total = M * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        T acc = T(0);
        for (int s = 0; s < splits; ++s) acc += ws[s * total + e];
        const int64_t i = e % M, j = e / M;
        T* c = C + i + j * ldc;
        *c = beta == T(0) ? alpha * acc : alpha * acc + beta * *c;
    }
}

// the gemm_kernel instance for d's transposes on prepared parameters
template <typename T, int BM, int BN, int BK, int PART>
rmh_status launch_kparams(const GemmDesc<T>& d, const KParams<T>& p, dim3 grid, size_t dyn, hipStream_t stream) {
    const dim3 block(kThreads);
    if (!d.transA && !d.transB)
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, false, PART>), grid, block, dyn, stream, p);
    else if (!d.transA && d.transB)
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, true, PART>), grid, block, dyn, stream, p);
    else if (d.transA && !d.transB)
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, true, false, PART>), grid, block, dyn, stream, p);
    else
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, true, true, PART>), grid, block, dyn, stream, p);
    return check_launch("gemm_kernel");
}

// compute units of the current device (cached per device)
int device_cu_count() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int c = cache[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        cache[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

template <typename T, int BM, int BN, int BK, int PART>
rmh_status launch_cfg(const GemmDesc<T>& d, hipStream_t stream) {
    KParams<T> p{};
    p.M = d.M; p.N = d.N; p.K = d.K;
    p.alpha = d.alpha; p.beta = d.beta;
    p.A = d.A; p.lda = d.lda; p.B = d.B; p.ldb = d.ldb; p.C = d.C; p.ldc = d.ldc;
    p.tiles_m = (d.M + BM - 1) / BM;
    p.tiles_n = (d.N + BN - 1) / BN;
    p.abort_flag = d.abort_flag;
    p.batch_ny = d.batch_ny > 0 ? d.batch_ny : 1;
    p.sAy = d.sAy; p.sAz = d.sAz; p.sBy = d.sBy; p.sBz = d.sBz; p.sCy = d.sCy; p.sCz = d.sCz;
    p.k_chunk = d.k_chunk;
    p.split_ws = d.split_ws;
    p.k_from_row = d.k_from_row ? 1 : 0;
    {
        const int64_t a_max = d.transA ? (BK - 1) + (int64_t)(BM - 1) * d.lda : (BM - 1) + (int64_t)(BK - 1) * d.lda;
        const int64_t b_max = d.transB ? (BN - 1) + (int64_t)(BK - 1) * d.ldb : (BK - 1) + (int64_t)(BN - 1) * d.ldb;
        const int64_t lim = ((int64_t)1 << 31) / (int64_t)sizeof(T) - 1;  // buffer range: 2^31 - 1 bytes
        p.fast32 = (a_max < lim && b_max < lim) ? 1 : 0;
    }
    int64_t ntiles;
    if (PART == 0) {
        ntiles = p.tiles_m * p.tiles_n;
    } else if (PART == 1) {  // rows >= cols: tile-row i has min(i+1, tiles_n) tiles
        const int64_t tn = p.tiles_n, tmm = p.tiles_m;
        ntiles = tn * (tn + 1) / 2 + (tmm - tn) * tn;
    } else {
        const int64_t tn = p.tiles_m, tmm = p.tiles_n;
        ntiles = tn * (tn + 1) / 2 + (tmm - tn) * tn;
    }
    if (ntiles <= 0) return RMH_STATUS_SUCCESS;
    if (ntiles > 0x7fffffffLL) return fail(RMH_STATUS_INVALID_VALUE, "gemm: problem too large for one launch");
    p.ntiles = ntiles;
    // Trapezoid on 128 x 128 tiles whose last round of workgroups would fill
    // only part of the device: the full rounds run here, and the remaining
    // tiles of the order run as 64 x 64 quarters in a second launch, four
    // times as many workgroups of a quarter of the work (four fit on a
    // compute unit where two 128 x 128 ones do).  Both tile shapes form each
    // element's sum in the same order (same MFMA, same depth step, same
    // epilogue), so the result is bitwise the same.  The FP64 quarters keep
    // most of the large tiles' MFMA rate and pay off up to half a round; the
    // FP32 ones run at about half of it (38% against 74% busy), so four of
    // them per compute unit take as long as the two tiles they replace and
    // FP32 splits up to a quarter of a round only.
    if constexpr (BM == 128 && BN == 128 && PART != 0) {
        const int64_t slots = 2 * (int64_t)device_cu_count();  // two workgroups per compute unit
        const int64_t rounds = ntiles / slots, rest = ntiles - rounds * slots;
        const int64_t max_rest = sizeof(T) == 8 ? slots / 2 : slots / 4;
        if (d.tail_split && d.splits <= 1 && d.batch <= 1 && !d.k_from_row && !d.one_per_cu && slots > 0 &&
            rounds >= 1 && rest > 0 && rest <= max_rest) {
            p.ntiles = rounds * slots;
            RMH_TRY((launch_kparams<T, 128, 128, BK, PART>(d, p, dim3((unsigned)p.ntiles), 0, stream)));
            KParams<T> q = p;
            q.tiles_m = (d.M + 63) / 64;
            q.tiles_n = (d.N + 63) / 64;
            q.sub_t0 = rounds * slots;
            q.sub_tbig = PART == 1 ? p.tiles_n : p.tiles_m;
            q.ntiles = 4 * rest;
            return launch_kparams<T, 64, 64, BK, PART>(d, q, dim3((unsigned)q.ntiles), 0, stream);
        }
    }
    const dim3 grid((unsigned)ntiles, (unsigned)std::max(1, d.splits), (unsigned)std::max<int64_t>(1, d.batch));
    // one workgroup per compute unit: pad the LDS request past half of it
    constexpr size_t kStatic = 2 * BK * (BM + kPad) * sizeof(T) + 2 * BK * (BN + kPad) * sizeof(T);
    constexpr size_t kOnePerCu = 96 * 1024;
    const size_t dyn = (d.one_per_cu && kStatic < kOnePerCu) ? kOnePerCu - kStatic : 0;
    RMH_TRY((launch_kparams<T, BM, BN, BK, PART>(d, p, grid, dyn, stream)));
    if (d.splits > 1) {
        const int64_t total = d.M * d.N;
        const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL(split_reduce_kernel<T>, dim3(g), dim3(256), 0, stream, d.split_ws, d.splits, d.M, d.N, d.alpha,
                           d.beta, d.C, d.ldc);
        RMH_TRY(check_launch("split_reduce_kernel"));
    }
    return RMH_STATUS_SUCCESS;
}

template <typename T, int BM, int BN, int BK>
rmh_status launch_part(const GemmDesc<T>& d, hipStream_t stream) {
    switch (d.part) {
        case Part::Full: return launch_cfg<T, BM, BN, BK, 0>(d, stream);
        case Part::Lower: return launch_cfg<T, BM, BN, BK, 1>(d, stream);
        case Part::Upper: return launch_cfg<T, BM, BN, BK, 2>(d, stream);
    }
    return fail(RMH_STATUS_INVALID_VALUE, "gemm: bad part");
}

template <typename T, int BM, int BN, int BK, int SYM>
rmh_status launch_symm_cfg(const GemmDesc<T>& d, hipStream_t stream) {
    KParams<T> p{};
    p.M = d.M; p.N = d.N; p.K = d.K;
    p.alpha = d.alpha; p.beta = d.beta;
    p.A = d.A; p.lda = d.lda; p.B = d.B; p.ldb = d.ldb; p.C = d.C; p.ldc = d.ldc;
    p.tiles_m = (d.M + BM - 1) / BM;
    p.tiles_n = (d.N + BN - 1) / BN;
    p.abort_flag = d.abort_flag;
    p.batch_ny = 1;
    p.sAy = p.sAz = p.sBy = p.sBz = p.sCy = p.sCz = 0;
    p.k_from_row = 0;
    p.fast32 = 0;
    int splits = 1;
    p.k_chunk = d.K;
    p.split_ws = nullptr;
    if (d.splits > 1 && d.split_ws != nullptr && d.K > 0) {  // split-K: partials, then one reduction
        p.k_chunk = ((d.K + d.splits - 1) / d.splits + BK - 1) / BK * BK;
        splits = (int)((d.K + p.k_chunk - 1) / p.k_chunk);
        p.split_ws = d.split_ws;
        if (splits == 1) p.k_chunk = d.K;
    }
    const int64_t ntiles = p.tiles_m * p.tiles_n;
    if (ntiles <= 0) return RMH_STATUS_SUCCESS;
    p.ntiles = ntiles;
    const dim3 grid((unsigned)ntiles, (unsigned)splits);
    if (d.transB)
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, true, 0, SYM>), grid, dim3(kThreads), 0, stream, p);
    else
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, false, 0, SYM>), grid, dim3(kThreads), 0, stream, p);
    RMH_TRY(check_launch("gemm_kernel(symm)"));
    if (splits > 1) {
        const int64_t total = d.M * d.N;
        const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL(split_reduce_kernel<T>, dim3(g), dim3(256), 0, stream, d.split_ws, splits, d.M, d.N, d.alpha,
                           d.beta, d.C, d.ldc);
        RMH_TRY(check_launch("split_reduce_kernel"));
    }
    return RMH_STATUS_SUCCESS;
}

}  // namespace

template <typename T>
rmh_status launch_symm(const GemmDesc<T>& d, bool upper, hipStream_t stream) {
    if (d.M <= 0 || d.N <= 0) return RMH_STATUS_SUCCESS;
    if (d.M != d.K) return fail(RMH_STATUS_INVALID_VALUE, "symm: A must be square");
    if (d.lda > kMaxLd || d.ldb > kMaxLd || d.ldc > kMaxLd)
        return fail(RMH_STATUS_INVALID_VALUE, "symm: leading dimension above 2^23");
    constexpr int BK = sizeof(T) == 8 ? 16 : 32;
    const bool big = pick_tile(d.M, d.N, Part::Full) == Tile::T128x128;
    if (big) return upper ? launch_symm_cfg<T, 128, 128, BK, 2>(d, stream) : launch_symm_cfg<T, 128, 128, BK, 1>(d, stream);
    return upper ? launch_symm_cfg<T, 64, 64, BK, 2>(d, stream) : launch_symm_cfg<T, 64, 64, BK, 1>(d, stream);
}
template rmh_status launch_symm<double>(const GemmDesc<double>&, bool, hipStream_t);
template rmh_status launch_symm<float>(const GemmDesc<float>&, bool, hipStream_t);

static rmh_status launch_bf16x3(const GemmDesc<float>& d, const SplitPanel* sp, hipStream_t stream) {
    if (d.M <= 0 || d.N <= 0) return RMH_STATUS_SUCCESS;
    if (d.part != Part::Lower || d.transA || !d.transB || d.splits > 1 || d.batch > 1 || d.k_from_row || d.M < d.N)
        return fail(RMH_STATUS_INVALID_VALUE, "gemm_bf16x3: only C -= A B^T on a lower trapezoid");
    if (d.lda > kMaxLd || d.ldb > kMaxLd || d.ldc > kMaxLd)
        return fail(RMH_STATUS_INVALID_VALUE, "gemm_bf16x3: leading dimension above 2^23");
    KParams<float> p{};
    p.M = d.M; p.N = d.N; p.K = d.K;
    p.alpha = d.alpha; p.beta = d.beta;
    p.A = d.A; p.lda = d.lda; p.B = d.B; p.ldb = d.ldb; p.C = d.C; p.ldc = d.ldc;
    p.tiles_m = (d.M + 127) / 128;
    p.tiles_n = (d.N + 127) / 128;
    const int64_t tn = p.tiles_n, tmm = p.tiles_m;
    const int64_t ntiles = tn * (tn + 1) / 2 + (tmm - tn) * tn;
    if (ntiles <= 0) return RMH_STATUS_SUCCESS;
    p.ntiles = ntiles;
    p.batch_ny = 1;
    p.k_chunk = d.K;
    constexpr size_t kStatic = 4 * 128 * 40 * 2 * 2;  // Ah, Al, Bh, Bl, double-buffered
    const size_t dyn = (d.one_per_cu && kStatic < 96 * 1024) ? 96 * 1024 - kStatic : 0;
    if (sp != nullptr) {
        if (sp->hi == nullptr || sp->lo == nullptr || sp->kstride < 32 * (sp->a_row0 + d.M) ||
            sp->kstride < 32 * (sp->b_row0 + d.N) || sp->a_row0 < 0 || sp->b_row0 < 0)
            return fail(RMH_STATUS_INVALID_VALUE, "gemm_bf16x3: split panel does not cover the operands");
        hipLaunchKernelGGL(gemm_bf16x3_kernel<true>, dim3((unsigned)ntiles), dim3(kThreads), dyn, stream, p, *sp);
    } else {
        hipLaunchKernelGGL(gemm_bf16x3_kernel<false>, dim3((unsigned)ntiles), dim3(kThreads), dyn, stream, p,
                           SplitPanel{});
    }
    return check_launch("gemm_bf16x3_kernel");
}

rmh_status launch_gemm_bf16x3(const GemmDesc<float>& d, hipStream_t stream) { return launch_bf16x3(d, nullptr, stream); }

rmh_status launch_gemm_bf16x3_split(const GemmDesc<float>& d, const SplitPanel& sp, hipStream_t stream) {
    return launch_bf16x3(d, &sp, stream);
}

Tile pick_tile(int64_t M, int64_t N, Part part) {
    const int64_t t128 = ((M + 127) / 128) * ((N + 127) / 128) / (part == Part::Full ? 1 : 2);
    if (t128 >= 512) return Tile::T128x128;
    return Tile::T64x64;
}

template <typename T>
rmh_status launch_gemm(const GemmDesc<T>& d_in, Tile tile, hipStream_t stream) {
    GemmDesc<T> d = d_in;
    if (d.M <= 0 || d.N <= 0) return RMH_STATUS_SUCCESS;
    if (d.k_from_row && (d.splits > 1 || d.batch > 1))
        return fail(RMH_STATUS_INVALID_VALUE, "gemm: triangular depth skip needs one plain launch");
    if (d.splits > 1) {
        if (d.split_ws == nullptr || d.part != Part::Full || d.batch > 1 || d.K <= 0)
            return fail(RMH_STATUS_INVALID_VALUE, "gemm: split-K needs a workspace and a plain rectangle");
        d.k_chunk = (d.K + d.splits - 1) / d.splits;
        d.splits = (int)((d.K + d.k_chunk - 1) / d.k_chunk);  // no empty chunk
        if (d.splits == 1) d.split_ws = nullptr;
    } else {
        d.splits = 1;
        d.k_chunk = d.K;
    }
    // tile-relative offsets are 32-bit: (tile extent) * ld must stay below 2^31
    if (d.lda > kMaxLd || d.ldb > kMaxLd || d.ldc > kMaxLd)
        return fail(RMH_STATUS_INVALID_VALUE, "gemm: leading dimension above 2^23");
    if (d.part == Part::Lower && d.M < d.N) return fail(RMH_STATUS_INVALID_VALUE, "gemm: lower part needs M >= N");
    if (d.part == Part::Upper && d.N < d.M) return fail(RMH_STATUS_INVALID_VALUE, "gemm: upper part needs N >= M");
    if (d.part != Part::Full && (tile == Tile::T128x64 || tile == Tile::T64x128))
        return fail(RMH_STATUS_INVALID_VALUE, "gemm: triangular parts need square tiles");
    constexpr int BK = sizeof(T) == 8 ? 16 : 32;
    switch (tile) {
        case Tile::T128x128: return launch_part<T, 128, 128, BK>(d, stream);
        case Tile::T64x64: return launch_part<T, 64, 64, BK>(d, stream);
        case Tile::T128x64: return launch_cfg<T, 128, 64, BK, 0>(d, stream);
        case Tile::T64x128: return launch_cfg<T, 64, 128, BK, 0>(d, stream);
    }
    return fail(RMH_STATUS_INVALID_VALUE, "gemm: bad tile");
}

template rmh_status launch_gemm<double>(const GemmDesc<double>&, Tile, hipStream_t);
template rmh_status launch_gemm<float>(const GemmDesc<float>&, Tile, hipStream_t);

}  // namespace rmh
