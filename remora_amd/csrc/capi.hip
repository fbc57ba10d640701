// C-ABI entry points of libremora_hip and the host-side drivers behind them:
// handle lifetime, BLAS-3/2 front ends, the blocked TRSM, and the
// device-side blocked right-looking POTRF with its solve.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rmh_context.hpp"

namespace rmh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

rmh_status fail(rmh_status st, const std::string& msg) {
    set_error(msg);
    return st;
}

rmh_status check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RMH_STATUS_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    return RMH_STATUS_SUCCESS;
}

// ------------------------------------------------------------ context --
rmh_status Context::activate() {
    RMH_TRY_HIP(hipSetDevice(device));
    return RMH_STATUS_SUCCESS;
}

rmh_status Context::reserve(size_t bytes) {
    if (bytes <= ws_bytes) return RMH_STATUS_SUCCESS;
    // Kernels queued earlier may still use the old buffer.
    RMH_TRY_HIP(hipDeviceSynchronize());
    if (ws != nullptr) RMH_TRY_HIP(hipFree(ws));
    ws = nullptr;
    ws_bytes = 0;
    size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&ws, want) != hipSuccess) {
        (void)hipGetLastError();
        ws = nullptr;
        return fail(RMH_STATUS_ALLOC_FAILED, "workspace allocation of " + std::to_string(want) + " bytes failed");
    }
    ws_bytes = want;
    return RMH_STATUS_SUCCESS;
}

rmh_status Context::reserve_slot(int slot, size_t bytes, void** ptr) {
    if (slot == WS_MAIN) {
        RMH_TRY(reserve(bytes));
        *ptr = ws;
        return RMH_STATUS_SUCCESS;
    }
    if (bytes > aux_bytes[slot]) {
        RMH_TRY_HIP(hipDeviceSynchronize());
        if (aux[slot] != nullptr) RMH_TRY_HIP(hipFree(aux[slot]));
        aux[slot] = nullptr;
        aux_bytes[slot] = 0;
        size_t want = std::max(bytes, (size_t)1 << 20);
        if (hipMalloc(&aux[slot], want) != hipSuccess) {
            (void)hipGetLastError();
            aux[slot] = nullptr;
            return fail(RMH_STATUS_ALLOC_FAILED, "workspace allocation of " + std::to_string(want) + " bytes failed");
        }
        aux_bytes[slot] = want;
    }
    *ptr = aux[slot];
    return RMH_STATUS_SUCCESS;
}

rmh_status Context::side_stream(hipStream_t* st, hipEvent_t* ready, hipEvent_t* done) {
    if (side == nullptr) {
        RMH_TRY_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));  // default priority
        RMH_TRY_HIP(hipEventCreateWithFlags(&ev_a, hipEventDisableTiming));
        RMH_TRY_HIP(hipEventCreateWithFlags(&ev_b, hipEventDisableTiming));
    }
    *st = side;
    *ready = ev_a;
    *done = ev_b;
    return RMH_STATUS_SUCCESS;
}

// The stream of the look-ahead trailing updates: its CU mask leaves out
// `rsv` compute units of each of the 8 XCDs.  Mask bit b stands for CU b of
// the device; the bits left out are (per + 1) x + 8 y for XCD x and y < rsv
// (per = CUs per XCD), which are rsv units of XCD x both where the bits run
// through the XCDs in turn (bit b on XCD b % 8) and where each XCD owns a
// contiguous range of per bits.  Returns the caller's stream unchanged when
// the device does not divide that way.
rmh_status Context::trail_stream(int rsv, hipStream_t* st, hipEvent_t* in, hipEvent_t* out) {
    if (trail != nullptr && trail_rsv != rsv) {
        RMH_TRY_HIP(hipStreamSynchronize(trail));
        RMH_TRY_HIP(hipStreamDestroy(trail));
        trail = nullptr;
    }
    if (trail == nullptr) {
        constexpr int kXcd = 8;  // XCDs of an MI355X
        hipDeviceProp_t prop;
        RMH_TRY_HIP(hipGetDeviceProperties(&prop, device));
        const int ncu = prop.multiProcessorCount, per = ncu / kXcd;
        if (rsv <= 0 || ncu % kXcd != 0 || per % 8 != 0 || per < 8 * rsv) return RMH_STATUS_SUCCESS;
        std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
        for (int b = 0; b < ncu; ++b) mask[(size_t)b / 32] |= 1u << (b % 32);
        for (int x = 0; x < kXcd; ++x)
            for (int y = 0; y < rsv; ++y) {
                const int b = (per + 1) * x + 8 * y;
                mask[(size_t)b / 32] &= ~(1u << (b % 32));
            }
        RMH_TRY_HIP(hipExtStreamCreateWithCUMask(&trail, (uint32_t)mask.size(), mask.data()));  // default priority
        trail_rsv = rsv;
        if (ev_c == nullptr) {
            RMH_TRY_HIP(hipEventCreateWithFlags(&ev_c, hipEventDisableTiming));
            RMH_TRY_HIP(hipEventCreateWithFlags(&ev_d, hipEventDisableTiming));
        }
    }
    *st = trail;
    *in = ev_c;
    *out = ev_d;
    return RMH_STATUS_SUCCESS;
}

rmh_status Context::prof_begin(hipStream_t s, int cls, double flops, int* slot) {
    *slot = -1;
    if (((prof_mask >> cls) & 1) == 0) return RMH_STATUS_SUCCESS;
    ProfRec r;
    r.cls = cls;
    r.flops = flops;
    RMH_TRY_HIP(hipEventCreate(&r.a));
    RMH_TRY_HIP(hipEventCreate(&r.b));
    RMH_TRY_HIP(hipEventRecord(r.a, s));
    recs.push_back(r);
    *slot = (int)recs.size() - 1;
    return RMH_STATUS_SUCCESS;
}

rmh_status Context::prof_end(hipStream_t s, int slot) {
    if (slot < 0) return RMH_STATUS_SUCCESS;
    RMH_TRY_HIP(hipEventRecord(recs[slot].b, s));
    return RMH_STATUS_SUCCESS;
}

void Context::prof_clear() {
    for (auto& r : recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    recs.clear();
}

Context::~Context() {
    prof_clear();
    (void)hipSetDevice(device);
    if (side != nullptr) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
        (void)hipEventDestroy(ev_a);
        (void)hipEventDestroy(ev_b);
    }
    if (trail != nullptr) {
        (void)hipStreamSynchronize(trail);
        (void)hipStreamDestroy(trail);
    }
    if (ev_conv != nullptr) (void)hipEventDestroy(ev_conv);
    if (ev_conv0 != nullptr) (void)hipEventDestroy(ev_conv0);
    if (ev_c != nullptr) {
        (void)hipEventDestroy(ev_c);
        (void)hipEventDestroy(ev_d);
    }
    if (ws != nullptr) (void)hipFree(ws);
    for (int i = 0; i < WS_COUNT; ++i)
        if (aux[i] != nullptr) (void)hipFree(aux[i]);
    if (batch_ctr != nullptr) (void)hipFree(batch_ctr);
}

// ------------------------------------------------------- TRSM driver --
// Blocked recursive solve with inverted 64 x 64 diagonal blocks: the leaves
// multiply by the inverse block in place, and they and the off-diagonal
// work are MFMA GEMM.  The split mirrors the reference's trsm_recursive
// (include/remora/kernels/default/trsm.hpp:227-261): halves in whole
// blocks, lower solves front-to-back, upper back-to-front.
namespace {

constexpr int kTrsmNb = 64;

template <typename T>
struct TrsmPlan {
    bool left, eff_upper, trans;
    const T* A;
    int64_t lda;
    T* B;
    int64_t ldb;
    int64_t na, other;
    const T* Dinv;  // nb = 64: inv(op(A)_bb) per block, contiguous, ld 64
    hipStream_t s;
    int64_t nb = kTrsmNb;
    // nb > 64: D holds inv(S_bb) as an na x nb strip (ld na), S the lower
    // triangular A or A^T, and tmp (nb x other or other x nb) takes the
    // leaf product before it is copied back
    const T* D = nullptr;
    T* tmp = nullptr;
};

// Block T[r0:r1, c0:c1] of op(A) as a GEMM operand.
template <typename T>
void tri_operand(const TrsmPlan<T>& p, int64_t r0, int64_t c0, const T** ptr, bool* tr) {
    if (!p.trans) { *ptr = p.A + r0 + c0 * p.lda; *tr = false; }
    else { *ptr = p.A + c0 + r0 * p.lda; *tr = true; }
}

template <typename T>
rmh_status trsm_rec(const TrsmPlan<T>& p, int64_t b0, int64_t b1) {
    const int64_t nb = p.nb;
    if (b1 - b0 == 1 && nb > kTrsmNb) {
        // wide leaf: X_b = inv(op(A)_bb) B_b (left) or B_b inv(op(A)_bb)
        // (right) into tmp, then back into B; inv(op(A)_bb) = Db or Db^T
        const int64_t kb = std::min(nb, p.na - b0 * nb), r0 = b0 * nb;
        const T* Db = p.D + r0;
        GemmDesc<T> g;
        g.alpha = T(1);
        g.beta = T(0);
        g.K = kb;
        if (p.left) {
            g.M = kb; g.N = p.other;
            g.A = Db; g.lda = p.na; g.transA = p.eff_upper;
            g.B = p.B + r0; g.ldb = p.ldb;
            g.C = p.tmp; g.ldc = kb;
        } else {
            g.M = p.other; g.N = kb;
            g.A = p.B + r0 * p.ldb; g.lda = p.ldb;
            g.B = Db; g.ldb = p.na; g.transB = p.eff_upper;
            g.C = p.tmp; g.ldc = p.other;
        }
        RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), p.s));
        if (p.left)
            RMH_TRY_HIP(hipMemcpy2DAsync(p.B + r0, (size_t)p.ldb * sizeof(T), p.tmp, (size_t)kb * sizeof(T),
                                         (size_t)kb * sizeof(T), (size_t)p.other, hipMemcpyDeviceToDevice, p.s));
        else
            RMH_TRY_HIP(hipMemcpy2DAsync(p.B + r0 * p.ldb, (size_t)p.ldb * sizeof(T), p.tmp,
                                         (size_t)p.other * sizeof(T), (size_t)p.other * sizeof(T), (size_t)kb,
                                         hipMemcpyDeviceToDevice, p.s));
        return RMH_STATUS_SUCCESS;
    }
    if (b1 - b0 == 1) {
        // leaf: X_b = D B_b (left) or B_b D (right), D = inv(op(A)_bb), in
        // place on the MFMA GEMM.  The product is one 64-row (left) or
        // 64-column (right) tile high, so every workgroup reads the part of
        // B it overwrites, and reads all of it before its epilogue writes.
        const int64_t kb = std::min(nb, p.na - b0 * nb), r0 = b0 * nb;
        const T* D = p.Dinv + b0 * nb * nb;
        GemmDesc<T> g;
        g.alpha = T(1);
        g.beta = T(0);
        g.K = kb;
        if (p.left) {
            g.M = kb; g.N = p.other;
            g.A = D; g.lda = nb;
            g.B = p.B + r0; g.ldb = p.ldb;
            g.C = p.B + r0; g.ldc = p.ldb;
        } else {
            g.M = p.other; g.N = kb;
            g.A = p.B + r0 * p.ldb; g.lda = p.ldb;
            g.B = D; g.ldb = nb;
            g.C = p.B + r0 * p.ldb; g.ldc = p.ldb;
        }
        return launch_gemm<T>(g, Tile::T64x64, p.s);
    }
    const int64_t mid = b0 + (b1 - b0) / 2;
    const int64_t r0 = b0 * nb, rm = mid * nb, r1 = std::min(b1 * nb, p.na);
    GemmDesc<T> g;
    g.alpha = T(-1);
    g.beta = T(1);
    if (p.left) {
        // forward (eff lower): B[rm:r1] -= T[rm:r1, r0:rm] X[r0:rm]
        // backward (eff upper): B[r0:rm] -= T[r0:rm, rm:r1] X[rm:r1]
        const bool fwd = !p.eff_upper;
        const int64_t cr0 = fwd ? rm : r0, cr1 = fwd ? r1 : rm;  // rows of C
        const int64_t xr0 = fwd ? r0 : rm, xr1 = fwd ? rm : r1;  // rows of X
        if (fwd) RMH_TRY(trsm_rec(p, b0, mid));
        else RMH_TRY(trsm_rec(p, mid, b1));
        g.M = cr1 - cr0; g.N = p.other; g.K = xr1 - xr0;
        tri_operand(p, cr0, xr0, &g.A, &g.transA);
        g.lda = p.lda;
        g.B = p.B + xr0; g.ldb = p.ldb; g.transB = false;
        g.C = p.B + cr0; g.ldc = p.ldb;
        RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), p.s));
        if (fwd) return trsm_rec(p, mid, b1);
        return trsm_rec(p, b0, mid);
    }
    // right side, X T = B: eff lower runs back-to-front, eff upper front-to-back.
    //   eff lower: B[:, r0:rm] -= X[:, rm:r1] T[rm:r1, r0:rm]
    //   eff upper: B[:, rm:r1] -= X[:, r0:rm] T[r0:rm, rm:r1]
    const bool fwd = p.eff_upper;
    const int64_t cc0 = fwd ? rm : r0, cc1 = fwd ? r1 : rm;  // columns of C
    const int64_t xc0 = fwd ? r0 : rm, xc1 = fwd ? rm : r1;  // columns of X
    if (fwd) RMH_TRY(trsm_rec(p, b0, mid));
    else RMH_TRY(trsm_rec(p, mid, b1));
    g.M = p.other; g.N = cc1 - cc0; g.K = xc1 - xc0;
    g.A = p.B + xc0 * p.ldb; g.lda = p.ldb; g.transA = false;
    tri_operand(p, xc0, cc0, &g.B, &g.transB);
    g.ldb = p.lda;
    g.C = p.B + cc0 * p.ldb; g.ldc = p.ldb;
    RMH_TRY(launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), p.s));
    if (fwd) return trsm_rec(p, mid, b1);
    return trsm_rec(p, b0, mid);
}

}  // namespace

template <typename T>
rmh_status trsm_impl(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                     const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    if (trsm_blocked_applies(left ? m : n, left ? n : m))
        return trsm_blocked<T>(h, left, upper, trans, unit, m, n, alpha, A, lda, B, ldb, s);
    if (alpha != T(1)) RMH_TRY(launch_scale<T>(B, ldb, m, n, alpha, s));
    if (alpha == T(0)) return RMH_STATUS_SUCCESS;
    const int64_t na = left ? m : n, other = left ? n : m;
    if (na >= 2048 && other >= 512) {
        // many right-hand sides: 512-wide leaves from block inverses built
        // by doubling, so the recursion stops before its GEMMs get thin
        constexpr int64_t nbd = 512;
        auto up = [](size_t e) { return (e + 63) / 64 * 64; };
        const size_t d_el = up((size_t)na * nbd), w_el = up((size_t)na * nbd / 2 + 64 * 64),
                     t_el = up((size_t)nbd * other);
        RMH_TRY(h->reserve((d_el + w_el + t_el) * sizeof(T)));
        T* D = static_cast<T*>(h->ws);
        RMH_TRY(diag_block_inverses<T>(A, lda, na, upper, unit, nbd, D, D + d_el, s));
        TrsmPlan<T> p{left, upper != trans, trans, A, lda, B, ldb, na, other, nullptr, s};
        p.nb = nbd;
        p.D = D;
        p.tmp = D + d_el + w_el;
        return trsm_rec(p, 0, (na + nbd - 1) / nbd);
    }
    const int64_t nblk = (na + kTrsmNb - 1) / kTrsmNb;
    RMH_TRY(h->reserve((size_t)nblk * kTrsmNb * kTrsmNb * sizeof(T)));
    T* Dinv = static_cast<T*>(h->ws);
    RMH_TRY(launch_trtri_blocks<T>(A, lda, na, kTrsmNb, upper, trans, unit, Dinv, s));
    TrsmPlan<T> p{left, upper != trans, trans, A, lda, B, ldb, na, other, Dinv, s};
    return trsm_rec(p, 0, nblk);
}

// ------------------------------------------------------ POTRF driver --
// Blocked right-looking Cholesky on the lower view V of A: V(i,j) is
// A[i + j*lda] for uplo = lower and A[j + i*lda] for uplo = upper (so the
// upper factor U is computed as the lower factor of A^T, as the reference's
// potrf_recursive does through trans(), include/remora/kernels/default/potrf.hpp:136-145).
//
// For each panel of width nb:
//   for each diagonal block of width nbi inside the panel:
//     inner  -- update of this block column by the panel's earlier columns
//     panel  -- factor V_jj in LDS and solve the rows below against it
//               (one fused launch, potrf_panel_kernel)
//   trailing -- V22 -= P P^T with K = nb (MFMA SYRK on the lower trapezoid),
//               the next panel's columns first so that panel can be factored
//               on a second stream meanwhile (look-ahead)
// A failed pivot records LAPACK's info (the first failure wins); the
// remaining launches still run, on values that no longer matter.
// Trailing update of potrf: the MFMA GEMM in T, or for FP32 factors the
// split-BF16x3 kernel when the handle asks for it (lower, 128 x 128 tiles).
template <typename T>
static rmh_status launch_trailing(Context* h, const GemmDesc<T>& g, Tile tile, const SplitPanel* sp, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if ((h->bf16x3_trailing || h->spotrf_bf16x3) && tile == Tile::T128x128 && g.part == Part::Lower)
            return sp != nullptr ? launch_gemm_bf16x3_split(g, *sp, st) : launch_gemm_bf16x3(g, st);
    }
    return launch_gemm<T>(g, tile, st);
}

// Tile of the factorisation's own updates (trapezoids).  min_tiles 0: the
// general rule (pick_tile).  Otherwise 128 x 128 tiles once the trapezoid has
// min_tiles of them, counted exactly.  Split-BF16x3 factors switch at 256,
// which sends the next panel's updates to the split-BF16x3 kernel instead of
// FP32 64 x 64 tiles (13.6 -> 12.7 ms at n = 16384).  FP64 and FP32 keep the
// general rule: an exact count from 512 on gave FP64 nothing and FP32
// 0.2 ms, within the benchmark's noise.
static Tile factor_tile(int64_t M, int64_t N, Part part, int64_t min_tiles) {
    if (min_tiles <= 0) return pick_tile(M, N, part);
    const int64_t tl = ((part == Part::Upper ? N : M) + 127) / 128, ts = ((part == Part::Upper ? M : N) + 127) / 128;
    const int64_t t128 = ts * (ts + 1) / 2 + (tl - ts) * ts;
    return t128 >= min_tiles ? Tile::T128x128 : Tile::T64x64;
}

template <typename T>
size_t potrf_workspace(const Context* h, int64_t n) {
    return 256 + (size_t)std::max<int64_t>(n, 1) * h->nbi * sizeof(T);
}

// The factorisation on a caller-provided workspace of potrf_workspace<T>(h, n)
// bytes: [0, 256) info when the caller passes none, then the side buffer of
// the diagonal-block factors (n x nbi, ld n).  With `plain` the call runs on
// the caller's stream alone and records no profile events, so it creates and
// allocates nothing.
template <typename T>
rmh_status potrf_core(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* d_info, void* ws, bool plain,
                      hipStream_t s) {
    const bool bf16x3 = sizeof(T) == 4 && (h->bf16x3_trailing || h->spotrf_bf16x3);
    const int64_t nb = potrf_panel_width(h, sizeof(T) == 8, bf16x3);
    const int64_t nbi = (int64_t)h->nbi;
    const int64_t min_tiles = bf16x3 ? 256 : 0;  // factor_tile
    // look-ahead mode: 0 single stream; 1 look-ahead; 2 with one trailing
    // workgroup per compute unit; 3 and 4: modes 2 and 1 with four-wavefront
    // panel workgroups
    const int mode = h->la_mode >= 0 ? h->la_mode : (sizeof(T) == 8 ? 1 : 4);
    const bool share_cus = mode == 2 || mode == 3, panel_narrow = mode == 3 || mode == 4;
    int64_t* scratch = static_cast<int64_t*>(ws);
    int64_t* info = d_info != nullptr ? d_info : scratch;
    T* dfac = reinterpret_cast<T*>(static_cast<char*>(ws) + 256);
    RMH_TRY(launch_fill_i64(info, 1, 0, s));
    if (n == 0) return RMH_STATUS_SUCCESS;
    const int64_t rs = upper ? lda : 1, cs = upper ? 1 : lda;
    auto V = [&](int64_t i, int64_t j) { return A + i * rs + j * cs; };
    const bool lookahead = !plain && mode != 0 && n > 2 * nb;
    // split-BF16x3 trailing updates (FP32 factors, lower storage): the panel
    // kernel also writes its solved rows split into BF16 hi + lo, in one of
    // two buffers by panel parity (look-ahead factors panel k+1 while the
    // update reads panel k's), laid out in 32-deep blocks of rows:
    // panel column k of row r at (k / 32) * (32 n) + 32 r + k % 32
    __bf16* split[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    const int64_t sp_kstride = 32 * n;
    if constexpr (sizeof(T) == 4) {
        if (!plain && !upper && (h->bf16x3_trailing || h->spotrf_bf16x3) && nbi <= 32) {
            const size_t per = (size_t)((nb + 31) / 32) * (size_t)sp_kstride;  // elements per array
            void* base = nullptr;
            RMH_TRY(h->reserve_slot(WS_SPLIT, 4 * per * sizeof(__bf16), &base));
            __bf16* b = static_cast<__bf16*>(base);
            split[0][0] = b;
            split[0][1] = b + per;
            split[1][0] = b + 2 * per;
            split[1][1] = b + 3 * per;
        }
    }
    auto parity = [&](int64_t k0) { return (int)((k0 / nb) & 1); };
    auto prof_begin = [&](hipStream_t st, int cls, double flops, int* slot) {
        *slot = -1;
        return plain ? RMH_STATUS_SUCCESS : h->prof_begin(st, cls, flops, slot);
    };

    // Factor the panel V[k0:, k0:k0+kw] left-looking, one fused kernel per
    // block column [j0, j0+jb): it applies the updates of the panel's earlier
    // columns [k0, j0), factors the diagonal block and solves the rows below.
    // chain of fused panel launches over the block columns of [s0, s0+sw),
    // each left-looking from s0; the panel starts at p0
    auto chain = [&](int64_t p0, int64_t s0, int64_t sw, hipStream_t st) -> rmh_status {
        for (int64_t j0 = s0; j0 < s0 + sw; j0 += nbi) {
            const int64_t jb = std::min(nbi, s0 + sw - j0);
            const int64_t rows = n - j0, K = j0 - s0;
            const double flops = 2.0 * K * ((double)jb * (jb + 1) / 2.0 + (double)(rows - jb) * jb) +
                                 (double)jb * jb * jb / 3.0 + (double)(rows - jb) * jb * jb;
            int slot;
            RMH_TRY(prof_begin(st, PROF_PANEL, flops, &slot));
            // split output: panel column (j0 - p0) of the row; the chain's
            // first column (a multiple of 32 past p0) folded into the base
            const int64_t soff = ((s0 - p0) >> 5) * sp_kstride + j0 * 32;
            __bf16* sh = split[parity(p0)][0];
            __bf16* sl = split[parity(p0)][1];
            RMH_TRY(launch_potrf_panel<T>(V(j0, s0), rs, cs, rows, K, (int)jb, info, j0, dfac + j0, n, st,
                                          sh != nullptr ? sh + soff : nullptr, sl != nullptr ? sl + soff : nullptr,
                                          sp_kstride, panel_narrow));
            RMH_TRY(h->prof_end(st, slot));
        }
        return RMH_STATUS_SUCCESS;
    };
    // V[c0:, c0:c1] -= S S[c0:c1]^T, S = V[c0:, s0:s0+sw]: the panel's columns
    // right of a factored part take its update
    auto in_panel_update = [&](int64_t s0, int64_t sw, int64_t c0, int64_t c1, hipStream_t st) -> rmh_status {
        if (c0 >= c1) return RMH_STATUS_SUCCESS;
        GemmDesc<T> g;
        g.alpha = T(-1); g.beta = T(1); g.K = sw;
        g.M = n - c0; g.N = c1 - c0; g.part = Part::Lower;
        if (!upper) {
            g.A = A + c0 + s0 * lda; g.lda = lda; g.transA = false;
            g.B = g.A; g.ldb = lda; g.transB = true;
        } else {
            g.M = c1 - c0; g.N = n - c0; g.part = Part::Upper;
            g.A = A + s0 + c0 * lda; g.lda = lda; g.transA = true;
            g.B = g.A; g.ldb = lda; g.transB = false;
        }
        g.C = A + c0 + c0 * lda; g.ldc = lda;
        const double flops = 2.0 * sw * ((double)(c1 - c0) * (c1 - c0 + 1) / 2.0 + (double)(n - c1) * (c1 - c0));
        int slot;
        RMH_TRY(prof_begin(st, PROF_PANEL, flops, &slot));
        g.tail_split = true;
        RMH_TRY(launch_gemm<T>(g, factor_tile(g.M, g.N, g.part, min_tiles), st));
        return h->prof_end(st, slot);
    };
    // recursive halving of [s0, s0+sw) down to leaves of at most `leafw`
    // columns (a multiple of nbi), each a chain left-looking from its first
    // column: factor the left half, update the right half with it, factor
    // the right half.  The chains' K loops stay short (at most leafw - nbi
    // deep); the depth above a leaf is taken by MFMA GEMMs.
    const int64_t lq = std::max<int64_t>(nbi, 32);  // split output needs 32-aligned chain starts
    const int64_t leafw = std::max<int64_t>(lq, (int64_t)h->nb_leaf / lq * lq);
    auto rec = [&](auto& self, int64_t p0, int64_t s0, int64_t sw, hipStream_t st) -> rmh_status {
        if (sw <= leafw) return chain(p0, s0, sw, st);
        const int64_t h1 = ((sw / 2 + leafw - 1) / leafw) * leafw;
        RMH_TRY(self(self, p0, s0, h1, st));
        RMH_TRY(in_panel_update(s0, h1, s0 + h1, s0 + sw, st));
        return self(self, p0, s0 + h1, sw - h1, st);
    };
    // Factor the panel V[k0:, k0:k0+kw].
    auto factor_panel = [&](int64_t k0, int64_t kw, hipStream_t st) -> rmh_status {
        return rec(rec, k0, k0, kw, st);
    };
    // V[c0:, c0:c1] -= P P[c0:c1]^T on the lower trapezoid, P = V[c0:, k0:k0+kw]
    auto trailing = [&](int64_t k0, int64_t kw, int64_t c0, int64_t c1, hipStream_t st) -> rmh_status {
        const int64_t rows = n - c0, cols = c1 - c0;
        if (rows <= 0 || cols <= 0) return RMH_STATUS_SUCCESS;
        GemmDesc<T> g;
        g.alpha = T(-1); g.beta = T(1); g.K = kw;
        if (!upper) {
            g.M = rows; g.N = cols; g.part = Part::Lower;
            g.A = A + c0 + k0 * lda; g.lda = lda; g.transA = false;
            g.B = g.A; g.ldb = lda; g.transB = true;
        } else {
            g.M = cols; g.N = rows; g.part = Part::Upper;
            g.A = A + k0 + c0 * lda; g.lda = lda; g.transA = true;
            g.B = g.A; g.ldb = lda; g.transB = false;
        }
        g.C = A + c0 + c0 * lda; g.ldc = lda;
        const double flops = 2.0 * kw * ((double)cols * (cols + 1) / 2.0 + (double)(rows - cols) * cols);
        const Tile tile = factor_tile(g.M, g.N, g.part, min_tiles);
        g.one_per_cu = lookahead && share_cus && tile == Tile::T128x128;
        g.tail_split = true;
        int slot;
        RMH_TRY(prof_begin(st, tile == Tile::T128x128 ? PROF_TRAILING : PROF_TRAILING_SMALL, flops, &slot));
        SplitPanel sp;
        sp.hi = split[parity(k0)][0];
        sp.lo = split[parity(k0)][1];
        sp.kstride = sp_kstride;
        sp.a_row0 = c0;
        sp.b_row0 = c0;
        RMH_TRY(launch_trailing<T>(h, g, tile, sp.hi != nullptr ? &sp : nullptr, st));
        return h->prof_end(st, slot);
    };

    // Look-ahead: while the caller's stream s applies panel k's update to
    // the columns right of panel k+1, the side stream factors panel k+1
    // (whose columns s updated first).  Both read panel k only; the column
    // ranges they write are disjoint.
    // With compute units reserved (rmh_set_potrf_reserve) the steps whose
    // trailing rows fall in the reservation window run their updates on m, a
    // stream whose CU mask leaves those units to the side stream's panel
    // launches; `cur` is the stream of the current step's updates, and work
    // moves between s and m behind an event.
    hipStream_t s2 = nullptr, m = s;
    hipEvent_t ev_ready = nullptr, ev_done = nullptr, ev_in = nullptr, ev_out = nullptr;
    if (lookahead) RMH_TRY(h->side_stream(&s2, &ev_ready, &ev_done));
    const int rsv = h->reserve_cus;
    const int64_t rsv_lo = (int64_t)h->rsv_rows_lo, rsv_hi = (int64_t)h->rsv_rows_hi;
    if (lookahead && rsv > 0) RMH_TRY(h->trail_stream(rsv, &m, &ev_in, &ev_out));
    hipStream_t cur = s;
    auto move_to = [&](hipStream_t want) -> rmh_status {
        if (want == cur) return RMH_STATUS_SUCCESS;
        hipEvent_t e = want == m ? ev_in : ev_out;
        RMH_TRY_HIP(hipEventRecord(e, cur));
        RMH_TRY_HIP(hipStreamWaitEvent(want, e, 0));
        cur = want;
        return RMH_STATUS_SUCCESS;
    };
    RMH_TRY(factor_panel(0, std::min(nb, n), s));
    if (h->cols_ready != nullptr) RMH_TRY_HIP(hipStreamWaitEvent(s, h->cols_ready, 0));
    for (int64_t k0 = 0; k0 < n; k0 += nb) {
        const int64_t kw = std::min(nb, n - k0);
        const int64_t k1 = k0 + kw;  // next panel
        if (k1 >= n) break;
        const int64_t kw1 = std::min(nb, n - k1);
        const int64_t rows = n - k1;
        RMH_TRY(move_to(m != s && rows > rsv_lo && rows <= rsv_hi ? m : s));
        RMH_TRY(trailing(k0, kw, k1, k1 + kw1, cur));  // next panel's columns first
        if (lookahead) {
            RMH_TRY_HIP(hipEventRecord(ev_ready, cur));
            RMH_TRY_HIP(hipStreamWaitEvent(s2, ev_ready, 0));
            RMH_TRY(factor_panel(k1, kw1, s2));
            RMH_TRY_HIP(hipEventRecord(ev_done, s2));
            RMH_TRY(trailing(k0, kw, k1 + kw1, n, cur));
            RMH_TRY_HIP(hipStreamWaitEvent(cur, ev_done, 0));
        } else {
            RMH_TRY(trailing(k0, kw, k1 + kw1, n, cur));
            RMH_TRY(factor_panel(k1, kw1, cur));
        }
    }
    RMH_TRY(move_to(s));
    return launch_scatter_diag<T>(A, rs, cs, n, (int)nbi, dfac, n, info, s);
}

template <typename T>
rmh_status potrf_impl(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* d_info, hipStream_t s) {
    RMH_TRY(h->reserve(potrf_workspace<T>(h, n)));
    return potrf_core<T>(h, upper, n, A, lda, d_info, h->ws, false, s);
}

template <typename T>
rmh_status potrs_impl(Context* h, bool upper, int64_t n, int64_t nrhs, const T* A, int64_t lda, T* B, int64_t ldb,
                      hipStream_t s) {
    if (n == 0 || nrhs == 0) return RMH_STATUS_SUCCESS;
    // lower: L y = b, L^T x = y ; upper: U^T y = b, U x = y
    if (trsm_blocked_applies(n, nrhs)) {  // one set of block inverses for both solves
        RMH_TRY(h->reserve(trsm_inverse_workspace<T>(n, nrhs)));
        TrsmInverses<T> inv;
        RMH_TRY(trsm_prepare<T>(A, lda, n, upper, false, h->ws, &inv, s));
        RMH_TRY(trsm_apply<T>(inv, true, upper, n, nrhs, T(1), B, ldb, s));
        return trsm_apply<T>(inv, true, !upper, n, nrhs, T(1), B, ldb, s);
    }
    RMH_TRY(trsm_impl<T>(h, true, upper, upper, false, n, nrhs, T(1), A, lda, B, ldb, s));
    return trsm_impl<T>(h, true, upper, !upper, false, n, nrhs, T(1), A, lda, B, ldb, s);
}

template <typename T>
rmh_status trsv_impl(Context* h, bool upper, bool trans, bool unit, int64_t n, const T* A, int64_t lda, T* x,
                     int64_t incx, hipStream_t s) {
    if (n == 0) return RMH_STATUS_SUCCESS;
    if (incx == 1) return trsm_impl<T>(h, true, upper, trans, unit, n, 1, T(1), A, lda, x, n, s);
    // gather a strided vector into the tail of the workspace, solve, scatter
    const int64_t nblk = (n + kTrsmNb - 1) / kTrsmNb;
    const size_t dinv_bytes = (size_t)nblk * kTrsmNb * kTrsmNb * sizeof(T);
    const size_t off = (dinv_bytes + 255) / 256 * 256;
    RMH_TRY(h->reserve(off + (size_t)n * sizeof(T)));
    T* tmp = reinterpret_cast<T*>(static_cast<char*>(h->ws) + off);
    RMH_TRY(launch_strided_copy<T>(x, incx, tmp, 1, n, s));
    RMH_TRY(trsm_impl<T>(h, true, upper, trans, unit, n, 1, T(1), A, lda, tmp, n, s));
    return launch_strided_copy<T>(tmp, 1, x, incx, n, s);
}

template rmh_status potrf_impl<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, hipStream_t);
template size_t potrf_workspace<double>(const Context*, int64_t);
template rmh_status potrf_core<double>(Context*, bool, int64_t, double*, int64_t, int64_t*, void*, bool, hipStream_t);
template rmh_status potrf_impl<float>(Context*, bool, int64_t, float*, int64_t, int64_t*, hipStream_t);
template rmh_status potrs_impl<double>(Context*, bool, int64_t, int64_t, const double*, int64_t, double*, int64_t,
                                       hipStream_t);
template rmh_status potrs_impl<float>(Context*, bool, int64_t, int64_t, const float*, int64_t, float*, int64_t,
                                      hipStream_t);

}  // namespace rmh

using namespace rmh;

// ------------------------------------------------------- validation ---
namespace {

Context* ctx(rmh_handle h) { return reinterpret_cast<Context*>(h); }

#define RMH_CHECK_HANDLE(h)                                                          \
    do {                                                                             \
        if ((h) == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");   \
        RMH_TRY(ctx(h)->activate());                                                 \
    } while (0)

#define RMH_REQUIRE(cond, msg)                                                       \
    do {                                                                             \
        if (!(cond)) return fail(RMH_STATUS_INVALID_VALUE, (msg));                   \
    } while (0)

inline size_t mx1(size_t v) { return v > 1 ? v : 1; }

template <typename T>
rmh_status gemm_entry(rmh_handle h, int transA, int transB, size_t m, size_t n, size_t k, T alpha, const T* A,
                      size_t lda, const T* B, size_t ldb, T beta, T* C, size_t ldc, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(transA ? k : m), "gemm: lda too small");
    RMH_REQUIRE(ldb >= mx1(transB ? n : k), "gemm: ldb too small");
    RMH_REQUIRE(ldc >= mx1(m), "gemm: ldc too small");
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(C != nullptr, "gemm: C is NULL");
    RMH_REQUIRE(k == 0 || alpha == T(0) || (A != nullptr && B != nullptr), "gemm: A or B is NULL");
    GemmDesc<T> g;
    g.M = (int64_t)m; g.N = (int64_t)n; g.K = (alpha == T(0)) ? 0 : (int64_t)k;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = (int64_t)lda; g.transA = transA != 0;
    g.B = B; g.ldb = (int64_t)ldb; g.transB = transB != 0;
    g.C = C; g.ldc = (int64_t)ldc;
    return launch_gemm<T>(g, pick_tile(g.M, g.N, Part::Full), hs(stream));
}

template <typename T>
rmh_status syrk_entry(rmh_handle h, int upper, int trans, size_t n, size_t k, T alpha, const T* A, size_t lda,
                      T beta, T* C, size_t ldc, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(trans ? k : n), "syrk: lda too small");
    RMH_REQUIRE(ldc >= mx1(n), "syrk: ldc too small");
    if (n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(C != nullptr, "syrk: C is NULL");
    RMH_REQUIRE(k == 0 || alpha == T(0) || A != nullptr, "syrk: A is NULL");
    GemmDesc<T> g;
    g.M = (int64_t)n; g.N = (int64_t)n; g.K = (alpha == T(0)) ? 0 : (int64_t)k;
    g.alpha = alpha; g.beta = beta;
    g.part = upper ? Part::Upper : Part::Lower;
    // C += alpha op(A) op(A)^T: op(A) = A (n x k) or A^T (A is k x n)
    g.A = A; g.lda = (int64_t)lda; g.transA = trans != 0;
    g.B = A; g.ldb = (int64_t)lda; g.transB = trans == 0;
    g.C = C; g.ldc = (int64_t)ldc;
    g.tail_split = true;
    return launch_gemm<T>(g, pick_tile(g.M, g.N, g.part), hs(stream));
}

template <typename T>
rmh_status trsm_entry(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, T alpha,
                      const T* A, size_t lda, T* B, size_t ldb, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    const size_t na = left ? m : n;
    RMH_REQUIRE(lda >= mx1(na), "trsm: lda too small");
    RMH_REQUIRE(ldb >= mx1(m), "trsm: ldb too small");
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(B != nullptr && (A != nullptr || alpha == T(0)), "trsm: NULL pointer");
    return trsm_impl<T>(ctx(h), left != 0, upper != 0, trans != 0, unit != 0, (int64_t)m, (int64_t)n, alpha, A,
                        (int64_t)lda, B, (int64_t)ldb, hs(stream));
}

template <typename T>
rmh_status trsv_entry(rmh_handle h, int upper, int trans, int unit, size_t n, const T* A, size_t lda, T* x,
                      size_t incx, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n), "trsv: lda too small");
    RMH_REQUIRE(incx >= 1, "trsv: incx must be positive");
    if (n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && x != nullptr, "trsv: NULL pointer");
    return trsv_impl<T>(ctx(h), upper != 0, trans != 0, unit != 0, (int64_t)n, A, (int64_t)lda, x, (int64_t)incx,
                        hs(stream));
}

template <typename T>
rmh_status copy_matrix_entry(rmh_handle h, int trans, size_t m, size_t n, const T* A, size_t lda, T* B, size_t ldb,
                             rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(trans ? n : m), "copy_matrix: lda too small");
    RMH_REQUIRE(ldb >= mx1(m), "copy_matrix: ldb too small");
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && B != nullptr, "copy_matrix: NULL pointer");
    return copy_matrix_impl<T>(trans != 0, (int64_t)m, (int64_t)n, A, (int64_t)lda, B, (int64_t)ldb, hs(stream));
}

template <typename T>
rmh_status trmv_entry(rmh_handle h, int upper, int trans, int unit, size_t n, const T* A, size_t lda, T* x,
                      size_t incx, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n), "trmv: lda too small");
    RMH_REQUIRE(incx >= 1, "trmv: incx must be positive");
    if (n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && x != nullptr, "trmv: NULL pointer");
    return trmv_impl<T>(ctx(h), upper != 0, trans != 0, unit != 0, (int64_t)n, A, (int64_t)lda, x, (int64_t)incx,
                        hs(stream));
}

template <typename T>
rmh_status trmm_entry(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, T alpha,
                      const T* A, size_t lda, T* B, size_t ldb, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    const size_t na = left ? m : n;
    RMH_REQUIRE(lda >= mx1(na), "trmm: lda too small");
    RMH_REQUIRE(ldb >= mx1(m), "trmm: ldb too small");
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && B != nullptr, "trmm: NULL pointer");
    return trmm_impl<T>(ctx(h), left != 0, upper != 0, trans != 0, unit != 0, (int64_t)m, (int64_t)n, alpha, A,
                        (int64_t)lda, B, (int64_t)ldb, hs(stream));
}

template <typename T>
rmh_status gemv_entry(rmh_handle h, int trans, size_t m, size_t n, T alpha, const T* A, size_t lda, const T* x,
                      size_t incx, T beta, T* y, size_t incy, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(m), "gemv: lda too small");
    RMH_REQUIRE(incx >= 1 && incy >= 1, "gemv: increments must be positive");
    const size_t ylen = trans ? n : m;
    if (ylen == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(y != nullptr, "gemv: y is NULL");
    const size_t xlen = trans ? m : n;
    RMH_REQUIRE(xlen == 0 || (A != nullptr && x != nullptr), "gemv: NULL pointer");
    return launch_gemv<T>(trans != 0, (int64_t)m, (int64_t)n, alpha, A, (int64_t)lda, x, (int64_t)incx, beta, y,
                          (int64_t)incy, hs(stream));
}

template <typename T>
rmh_status adjoint_entry(rmh_handle h, int upper, size_t n, const T* L, size_t ldl, const T* Lbar, size_t ldlb, T* Abar,
                         size_t ldab, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(ldl >= mx1(n) && ldlb >= mx1(n) && ldab >= mx1(n), "potrf_adjoint: leading dimension too small");
    if (n == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(L != nullptr && Lbar != nullptr && Abar != nullptr, "potrf_adjoint: NULL pointer");
    RMH_REQUIRE(Abar != L && Abar != Lbar, "potrf_adjoint: Abar aliases an input");
    return potrf_adjoint_impl<T>(ctx(h), upper != 0, (int64_t)n, L, (int64_t)ldl, Lbar, (int64_t)ldlb, Abar,
                                 (int64_t)ldab, hs(stream));
}
template <typename T>
rmh_status pstrf_entry(rmh_handle h, int upper, size_t n, T* A, size_t lda, int64_t* d_piv, int64_t* d_rank,
                       rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n), "pstrf: lda too small");
    RMH_REQUIRE(lda <= (size_t(1) << 23), "pstrf: leading dimension above 2^23");
    RMH_REQUIRE(d_rank != nullptr, "pstrf: d_rank is NULL");
    RMH_REQUIRE(n == 0 || (A != nullptr && d_piv != nullptr), "pstrf: NULL pointer");
    return pstrf_impl<T>(ctx(h), upper != 0, (int64_t)n, A, (int64_t)lda, d_piv, d_rank, hs(stream));
}
template <typename T>
rmh_status chol_update_entry(rmh_handle h, int upper, size_t n, T alpha, T beta, T* L, size_t ldl, const T* v,
                             size_t incv, int64_t* d_info, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(ldl >= mx1(n), "chol_update: ldl too small");
    RMH_REQUIRE(incv >= 1, "chol_update: incv must be positive");
    RMH_REQUIRE(d_info != nullptr, "chol_update: d_info is NULL");
    RMH_REQUIRE(n == 0 || (L != nullptr && (v != nullptr || beta == T(0))), "chol_update: NULL pointer");
    return chol_update_impl<T>(ctx(h), upper != 0, (int64_t)n, alpha, beta, L, (int64_t)ldl, v, (int64_t)incv, d_info,
                               hs(stream));
}
template <typename T>
rmh_status potrf_entry(rmh_handle h, int upper, size_t n, T* A, size_t lda, int64_t* d_info, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n), "potrf: lda too small");
    RMH_REQUIRE(n == 0 || A != nullptr, "potrf: A is NULL");
    return potrf_impl<T>(ctx(h), upper != 0, (int64_t)n, A, (int64_t)lda, d_info, hs(stream));
}

template <typename T>
rmh_status potrs_entry(rmh_handle h, int upper, size_t n, size_t nrhs, const T* A, size_t lda, T* B, size_t ldb,
                       rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n), "potrs: lda too small");
    RMH_REQUIRE(ldb >= mx1(n), "potrs: ldb too small");
    if (n == 0 || nrhs == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && B != nullptr, "potrs: NULL pointer");
    return potrs_impl<T>(ctx(h), upper != 0, (int64_t)n, (int64_t)nrhs, A, (int64_t)lda, B, (int64_t)ldb,
                         hs(stream));
}

}  // namespace

extern "C" {

rmh_status rmh_create(rmh_handle* handle, int device) {
    if (handle == nullptr) return fail(RMH_STATUS_INVALID_VALUE, "create: handle pointer is NULL");
    *handle = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return fail(RMH_STATUS_HIP_ERROR, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(RMH_STATUS_INVALID_VALUE, "create: device " + std::to_string(device) + " not present");
    Context* c = new Context();
    c->device = device;
    rmh_status st = c->activate();
    if (st == RMH_STATUS_SUCCESS) st = noalloc_scratch_init(device);
    if (st != RMH_STATUS_SUCCESS) {
        delete c;
        return st;
    }
    *handle = reinterpret_cast<rmh_handle>(c);
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_destroy(rmh_handle handle) {
    if (handle == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    delete ctx(handle);
    return RMH_STATUS_SUCCESS;
}

const char* rmh_status_string(rmh_status status) {
    switch (status) {
        case RMH_STATUS_SUCCESS: return "success";
        case RMH_STATUS_INVALID_VALUE: return "invalid value";
        case RMH_STATUS_INVALID_HANDLE: return "invalid handle";
        case RMH_STATUS_HIP_ERROR: return "HIP runtime error";
        case RMH_STATUS_ALLOC_FAILED: return "allocation failed";
        case RMH_STATUS_NOT_POSITIVE_DEFINITE: return "matrix is not positive definite";
        case RMH_STATUS_NOT_CONVERGED: return "iterative refinement did not converge";
        case RMH_STATUS_NOT_SUPPORTED: return "not supported";
    }
    return "unknown status";
}

const char* rmh_last_error(void) { return g_last_error.c_str(); }

rmh_status rmh_dgemm(rmh_handle h, int transA, int transB, size_t m, size_t n, size_t k, double alpha,
                     const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C, size_t ldc,
                     rmh_stream_t stream) {
    return gemm_entry<double>(h, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream);
}
rmh_status rmh_sgemm(rmh_handle h, int transA, int transB, size_t m, size_t n, size_t k, float alpha,
                     const float* A, size_t lda, const float* B, size_t ldb, float beta, float* C, size_t ldc,
                     rmh_stream_t stream) {
    return gemm_entry<float>(h, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream);
}
rmh_status rmh_dsyrk(rmh_handle h, int upper, int trans, size_t n, size_t k, double alpha, const double* A,
                     size_t lda, double beta, double* C, size_t ldc, rmh_stream_t stream) {
    return syrk_entry<double>(h, upper, trans, n, k, alpha, A, lda, beta, C, ldc, stream);
}
rmh_status rmh_ssyrk(rmh_handle h, int upper, int trans, size_t n, size_t k, float alpha, const float* A,
                     size_t lda, float beta, float* C, size_t ldc, rmh_stream_t stream) {
    return syrk_entry<float>(h, upper, trans, n, k, alpha, A, lda, beta, C, ldc, stream);
}
rmh_status rmh_dtrsm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, double alpha,
                     const double* A, size_t lda, double* B, size_t ldb, rmh_stream_t stream) {
    return trsm_entry<double>(h, left, upper, trans, unit, m, n, alpha, A, lda, B, ldb, stream);
}
rmh_status rmh_strsm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, float alpha,
                     const float* A, size_t lda, float* B, size_t ldb, rmh_stream_t stream) {
    return trsm_entry<float>(h, left, upper, trans, unit, m, n, alpha, A, lda, B, ldb, stream);
}
rmh_status rmh_dtrsv(rmh_handle h, int upper, int trans, int unit, size_t n, const double* A, size_t lda, double* x,
                     size_t incx, rmh_stream_t stream) {
    return trsv_entry<double>(h, upper, trans, unit, n, A, lda, x, incx, stream);
}
rmh_status rmh_strsv(rmh_handle h, int upper, int trans, int unit, size_t n, const float* A, size_t lda, float* x,
                     size_t incx, rmh_stream_t stream) {
    return trsv_entry<float>(h, upper, trans, unit, n, A, lda, x, incx, stream);
}
rmh_status rmh_dcopy_matrix(rmh_handle h, int trans, size_t m, size_t n, const double* A, size_t lda, double* B,
                            size_t ldb, rmh_stream_t stream) {
    return copy_matrix_entry<double>(h, trans, m, n, A, lda, B, ldb, stream);
}
rmh_status rmh_scopy_matrix(rmh_handle h, int trans, size_t m, size_t n, const float* A, size_t lda, float* B,
                            size_t ldb, rmh_stream_t stream) {
    return copy_matrix_entry<float>(h, trans, m, n, A, lda, B, ldb, stream);
}
rmh_status rmh_dtrmv(rmh_handle h, int upper, int trans, int unit, size_t n, const double* A, size_t lda, double* x,
                     size_t incx, rmh_stream_t stream) {
    return trmv_entry<double>(h, upper, trans, unit, n, A, lda, x, incx, stream);
}
rmh_status rmh_strmv(rmh_handle h, int upper, int trans, int unit, size_t n, const float* A, size_t lda, float* x,
                     size_t incx, rmh_stream_t stream) {
    return trmv_entry<float>(h, upper, trans, unit, n, A, lda, x, incx, stream);
}
rmh_status rmh_dtrmm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, double alpha,
                     const double* A, size_t lda, double* B, size_t ldb, rmh_stream_t stream) {
    return trmm_entry<double>(h, left, upper, trans, unit, m, n, alpha, A, lda, B, ldb, stream);
}
rmh_status rmh_strmm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, float alpha,
                     const float* A, size_t lda, float* B, size_t ldb, rmh_stream_t stream) {
    return trmm_entry<float>(h, left, upper, trans, unit, m, n, alpha, A, lda, B, ldb, stream);
}
rmh_status rmh_dgemv(rmh_handle h, int trans, size_t m, size_t n, double alpha, const double* A, size_t lda,
                     const double* x, size_t incx, double beta, double* y, size_t incy, rmh_stream_t stream) {
    return gemv_entry<double>(h, trans, m, n, alpha, A, lda, x, incx, beta, y, incy, stream);
}
rmh_status rmh_sgemv(rmh_handle h, int trans, size_t m, size_t n, float alpha, const float* A, size_t lda,
                     const float* x, size_t incx, float beta, float* y, size_t incy, rmh_stream_t stream) {
    return gemv_entry<float>(h, trans, m, n, alpha, A, lda, x, incx, beta, y, incy, stream);
}
rmh_status rmh_dpotrf(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_info,
                      rmh_stream_t stream) {
    return potrf_entry<double>(h, upper, n, A, lda, d_info, stream);
}
rmh_status rmh_spotrf(rmh_handle h, int upper, size_t n, float* A, size_t lda, int64_t* d_info,
                      rmh_stream_t stream) {
    return potrf_entry<float>(h, upper, n, A, lda, d_info, stream);
}
rmh_status rmh_dpotrs(rmh_handle h, int upper, size_t n, size_t nrhs, const double* A, size_t lda, double* B,
                      size_t ldb, rmh_stream_t stream) {
    return potrs_entry<double>(h, upper, n, nrhs, A, lda, B, ldb, stream);
}
rmh_status rmh_spotrs(rmh_handle h, int upper, size_t n, size_t nrhs, const float* A, size_t lda, float* B,
                      size_t ldb, rmh_stream_t stream) {
    return potrs_entry<float>(h, upper, n, nrhs, A, lda, B, ldb, stream);
}

rmh_status rmh_dsposv(rmh_handle h, int upper, rmh_factor_precision prec, size_t n, size_t nrhs, const double* A,
                      size_t lda, const double* B, size_t ldb, double* X, size_t ldx, double tol, int max_iter,
                      int* iters, double* berr, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(lda >= mx1(n) && ldb >= mx1(n) && ldx >= mx1(n), "dsposv: leading dimension too small");
    RMH_REQUIRE(max_iter >= 0, "dsposv: max_iter must be >= 0");
    RMH_REQUIRE(prec == RMH_FACTOR_FP32 || prec == RMH_FACTOR_BF16X3, "dsposv: unknown factor precision");
    if (iters) *iters = 0;
    if (berr) *berr = 0.0;
    if (n == 0 || nrhs == 0) return RMH_STATUS_SUCCESS;
    RMH_REQUIRE(A != nullptr && B != nullptr && X != nullptr, "dsposv: NULL pointer");
    return dsposv_impl(ctx(h), upper != 0, prec, (int64_t)n, (int64_t)nrhs, A, (int64_t)lda, B, (int64_t)ldb, X,
                       (int64_t)ldx, tol, max_iter, iters, berr, hs(stream));
}

rmh_status rmh_dpstrf(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_piv, int64_t* d_rank,
                      rmh_stream_t stream) {
    return pstrf_entry<double>(h, upper, n, A, lda, d_piv, d_rank, stream);
}
rmh_status rmh_spstrf(rmh_handle h, int upper, size_t n, float* A, size_t lda, int64_t* d_piv, int64_t* d_rank,
                      rmh_stream_t stream) {
    return pstrf_entry<float>(h, upper, n, A, lda, d_piv, d_rank, stream);
}
rmh_status rmh_dlaswp(rmh_handle h, int inverse, size_t m, size_t ncols, double* B, size_t ldb, int row_major,
                      const int64_t* d_piv, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(ldb >= mx1(row_major ? ncols : m), "laswp: ldb too small");
    RMH_REQUIRE(m == 0 || ncols == 0 || (B != nullptr && d_piv != nullptr), "laswp: NULL pointer");
    return laswp_impl<double>(inverse != 0, (int64_t)m, (int64_t)ncols, B, (int64_t)ldb, row_major != 0, d_piv,
                              hs(stream));
}
rmh_status rmh_slaswp(rmh_handle h, int inverse, size_t m, size_t ncols, float* B, size_t ldb, int row_major,
                      const int64_t* d_piv, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(ldb >= mx1(row_major ? ncols : m), "laswp: ldb too small");
    RMH_REQUIRE(m == 0 || ncols == 0 || (B != nullptr && d_piv != nullptr), "laswp: NULL pointer");
    return laswp_impl<float>(inverse != 0, (int64_t)m, (int64_t)ncols, B, (int64_t)ldb, row_major != 0, d_piv,
                             hs(stream));
}
rmh_status rmh_dchol_update(rmh_handle h, int upper, size_t n, double alpha, double beta, double* L, size_t ldl,
                            const double* v, size_t incv, int64_t* d_info, rmh_stream_t stream) {
    return chol_update_entry<double>(h, upper, n, alpha, beta, L, ldl, v, incv, d_info, stream);
}
rmh_status rmh_schol_update(rmh_handle h, int upper, size_t n, float alpha, float beta, float* L, size_t ldl,
                            const float* v, size_t incv, int64_t* d_info, rmh_stream_t stream) {
    return chol_update_entry<float>(h, upper, n, alpha, beta, L, ldl, v, incv, d_info, stream);
}

// ----------------------------------------------------- differentiation --
rmh_status rmh_dpotrf_adjoint(rmh_handle h, int upper, size_t n, const double* L, size_t ldl, const double* Lbar,
                              size_t ldlb, double* Abar, size_t ldab, rmh_stream_t stream) {
    return adjoint_entry<double>(h, upper, n, L, ldl, Lbar, ldlb, Abar, ldab, stream);
}
rmh_status rmh_spotrf_adjoint(rmh_handle h, int upper, size_t n, const float* L, size_t ldl, const float* Lbar,
                              size_t ldlb, float* Abar, size_t ldab, rmh_stream_t stream) {
    return adjoint_entry<float>(h, upper, n, L, ldl, Lbar, ldlb, Abar, ldab, stream);
}

// ------------------------------------------------------------ batched --
rmh_status rmh_dposv_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, double* const* d_A,
                             const int64_t* d_lda, size_t nrhs, double* const* d_B, const int64_t* d_ldb,
                             int64_t* d_info, const int64_t* d_order, const rmh_batch_config* cfg,
                             rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "posv_batched: NULL pointer");
    return batched_posv_impl<double>(ctx(h), (int64_t)batch, max_n, d_n, d_A, d_lda, (int64_t)nrhs, d_B, d_ldb,
                                     d_info, d_order, cfg, true, hs(stream));
}
rmh_status rmh_sposv_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, float* const* d_A,
                             const int64_t* d_lda, size_t nrhs, float* const* d_B, const int64_t* d_ldb,
                             int64_t* d_info, const int64_t* d_order, const rmh_batch_config* cfg,
                             rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "posv_batched: NULL pointer");
    return batched_posv_impl<float>(ctx(h), (int64_t)batch, max_n, d_n, d_A, d_lda, (int64_t)nrhs, d_B, d_ldb, d_info,
                                    d_order, cfg, true, hs(stream));
}
rmh_status rmh_dpotrf_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, double* const* d_A,
                              const int64_t* d_lda, int64_t* d_info, const int64_t* d_order,
                              const rmh_batch_config* cfg, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "potrf_batched: NULL pointer");
    return batched_posv_impl<double>(ctx(h), (int64_t)batch, max_n, d_n, d_A, d_lda, 0, nullptr, nullptr, d_info,
                                     d_order, cfg, true, hs(stream));
}
rmh_status rmh_spotrf_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, float* const* d_A,
                              const int64_t* d_lda, int64_t* d_info, const int64_t* d_order,
                              const rmh_batch_config* cfg, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "potrf_batched: NULL pointer");
    return batched_posv_impl<float>(ctx(h), (int64_t)batch, max_n, d_n, d_A, d_lda, 0, nullptr, nullptr, d_info,
                                    d_order, cfg, true, hs(stream));
}
rmh_status rmh_dpotrs_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, const double* const* d_A,
                              const int64_t* d_lda, size_t nrhs, double* const* d_B, const int64_t* d_ldb,
                              const int64_t* d_order, const rmh_batch_config* cfg, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "potrs_batched: NULL pointer");
    // the solve only reads A (factor == false)
    return batched_posv_impl<double>(ctx(h), (int64_t)batch, max_n, d_n, const_cast<double* const*>(d_A), d_lda,
                                     (int64_t)nrhs, d_B, d_ldb, nullptr, d_order, cfg, false, hs(stream));
}
rmh_status rmh_spotrs_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, const float* const* d_A,
                              const int64_t* d_lda, size_t nrhs, float* const* d_B, const int64_t* d_ldb,
                              const int64_t* d_order, const rmh_batch_config* cfg, rmh_stream_t stream) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(batch == 0 || (d_n && d_A && d_lda), "potrs_batched: NULL pointer");
    return batched_posv_impl<float>(ctx(h), (int64_t)batch, max_n, d_n, const_cast<float* const*>(d_A), d_lda,
                                    (int64_t)nrhs, d_B, d_ldb, nullptr, d_order, cfg, false, hs(stream));
}
rmh_status rmh_batched_lpt_order(size_t batch, const int64_t* h_n, int64_t* h_order) {
    RMH_REQUIRE(batch == 0 || (h_n && h_order), "batched_lpt_order: NULL pointer");
    for (size_t i = 0; i < batch; ++i) h_order[i] = (int64_t)i;
    std::stable_sort(h_order, h_order + batch, [&](int64_t a, int64_t b) { return h_n[a] > h_n[b]; });
    return RMH_STATUS_SUCCESS;
}
rmh_status rmh_batched_autotune(rmh_handle h, int fp64, size_t batch, const int64_t* h_n, size_t nrhs,
                                rmh_batch_config* best, double* best_ms) {
    RMH_CHECK_HANDLE(h);
    RMH_REQUIRE(h_n != nullptr, "batched_autotune: NULL sizes");
    if (fp64) return batched_autotune_impl<double>(ctx(h), (int64_t)batch, h_n, (int64_t)nrhs, best, best_ms);
    return batched_autotune_impl<float>(ctx(h), (int64_t)batch, h_n, (int64_t)nrhs, best, best_ms);
}

rmh_status rmh_set_spotrf_precision(rmh_handle h, rmh_factor_precision prec) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    RMH_REQUIRE(prec == RMH_FACTOR_FP32 || prec == RMH_FACTOR_BF16X3, "set_spotrf_precision: unknown precision");
    ctx(h)->spotrf_bf16x3 = prec == RMH_FACTOR_BF16X3;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_set_potrf_lookahead(rmh_handle h, int enable) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    RMH_REQUIRE(enable >= -1 && enable <= 4, "set_potrf_lookahead: mode must be -1 to 4");
    ctx(h)->la_mode = enable;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_set_potrf_blocking(rmh_handle h, size_t nb, size_t nb_inner) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    if (nb_inner == 0) nb_inner = Context::kDefaultNbi;
    RMH_REQUIRE(nb_inner <= 64, "set_potrf_blocking: nb_inner must be <= 64");
    RMH_REQUIRE(nb % nb_inner == 0 && Context::kDefaultNbF64 % nb_inner == 0 &&
                    Context::kDefaultNbF32 % nb_inner == 0,
                "set_potrf_blocking: nb (and the default widths) must be multiples of nb_inner");
    ctx(h)->nb = nb;
    ctx(h)->nbi = nb_inner;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_set_potrf_reserve(rmh_handle h, int cus_per_xcd, size_t rows_lo, size_t rows_hi) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    RMH_REQUIRE(cus_per_xcd >= 0 && cus_per_xcd <= 4, "set_potrf_reserve: units per XCD must be 0 to 4");
    RMH_REQUIRE(rows_lo <= rows_hi, "set_potrf_reserve: rows_lo above rows_hi");
    ctx(h)->reserve_cus = cus_per_xcd;
    ctx(h)->rsv_rows_lo = rows_lo;
    ctx(h)->rsv_rows_hi = rows_hi;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_set_potrf_leaf(rmh_handle h, size_t leaf) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    ctx(h)->nb_leaf = leaf == 0 ? Context::kDefaultNbLeaf : leaf;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_profile_enable(rmh_handle h, int class_mask) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    ctx(h)->prof_mask = class_mask;
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_profile_reset(rmh_handle h) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    RMH_TRY(ctx(h)->activate());
    RMH_TRY_HIP(hipDeviceSynchronize());
    ctx(h)->prof_clear();
    return RMH_STATUS_SUCCESS;
}

rmh_status rmh_profile_read(rmh_handle h, const char* kernel_class, double* total_ms, double* total_flops,
                            int64_t* launches) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    RMH_REQUIRE(kernel_class != nullptr, "profile_read: class is NULL");
    int cls = -1;
    for (int i = 0; i < PROF_COUNT; ++i)
        if (std::strcmp(kernel_class, kProfNames[i]) == 0) cls = i;
    RMH_REQUIRE(cls >= 0, std::string("profile_read: unknown class ") + kernel_class);
    RMH_TRY(ctx(h)->activate());
    double ms = 0.0, fl = 0.0;
    int64_t cnt = 0;
    for (auto& r : ctx(h)->recs) {
        if (r.cls != cls) continue;
        RMH_TRY_HIP(hipEventSynchronize(r.b));
        float t = 0.f;
        RMH_TRY_HIP(hipEventElapsedTime(&t, r.a, r.b));
        ms += t;
        fl += r.flops;
        ++cnt;
    }
    if (total_ms) *total_ms = ms;
    if (total_flops) *total_flops = fl;
    if (launches) *launches = cnt;
    return RMH_STATUS_SUCCESS;
}

}  // extern "C"
