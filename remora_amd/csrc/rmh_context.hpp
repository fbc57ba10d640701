// The object behind an rmh_handle: device, scratch memory, tuning and the
// optional launch profiler.
#pragma once

#include <vector>

#include "rmh_internal.hpp"

namespace rmh {

enum ProfClass { PROF_TRAILING = 0, PROF_PANEL = 1, PROF_TRAILING_SMALL = 2, PROF_COUNT = 3 };
static const char* const kProfNames[PROF_COUNT] = {"trailing_update", "panel", "trailing_update_small"};

struct ProfRec {
    hipEvent_t a = nullptr, b = nullptr;
    double flops = 0.0;
    int cls = 0;
};

// Scratch slots of a handle: each grows on demand and is reused by every
// call that uses it.  Routines that call other routines use different slots.
enum WsSlot { WS_MAIN = 0, WS_MIXED = 1, WS_FACTOR = 2, WS_ADJOINT = 3, WS_UPDATE = 4, WS_PSTRF = 5, WS_SPLIT = 6, WS_COUNT = 7 };

struct Context {
    // default panel width by precision (the K of the trailing updates; the
    // panels are factored recursively, so a wide panel costs little on the
    // panel chain): 1024 for FP64 and FP32, 768 for split-BF16x3 factors,
    // whose trailing updates are cheap enough that a panel's GEMMs and
    // chains balance there (measured at n = 16384)
    static constexpr size_t kDefaultNbF64 = 1024, kDefaultNbF32 = 1024, kDefaultNbBf16x3 = 768;
    static constexpr size_t kDefaultNbi = 32;
    static constexpr size_t kDefaultNbLeaf = 128;
    int device = 0;
    void* ws = nullptr;  // WS_MAIN (potrf / trsm scratch)
    size_t ws_bytes = 0;
    void* aux[WS_COUNT] = {};
    size_t aux_bytes[WS_COUNT] = {};
    size_t nb = 0, nbi = kDefaultNbi;  // nb 0: the default for the precision
    size_t nb_leaf = kDefaultNbLeaf;     // leaf width of the recursive panel factorisation
    int prof_mask = 0;  // bit i: record launches of ProfClass i
    // look-ahead mode (rmh_set_potrf_lookahead); -1: by precision, mode 1
    // for FP64 factorisations and mode 4 (four-wavefront panel workgroups)
    // for FP32 ones (measured fastest at n = 16384)
    int la_mode = -1;
    // FP32 factorisations run their 128 x 128 lower trailing updates on the
    // split-BF16x3 kernel (set by rmh_dsposv for RMH_FACTOR_BF16X3)
    bool bf16x3_trailing = false;
    bool spotrf_bf16x3 = false;  // rmh_set_spotrf_precision
    hipStream_t side = nullptr;  // look-ahead stream of potrf (default priority)
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // set by rmh_dsposv around its factorisation: the columns right of the
    // first panel are ready (converted on the side stream) once this event
    // has completed; potrf waits for it after factoring the first panel
    hipEvent_t cols_ready = nullptr;
    hipEvent_t ev_conv = nullptr, ev_conv0 = nullptr;  // rmh_dsposv's own events (made on first use)
    // look-ahead potrf: compute units per XCD kept free of the trailing
    // updates (rmh_set_potrf_reserve), so that the panel launches of the side
    // stream find room at once; the updates then run on `trail`, a stream
    // whose CU mask leaves those units out (made for `trail_rsv` units)
    int reserve_cus = 0;                       // 0: none (the default)
    size_t rsv_rows_lo = 0, rsv_rows_hi = 0;  // the steps whose rows n - k1 lie in (lo, hi] reserve
    hipStream_t trail = nullptr;
    int trail_rsv = 0;
    hipEvent_t ev_c = nullptr, ev_d = nullptr;
    std::vector<ProfRec> recs;
    // batched solver: tuned launch configuration ([0] FP32, [1] FP64) and the
    // work counters of its persistent kernel (zeroed once, kept zero by it)
    rmh_batch_config batch_cfg[2] = {{32, 0, 1}, {16, 0, 1}};
    void* batch_ctr = nullptr;

    rmh_status activate();
    rmh_status reserve(size_t bytes);
    rmh_status reserve_slot(int slot, size_t bytes, void** ptr);
    rmh_status side_stream(hipStream_t* st, hipEvent_t* ready, hipEvent_t* done);
    rmh_status trail_stream(int rsv, hipStream_t* st, hipEvent_t* in, hipEvent_t* out);
    rmh_status prof_begin(hipStream_t s, int cls, double flops, int* slot);
    rmh_status prof_end(hipStream_t s, int slot);
    void prof_clear();
    rmh_status batch_counters(unsigned** c);
    ~Context();
};

// panel width of a factorisation: the handle's setting, else the default of
// its precision (FP64, FP32, or FP32 with split-BF16x3 trailing updates)
inline int64_t potrf_panel_width(const Context* h, bool fp64, bool bf16x3) {
    if (h->nb != 0) return (int64_t)h->nb;
    return (int64_t)(fp64 ? Context::kDefaultNbF64 : (bf16x3 ? Context::kDefaultNbBf16x3 : Context::kDefaultNbF32));
}

template <typename T>
rmh_status trsm_impl(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                     const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s);
bool trsm_blocked_applies(int64_t na, int64_t other);
// Inverses of the diagonal blocks of a triangular matrix S (lower storage, or
// upper storage used as S = A^T), built once and shared by every blocked
// solve with S or S^T (both solves of potrs, every correction of dsposv).
template <typename T>
struct TrsmInverses {
    const T* A = nullptr;
    int64_t lda = 0, na = 0, nbd = 0;
    bool upper = false, unit = false;
    T* D = nullptr;  // na x nbd strip of block inverses (ld na)
    T* P = nullptr;  // split-K partials
};
// D (na x nbd strip, ld na) <- inverses of the nbd x nbd diagonal blocks of
// S = A (lower storage) or A^T (upper storage); nbd = 2^k >= 64; W needs
// na * nbd / 2 + 64 * 64 elements.
template <typename T>
rmh_status diag_block_inverses(const T* A, int64_t lda, int64_t na, bool upper, bool unit, int64_t nbd, T* D, T* W,
                               hipStream_t s);
// bytes of workspace trsm_prepare needs for order na and `other` right-hand sides
template <typename T>
size_t trsm_inverse_workspace(int64_t na, int64_t other);
template <typename T>
rmh_status trsm_prepare(const T* A, int64_t lda, int64_t na, bool upper, bool unit, void* ws, TrsmInverses<T>* inv,
                        hipStream_t s);
// op(S) X = alpha B (left) or X op(S) = alpha B (right), op = transpose if trans
template <typename T>
rmh_status trsm_apply(const TrsmInverses<T>& inv, bool left, bool trans, int64_t m, int64_t n, T alpha, T* B,
                      int64_t ldb, hipStream_t s);
template <typename T>
rmh_status trsm_blocked(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                        const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s);
template <typename T>
rmh_status copy_matrix_impl(bool trans, int64_t m, int64_t n, const T* A, int64_t lda, T* B, int64_t ldb,
                            hipStream_t s);
template <typename T>
rmh_status trmv_impl(Context* h, bool upper, bool trans, bool unit, int64_t n, const T* A, int64_t lda, T* x,
                     int64_t incx, hipStream_t s);
template <typename T>
rmh_status trmm_impl(Context* h, bool left, bool upper, bool trans, bool unit, int64_t m, int64_t n, T alpha,
                     const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s);
template <typename T>
size_t potrf_workspace(const Context* h, int64_t n);
template <typename T>
rmh_status potrf_core(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* d_info, void* ws, bool plain,
                      hipStream_t s);
// The static device scratch of the entry points that allocate nothing:
// noalloc_scratch_init resolves its address on `device` (which must be
// current) once; noalloc_scratch returns it (nullptr before the init).
rmh_status noalloc_scratch_init(int device);
void* noalloc_scratch(int device, size_t* bytes);
template <typename T>
rmh_status potrf_impl(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* d_info, hipStream_t s);
template <typename T>
rmh_status potrs_impl(Context* h, bool upper, int64_t n, int64_t nrhs, const T* A, int64_t lda, T* B, int64_t ldb,
                      hipStream_t s);

rmh_status dsposv_impl(Context* h, bool upper, rmh_factor_precision prec, int64_t n, int64_t nrhs, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double* X, int64_t ldx, double tol, int max_iter,
                       int* iters, double* berr, hipStream_t s);

template <typename T>
rmh_status batched_posv_impl(Context* h, int64_t batch, int max_n, const int64_t* n, T* const* A, const int64_t* lda,
                             int64_t nrhs, T* const* B, const int64_t* ldb, int64_t* info, const int64_t* order,
                             const rmh_batch_config* cfg, bool factor, hipStream_t s);
template <typename T>
rmh_status batched_autotune_impl(Context* h, int64_t batch, const int64_t* h_n, int64_t nrhs, rmh_batch_config* best,
                                 double* best_ms);

template <typename T>
rmh_status potrf_adjoint_impl(Context* h, bool upper, int64_t n, const T* L, int64_t ldl, const T* Lbar, int64_t ldlb,
                              T* Abar, int64_t ldab, hipStream_t s);

template <typename T>
rmh_status laswp_impl(bool inverse, int64_t m, int64_t ncols, T* B, int64_t ldb, bool row_major, const int64_t* piv,
                      hipStream_t s);
template <typename T>
rmh_status pstrf_impl(Context* h, bool upper, int64_t n, T* A, int64_t lda, int64_t* piv, int64_t* rank, hipStream_t s);

template <typename T>
rmh_status chol_update_impl(Context* h, bool upper, int64_t n, T alpha, T beta, T* L, int64_t ldl, const T* v,
                            int64_t incv, int64_t* info, hipStream_t s);

}  // namespace rmh
