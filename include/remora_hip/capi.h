/*
 * remora_hip/capi.h -- C-ABI of the MI355X (gfx950) dense-Cholesky library.
 *
 * Replaces the layer under Remora's hip_tag kernels: the cuBLAS-backed
 * hip::blas_device (include/remora/hip/cublas_backend.hpp:86-450) and the host
 * recursion of kernels::potrf on hip_tag (include/remora/kernels/hip/potrf.hpp:83-140).
 *
 * Conventions (kept from blas_device, cublas_backend.hpp:121-431):
 *   - BLAS column-major storage; element (i,j) of a matrix with leading
 *     dimension ld is at ptr[i + j*ld].  Row-major callers flip orientation
 *     the way Remora's glue does (include/remora/kernels/hip/{gemm,syrk,trsm,
 *     trsv}.hpp).
 *   - bool-like int flags, 64-bit sizes (the reference narrowed to int for
 *     cuBLAS, cublas_backend.hpp:136-141; these do not).
 *   - every call is asynchronous on the given stream; nothing synchronises
 *     unless the function says so.  A NULL stream is the default stream.
 *   - device pointers are caller-owned; no exception crosses the ABI: each
 *     function returns an rmh_status, and rmh_last_error() gives the detail
 *     of the last failure on the calling thread.
 *   - a handle owns scratch memory and must be used by one host thread /
 *     stream at a time (like a cuBLAS handle); create one per thread.
 */
#ifndef REMORA_HIP_CAPI_H
#define REMORA_HIP_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct ihipStream_t;                      /* hipStream_t without the HIP headers */
typedef struct ihipStream_t* rmh_stream_t;
typedef struct rmh_context* rmh_handle;

typedef enum rmh_status {
    RMH_STATUS_SUCCESS = 0,
    RMH_STATUS_INVALID_VALUE = 1,          /* bad size, leading dimension, flag or NULL pointer */
    RMH_STATUS_INVALID_HANDLE = 2,
    RMH_STATUS_HIP_ERROR = 3,              /* a HIP runtime call or kernel launch failed */
    RMH_STATUS_ALLOC_FAILED = 4,           /* scratch memory could not be allocated */
    RMH_STATUS_NOT_POSITIVE_DEFINITE = 5,  /* reported by synchronous entry points only */
    RMH_STATUS_NOT_CONVERGED = 6,          /* iterative refinement reached max_iter */
    RMH_STATUS_NOT_SUPPORTED = 7
} rmh_status;

/* ---------------------------------------------------------- lifetime -- */
/* One handle per device per host thread: replaces the thread-local
 * blas_device cache of get_blas (cublas_backend.hpp:442-450). */
rmh_status rmh_create(rmh_handle* handle, int device);
rmh_status rmh_destroy(rmh_handle handle);
const char* rmh_status_string(rmh_status status);
/* Detail of the last non-success status returned on the calling thread. */
const char* rmh_last_error(void);

/* ------------------------------------------------------------- BLAS-3 -- */
/* C = alpha*op(A)*op(B) + beta*C; replaces blas_device::gemm
 * (cublas_backend.hpp:121-166).  beta == 0 does not read C. */
rmh_status rmh_dgemm(rmh_handle h, int transA, int transB, size_t m, size_t n, size_t k, double alpha,
                     const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
                     size_t ldc, rmh_stream_t stream);
rmh_status rmh_sgemm(rmh_handle h, int transA, int transB, size_t m, size_t n, size_t k, float alpha,
                     const float* A, size_t lda, const float* B, size_t ldb, float beta, float* C, size_t ldc,
                     rmh_stream_t stream);

/* C = alpha*op(A)*op(A)^T + beta*C on the upper or lower triangle only;
 * op(A) = A (n x k) if !trans, A^T (A is k x n) if trans.  Replaces
 * blas_device::syrk (cublas_backend.hpp:389-431). */
rmh_status rmh_dsyrk(rmh_handle h, int upper, int trans, size_t n, size_t k, double alpha, const double* A,
                     size_t lda, double beta, double* C, size_t ldc, rmh_stream_t stream);
rmh_status rmh_ssyrk(rmh_handle h, int upper, int trans, size_t n, size_t k, float alpha, const float* A,
                     size_t lda, float beta, float* C, size_t ldc, rmh_stream_t stream);

/* Solve op(A) X = alpha B (left) or X op(A) = alpha B (right) in place in B
 * (m x n); A is triangular of order m (left) or n (right).  Replaces
 * blas_device::trsm (cublas_backend.hpp:303-344).  Like the cuBLAS routine it
 * replaces, a zero on the diagonal is not detected (the result holds inf/nan). */
rmh_status rmh_dtrsm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, double alpha,
                     const double* A, size_t lda, double* B, size_t ldb, rmh_stream_t stream);
rmh_status rmh_strsm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, float alpha,
                     const float* A, size_t lda, float* B, size_t ldb, rmh_stream_t stream);

/* B = alpha*op(A)*B (left) or alpha*B*op(A) (right) in place, A triangular
 * (upper or lower triangle read, diagonal taken as 1 when unit) of order m
 * (left) or n (right), B m x n.  Replaces blas_device::trmm
 * (cublas_backend.hpp:222-263), called by kernels::trmm on hip_tag
 * (kernels/hip/trmm.hpp:44-80). */
rmh_status rmh_dtrmm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, double alpha,
                     const double* A, size_t lda, double* B, size_t ldb, rmh_stream_t stream);
rmh_status rmh_strmm(rmh_handle h, int left, int upper, int trans, int unit, size_t m, size_t n, float alpha,
                     const float* A, size_t lda, float* B, size_t ldb, rmh_stream_t stream);

/* ------------------------------------------------------------- BLAS-2 -- */
/* Solve op(A) x = b in place (x has stride incx).  Replaces
 * blas_device::trsv (cublas_backend.hpp:348-383). */
rmh_status rmh_dtrsv(rmh_handle h, int upper, int trans, int unit, size_t n, const double* A, size_t lda,
                     double* x, size_t incx, rmh_stream_t stream);
rmh_status rmh_strsv(rmh_handle h, int upper, int trans, int unit, size_t n, const float* A, size_t lda,
                     float* x, size_t incx, rmh_stream_t stream);

/* x = op(A) x in place (x has stride incx), A triangular of order n.
 * Replaces blas_device::trmv (cublas_backend.hpp:265-299), called by
 * kernels::trmv on hip_tag (kernels/hip/trmv.hpp). */
rmh_status rmh_dtrmv(rmh_handle h, int upper, int trans, int unit, size_t n, const double* A, size_t lda, double* x,
                     size_t incx, rmh_stream_t stream);
rmh_status rmh_strmv(rmh_handle h, int upper, int trans, int unit, size_t n, const float* A, size_t lda, float* x,
                     size_t incx, rmh_stream_t stream);

/* y = alpha*op(A)*x + beta*y, A is m x n.  Replaces blas_device::gemv
 * (cublas_backend.hpp:170-213). */
rmh_status rmh_dgemv(rmh_handle h, int trans, size_t m, size_t n, double alpha, const double* A, size_t lda,
                     const double* x, size_t incx, double beta, double* y, size_t incy, rmh_stream_t stream);
rmh_status rmh_sgemv(rmh_handle h, int trans, size_t m, size_t n, float alpha, const float* A, size_t lda,
                     const float* x, size_t incx, float beta, float* y, size_t incy, rmh_stream_t stream);

/* ------------------------------------------------------------ copies -- */
/* B = op(A): B is m x n (ld ldb), A is m x n (!trans) or n x m (trans); only
 * those elements are read and written.  The device side of kernels::assign
 * between dense hip_tag matrices (include/remora/kernels/hip/matrix_assign.hpp),
 * which assignments between row- and column-major storage need.  A and B must
 * not overlap.  Asynchronous. */
rmh_status rmh_dcopy_matrix(rmh_handle h, int trans, size_t m, size_t n, const double* A, size_t lda, double* B,
                            size_t ldb, rmh_stream_t stream);
rmh_status rmh_scopy_matrix(rmh_handle h, int trans, size_t m, size_t n, const float* A, size_t lda, float* B,
                            size_t ldb, rmh_stream_t stream);

/* ------------------------------------------------------------ LAPACK -- */
/* In-place Cholesky factorisation A = L L^T (lower) or U^T U (upper); only
 * that triangle is read and written.  Replaces the host recursion of
 * kernels::potrf on hip_tag (include/remora/kernels/hip/potrf.hpp:83-140) with
 * a device-side blocked right-looking factorisation.
 * d_info (device int64, may be NULL) receives LAPACK's info: 0 on success,
 * k+1 if the leading minor of order k+1 is not positive definite (the
 * factorisation stops there).  Asynchronous: read d_info after the stream. */
rmh_status rmh_dpotrf(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_info,
                      rmh_stream_t stream);
rmh_status rmh_spotrf(rmh_handle h, int upper, size_t n, float* A, size_t lda, int64_t* d_info,
                      rmh_stream_t stream);

/* Solve A X = B with the factor from rmh_?potrf (two triangular solves, as
 * cholesky_decomposition::solve does, include/remora/decompositions.hpp:101-105). */
rmh_status rmh_dpotrs(rmh_handle h, int upper, size_t n, size_t nrhs, const double* A, size_t lda, double* B,
                      size_t ldb, rmh_stream_t stream);
rmh_status rmh_spotrs(rmh_handle h, int upper, size_t n, size_t nrhs, const float* A, size_t lda, float* B,
                      size_t ldb, rmh_stream_t stream);

/* Pivoted Cholesky factorisation of a symmetric positive semi-definite
 * matrix in place: P^T A P = L L^T (lower) or U^T U (upper), the reference's
 * algorithm and pivot order (kernels::pstrf, include/remora/kernels/pstrf.hpp:58-66,
 * include/remora/kernels/default/pstrf.hpp:40-160).  d_piv (device, n int64)
 * receives the transpositions: at step j row/column j was swapped with
 * d_piv[j] (j itself when not swapped; 0-based), as the reference's
 * permutation_matrix records them.  d_rank (device int64) receives the rank:
 * the factorisation stops when the largest remaining pivot is below
 * n^2 * epsilon * max|A_ii|, and the trailing square is then zero.  Only
 * the `upper`/lower triangle is read (A symmetric); the other triangle is
 * zeroed on return, as the reference leaves it.  Asynchronous. */
rmh_status rmh_dpstrf(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_piv, int64_t* d_rank,
                      rmh_stream_t stream);
rmh_status rmh_spstrf(rmh_handle h, int upper, size_t n, float* A, size_t lda, int64_t* d_piv, int64_t* d_rank,
                      rmh_stream_t stream);

/* Row interchanges of rmh_?pstrf's d_piv on an m x ncols matrix B (column-
 * major with leading dimension ldb, or row-major when row_major): rows j and
 * d_piv[j] are swapped for j = 0 .. m-1, or m-1 .. 0 when `inverse`.  These
 * are the reference's swap_rows / swap_rows_inverted with a permutation_matrix
 * (used by symm_pos_semi_definite_solver::solve,
 * include/remora/decompositions.hpp:360-375).  Asynchronous. */
rmh_status rmh_dlaswp(rmh_handle h, int inverse, size_t m, size_t ncols, double* B, size_t ldb, int row_major,
                      const int64_t* d_piv, rmh_stream_t stream);
rmh_status rmh_slaswp(rmh_handle h, int inverse, size_t m, size_t ncols, float* B, size_t ldb, int row_major,
                      const int64_t* d_piv, rmh_stream_t stream);

/* Rank-one update of a Cholesky factor in place: with A = L L^T on entry
 * (the factor from rmh_?potrf: L lower, or U = L^T upper; only that triangle
 * read and written), on return L L^T = alpha*A + beta*v*v^T.  Replaces
 * cholesky_decomposition::update (include/remora/decompositions.hpp:132-169),
 * same arguments (alpha >= 0; beta of either sign).  v: n elements at stride
 * incv, unchanged.  beta == 0 scales the whole n x n stored matrix by
 * sqrt(alpha), as the reference does.  d_info (device int64, not NULL)
 * receives 0, or j+1 when the update would make the matrix indefinite at
 * column j (the reference's std::invalid_argument); L is then unchanged.
 * Asynchronous: read d_info after the stream. */
rmh_status rmh_dchol_update(rmh_handle h, int upper, size_t n, double alpha, double beta, double* L, size_t ldl,
                            const double* v, size_t incv, int64_t* d_info, rmh_stream_t stream);
rmh_status rmh_schol_update(rmh_handle h, int upper, size_t n, float alpha, float beta, float* L, size_t ldl,
                            const float* v, size_t incv, int64_t* d_info, rmh_stream_t stream);

/* rmh_dpotrf without any allocation: the same factorisation, arguments and
 * result (bit for bit), on a static per-device scratch array of the library
 * instead of the handle's workspace, on the caller's stream alone (no
 * look-ahead stream, no profile events).  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Limit:
 * n <= 32768 with the default nb_inner 32 (n <= 16384 at nb_inner 64; the
 * scratch is 8 MiB + 256 bytes per GPU, held until the process ends); larger
 * orders return RMH_STATUS_INVALID_VALUE.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU race and must be ordered by
 * the caller.  Asynchronous, like rmh_dpotrf. */
rmh_status rmh_dpotrf_noalloc(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_info,
                              rmh_stream_t stream);

/* ----------------------------------------------------- mixed precision -- */
/* Precision of the factorisation inside rmh_dsposv. */
typedef enum rmh_factor_precision {
    RMH_FACTOR_FP32 = 0,   /* trailing updates on FP32 MFMA, FP32 accumulation */
    RMH_FACTOR_BF16X3 = 1  /* FP32 trailing updates with each operand split into two BF16 terms,
                              three BF16 MFMA products (hi*hi + hi*lo + lo*hi), FP32 accumulation
                              (lower storage; upper storage uses FP32 MFMA) */
} rmh_factor_precision;

/* Solve A X = B (A SPD, only the `upper`/lower triangle read, A unchanged)
 * by factoring a reduced-precision copy of A -- diagonal blocks in FP64,
 * trailing updates in `prec` -- and refining X with FP64 residuals until the
 * backward error  max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
 * is <= tol, at most max_iter corrections.  Like LAPACK's dsposv it falls
 * back to the FP64 factorisation when the reduced-precision factor is not
 * positive definite (*iters = -3) or refinement does not converge
 * (*iters = -(max_iter+1)); otherwise *iters is the number of corrections.
 * *berr receives the final backward error.  Synchronous: returns when X is
 * ready.  No counterpart in the reference (a new form beside the FP64 path). */
rmh_status rmh_dsposv(rmh_handle h, int upper, rmh_factor_precision prec, size_t n, size_t nrhs, const double* A,
                      size_t lda, const double* B, size_t ldb, double* X, size_t ldx, double tol, int max_iter,
                      int* iters, double* berr, rmh_stream_t stream);

/* ------------------------------------------------------ differentiation -- */
/* Reverse-mode derivative of the Cholesky factorisation: given the factor
 * from rmh_?potrf (L lower, or U = L^T upper; only that triangle read) and
 * the gradient Lbar of a scalar with respect to that triangle, writes the
 * symmetric gradient with respect to A to Abar (n x n, both triangles):
 *     Abar = sym(L^-T Phi(L^T Lbar) L^-1),  Phi = lower triangle, diagonal halved,
 * sym(G) = (G + G^T)/2.  The inverses of L's diagonal blocks are recomputed
 * inside the two triangular solves, not stored.  Abar must not alias L or
 * Lbar.  Asynchronous.  No counterpart in the reference (a new form beside the
 * FP64 drop-in). */
rmh_status rmh_dpotrf_adjoint(rmh_handle h, int upper, size_t n, const double* L, size_t ldl, const double* Lbar,
                              size_t ldlb, double* Abar, size_t ldab, rmh_stream_t stream);
rmh_status rmh_spotrf_adjoint(rmh_handle h, int upper, size_t n, const float* L, size_t ldl, const float* Lbar,
                              size_t ldlb, float* Abar, size_t ldab, rmh_stream_t stream);

/* ----------------------------------------------------------- batched -- */
/* Launch configuration of the batched solver: nb is the block column of the
 * per-matrix factorisation (16 or 32), wg_per_cu the workgroups per compute
 * unit of the persistent grid (0: as many as fit), lpt_order whether the
 * autotuner found a largest-first order (rmh_batched_lpt_order) faster. */
typedef struct rmh_batch_config {
    int nb;
    int wg_per_cu;
    int lpt_order;
} rmh_batch_config;

/* Factor and solve `batch` independent SPD systems A_i X_i = B_i in one
 * persistent launch: A_i is n_i x n_i (0 <= n_i <= max_n <= 512, column-major,
 * ld lda_i, lower triangle read and overwritten by L_i), B_i is n_i x nrhs
 * (nrhs <= 16, ld ldb_i, overwritten by X_i).  Every array argument is in
 * device memory -- d_n, d_lda, d_ldb, d_info (int64 per system), d_A and d_B
 * (device pointers per system), d_order (optional processing order, a
 * permutation of 0..batch-1, NULL for 0..batch-1) -- so the launch can be
 * captured in a graph and replayed with new data.  d_info[i] receives 0, k+1
 * for a non-positive pivot k (X_i is then not computed, and B_i holds the
 * forward substitution of the columns before block column k), or -1 for n_i
 * out of range (A_i and B_i untouched).  cfg NULL uses the handle's tuned configuration.  Asynchronous;
 * allocates nothing after the handle's first batched call.  No counterpart
 * in the reference (a new form beside the FP64 drop-in). */
rmh_status rmh_dposv_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, double* const* d_A,
                             const int64_t* d_lda, size_t nrhs, double* const* d_B, const int64_t* d_ldb,
                             int64_t* d_info, const int64_t* d_order, const rmh_batch_config* cfg,
                             rmh_stream_t stream);
rmh_status rmh_sposv_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, float* const* d_A,
                             const int64_t* d_lda, size_t nrhs, float* const* d_B, const int64_t* d_ldb,
                             int64_t* d_info, const int64_t* d_order, const rmh_batch_config* cfg,
                             rmh_stream_t stream);
/* Factor only (rmh_?potrf for every system), same arguments. */
rmh_status rmh_dpotrf_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, double* const* d_A,
                              const int64_t* d_lda, int64_t* d_info, const int64_t* d_order,
                              const rmh_batch_config* cfg, rmh_stream_t stream);
rmh_status rmh_spotrf_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, float* const* d_A,
                              const int64_t* d_lda, int64_t* d_info, const int64_t* d_order,
                              const rmh_batch_config* cfg, rmh_stream_t stream);
/* Solve only, with factors from rmh_?potrf_batched (rmh_?potrs for every system). */
rmh_status rmh_dpotrs_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, const double* const* d_A,
                              const int64_t* d_lda, size_t nrhs, double* const* d_B, const int64_t* d_ldb,
                              const int64_t* d_order, const rmh_batch_config* cfg, rmh_stream_t stream);
rmh_status rmh_spotrs_batched(rmh_handle h, size_t batch, int max_n, const int64_t* d_n, const float* const* d_A,
                              const int64_t* d_lda, size_t nrhs, float* const* d_B, const int64_t* d_ldb,
                              const int64_t* d_order, const rmh_batch_config* cfg, rmh_stream_t stream);
/* Largest-first processing order of host sizes h_n into host h_order
 * (stable for equal sizes). */
rmh_status rmh_batched_lpt_order(size_t batch, const int64_t* h_n, int64_t* h_order);
/* Autotuner: times every candidate configuration (nb in {16, 32}, each
 * power-of-two wg_per_cu up to occupancy, with and without largest-first
 * order) of rmh_?posv_batched on synthetic SPD systems of the host sizes
 * h_n with nrhs right-hand sides, stores the fastest in the handle (used
 * when cfg is NULL) and returns it with its time.  fp64 selects the
 * precision.  Synchronous; allocates and frees its own test data. */
rmh_status rmh_batched_autotune(rmh_handle h, int fp64, size_t batch, const int64_t* h_n, size_t nrhs,
                                rmh_batch_config* best, double* best_ms);

/* ------------------------------------------------------------ tuning -- */
/* Block sizes of rmh_?potrf: nb is the width of a panel (the K of the
 * trailing update), nb_inner the diagonal block inside a panel (<= 64).
 * 0 restores the defaults (nb 1024 for FP64 and FP32, 768 for split-BF16x3
 * FP32 factors; nb_inner 32).  Applies to later calls on this handle. */
rmh_status rmh_set_potrf_blocking(rmh_handle h, size_t nb, size_t nb_inner);
/* Leaves of the recursive panel factorisation of rmh_?potrf: a panel is
 * halved recursively (factor the left half, update the right half with it on
 * the MFMA GEMM, factor the right half) down to leaves of at most `leaf`
 * columns, each factored by a chain of fused block-column launches.  Rounded
 * down to a multiple of max(nb_inner, 32); 0 restores the default 128; a
 * leaf at least as wide as the panel factors every panel as one chain.  No
 * counterpart in the reference (a tuning knob of this library).  Applies to
 * later calls on this handle. */
rmh_status rmh_set_potrf_leaf(rmh_handle h, size_t leaf);
/* Precision of the trailing updates of rmh_spotrf on this handle:
 * RMH_FACTOR_FP32 (default, FP32 MFMA) or RMH_FACTOR_BF16X3 (split-BF16x3,
 * lower storage).  rmh_dsposv sets its own choice for its factorisation. */
rmh_status rmh_set_spotrf_precision(rmh_handle h, rmh_factor_precision prec);
/* Look-ahead mode 1: rmh_?potrf factors panel k+1 on a second,
 * handle-owned stream while the caller's stream applies the rest of panel
 * k's trailing update; the call still completes in the caller's stream
 * order.  Mode 2 also limits the trailing update to one workgroup per
 * compute unit so that the panel kernel's workgroups run beside it; mode 0
 * is a single stream.  Modes 3 and 4 are modes 2 and 1 with the panel
 * kernel on four wavefronts per workgroup instead of eight (small enough to
 * share a compute unit with one trailing-update workgroup); their factors
 * are identical to each other and differ from modes 0-2 in rounding only.
 * -1 (the default) picks by precision: mode 1 for rmh_dpotrf, mode 4 for
 * rmh_spotrf. */
rmh_status rmh_set_potrf_lookahead(rmh_handle h, int enable);
/* Look-ahead rmh_?potrf: number of compute units of each XCD (0 to 4) that
 * the trailing updates leave free, so that the second stream's panel
 * launches start without waiting for trailing-update workgroups to retire,
 * for the panel steps whose trailing rows (n minus the next panel's first
 * column) lie in (rows_lo, rows_hi].  Those steps run their updates on a
 * handle-owned stream created with that CU mask
 * (hipExtStreamCreateWithCUMask); the call still completes in the caller's
 * stream order, and the factors are identical for every setting.  0 (the
 * default) reserves nothing.  Measured at n = 16384: 4 units over
 * (4096, 12288] shorten rmh_dpotrf alone by about 2% and (4096, 16384] the
 * FP32 factorisation of rmh_dsposv by about 1%, while the masked updates run
 * on 224 of the 256 units.  No counterpart in the reference (a tuning knob of
 * this library). */
rmh_status rmh_set_potrf_reserve(rmh_handle h, int cus_per_xcd, size_t rows_lo, size_t rows_hi);

/* ---------------------------------------------------------- profiling -- */
/* Kernel classes of rmh_?potrf: bit 0 "trailing_update" (trailing updates on
 * 128 x 128 tiles), bit 1 "panel" (the fused panel kernel), bit 2
 * "trailing_update_small" (trailing updates on 64 x 64 tiles).  While
 * enabled, the handle records HIP events around every launch of a class in
 * class_mask on the stream it is launched on and accumulates the algorithmic
 * flops of each launch (class_mask 0 disables, -1 enables all).
 * rmh_profile_read waits for the recorded events and returns totals since
 * the last reset. */
rmh_status rmh_profile_enable(rmh_handle h, int class_mask);
rmh_status rmh_profile_reset(rmh_handle h);
rmh_status rmh_profile_read(rmh_handle h, const char* kernel_class, double* total_ms, double* total_flops,
                            int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* REMORA_HIP_CAPI_H */
