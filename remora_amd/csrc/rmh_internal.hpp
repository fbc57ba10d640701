// Internal declarations shared by the HIP translation units of libremora_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "remora_hip/capi.h"

namespace rmh {

// --------------------------------------------------------------- errors --
void set_error(const std::string& msg);
rmh_status fail(rmh_status st, const std::string& msg);
rmh_status check_launch(const char* what);

#define RMH_TRY_HIP(call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::rmh::fail(RMH_STATUS_HIP_ERROR,                                        \
                               std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define RMH_TRY(expr)                                                                       \
    do {                                                                                    \
        rmh_status s_ = (expr);                                                             \
        if (s_ != RMH_STATUS_SUCCESS) return s_;                                            \
    } while (0)

inline hipStream_t hs(rmh_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------ GEMM / SYRK core --
// Tile shapes of the MFMA GEMM family.  All use 256 threads (4 waves).
enum class Tile { T128x128, T64x64, T128x64, T64x128 };

// Which part of C a launch writes.
//   Full: the whole M x N rectangle.
//   Lower: the lower trapezoid row >= col of an M x N block (M >= N).
//   Upper: the upper trapezoid row <= col of an M x N block (N >= M).
enum class Part { Full, Lower, Upper };

template <typename T>
struct GemmDesc {
    int64_t M = 0, N = 0, K = 0;
    T alpha = T(1), beta = T(0);
    const T* A = nullptr;
    int64_t lda = 1;
    const T* B = nullptr;
    int64_t ldb = 1;
    T* C = nullptr;
    int64_t ldc = 1;
    bool transA = false, transB = false;
    Part part = Part::Full;
    const int64_t* abort_flag = nullptr;  // skip the launch's work when *abort_flag != 0
    // strided batch of `batch` products: product b uses y = b % batch_ny,
    // z = b / batch_ny and operand origins A + y*sAy + z*sAz (same for B, C)
    int64_t batch = 1, batch_ny = 1;
    int64_t sAy = 0, sAz = 0, sBy = 0, sBz = 0, sCy = 0, sCz = 0;
    // split-K over `splits` chunks of K; partials go to split_ws
    // (splits * M * N elements), then one reduction writes C
    int splits = 1;
    T* split_ws = nullptr;
    int64_t k_chunk = 0;  // set by launch_gemm
    // op(A) is upper triangular (op(A)(i, k) = 0 for k < i): each tile skips
    // the depth below its first row (no split-K, no batch)
    bool k_from_row = false;
    // reserve LDS so that one workgroup of this launch occupies a compute
    // unit's share for a 128 x 128 tile and leaves room for one workgroup of
    // a concurrent kernel (the look-ahead panel kernel)
    bool one_per_cu = false;
    // a triangular part on 128 x 128 tiles whose last round of workgroups
    // would fill at most half the device runs that round as 64 x 64 quarter
    // tiles in a second launch (bitwise the same result)
    bool tail_split = false;
};

// Launch C = alpha*op(A)*op(B) + beta*C (restricted to `part`).  `tile`
// selects the tile shape; Tile::T128x128 is the default for large problems.
template <typename T>
rmh_status launch_gemm(const GemmDesc<T>& d, Tile tile, hipStream_t stream);

// C = alpha*A*op(B) + beta*C with A symmetric (M x M), only its lower
// (upper = false) or upper triangle read.
template <typename T>
rmh_status launch_symm(const GemmDesc<T>& d, bool upper, hipStream_t stream);

// C = alpha*A*B^T + beta*C on a lower trapezoid (A M x K, B N x K, FP32)
// with each operand split into two BF16 terms and three BF16 MFMAs per
// product (hi*hi + hi*lo + lo*hi), FP32 accumulation; 128 x 128 tiles.
rmh_status launch_gemm_bf16x3(const GemmDesc<float>& d, hipStream_t stream);

// The same product with the operands already split (launch_potrf_panel's
// split output): d.A / d.B are not read; A's row r (0 <= r < M) is row
// a_row0 + r of the split panel, B's row r is row b_row0 + r, and depth k of
// row r is at (k / 32) * kstride + r * 32 + k % 32 of hi and lo.
struct SplitPanel {
    const __bf16* hi = nullptr;
    const __bf16* lo = nullptr;
    int64_t kstride = 0;
    int64_t a_row0 = 0, b_row0 = 0;
};
rmh_status launch_gemm_bf16x3_split(const GemmDesc<float>& d, const SplitPanel& sp, hipStream_t stream);

// Pick a tile shape for a problem of this size (fills 256 CUs).
Tile pick_tile(int64_t M, int64_t N, Part part);

// ------------------------------------------------- small LAPACK kernels --
// One block column of a left-looking Cholesky panel step (kb <= 64):
// V(i,c) = P[i*rs + c*cs], rows [0, rows) from the diagonal down; columns
// [0, K) are the panel's factored columns, [K, K+kb) the block column.
// Applies the earlier columns' updates, factors the diagonal block into
// dfac (element (i, c) at dfac[i + c*ldf]; the block itself keeps its
// values) and solves the rows below in place.  On a non-positive pivot j:
// *info = info_base + j + 1 (the first failure is kept).  Computes in FP64
// whatever T is.
// With sph/spl (FP32 factors, kb <= 32) the solved rows are also written
// split into BF16 hi + lo for the split-BF16x3 trailing update: view element
// (i, k) at (k / 32) * sp_kstride + i * 32 + k % 32 of sph and spl.  With
// `narrow`, blocks up to 32 wide run four wavefronts per workgroup instead of
// eight (same results), small enough to sit beside a trailing-update
// workgroup on one CU.
template <typename T>
rmh_status launch_potrf_panel(T* P, int64_t rs, int64_t cs, int64_t rows, int64_t K, int kb, int64_t* info,
                              int64_t info_base, T* dfac, int64_t ldf, hipStream_t stream, __bf16* sph = nullptr,
                              __bf16* spl = nullptr, int64_t sp_kstride = 0, bool narrow = false);

// Copy the diagonal-block factors written by launch_potrf_panel (blocks of
// nbi rows; block j0's element (i, c) at dfac[j0 + i + c*ldf]) into V of
// the n x n matrix; blocks at or after a failed pivot (*info != 0) are left.
template <typename T>
rmh_status launch_scatter_diag(T* A, int64_t rs, int64_t cs, int64_t n, int nbi, const T* dfac, int64_t ldf,
                               const int64_t* info, hipStream_t stream);

// Invert every nb x nb diagonal block of the triangular matrix op(A)
// (n x n).  eff_upper says whether op(A) is upper triangular.  Block b is
// written column-major to Dinv + b*nb*nb.
template <typename T>
rmh_status launch_trtri_blocks(const T* A, int64_t lda, int64_t n, int nb, bool upper, bool trans, bool unit,
                               T* Dinv, hipStream_t stream);

// Inverses of the 64-blocks of the lower-triangular S = (s_is_At ? A^T : A)
// into a block-diagonal strip D (n x nbd, ld ldd): element (i, j) of the
// inverse of an nbd-block lives at D[i + (j mod nbd)*ldd].
template <typename T>
rmh_status launch_trtri_strip(const T* A, int64_t lda, int64_t n, bool s_is_At, bool unit, T* D, int64_t ldd,
                              int64_t nbd, hipStream_t stream);

template <typename T>
rmh_status launch_scale(T* B, int64_t ldb, int64_t rows, int64_t cols, T alpha, hipStream_t stream);

template <typename T>
rmh_status launch_gemv(bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* x,
                       int64_t incx, T beta, T* y, int64_t incy, hipStream_t stream);

template <typename T>
rmh_status launch_strided_copy(const T* src, int64_t sinc, T* dst, int64_t dinc, int64_t n, hipStream_t stream);

rmh_status launch_fill_i64(int64_t* p, int64_t n, int64_t value, hipStream_t stream);

}  // namespace rmh
