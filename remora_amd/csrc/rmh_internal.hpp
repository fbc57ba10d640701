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
};

// Launch C = alpha*op(A)*op(B) + beta*C (restricted to `part`).  `tile`
// selects the tile shape; Tile::T128x128 is the default for large problems.
template <typename T>
rmh_status launch_gemm(const GemmDesc<T>& d, Tile tile, hipStream_t stream);

// C = alpha*A*op(B) + beta*C with A symmetric (M x M), only its lower
// (upper = false) or upper triangle read.
template <typename T>
rmh_status launch_symm(const GemmDesc<T>& d, bool upper, hipStream_t stream);

// Pick a tile shape for a problem of this size (fills 256 CUs).
Tile pick_tile(int64_t M, int64_t N, Part part);

// ------------------------------------------------- small LAPACK kernels --
// Factor the kb x kb (kb <= 64) diagonal block whose (i,j) element is at
// A[i*rs + j*cs] (lower triangle, i >= j), write L back, and write its
// inverse to W (column-major, ldw, full kb x kb with zeros above the
// diagonal).  On a non-positive pivot j: *info = info_base + j + 1.
// Computes in FP64 whatever T is.
template <typename T>
rmh_status launch_potrf_diag(T* A, int64_t rs, int64_t cs, int kb, T* W, int64_t ldw, int64_t* info,
                             int64_t info_base, hipStream_t stream);

// Invert every nb x nb diagonal block of the triangular matrix op(A)
// (n x n).  eff_upper says whether op(A) is upper triangular.  Block b is
// written column-major to Dinv + b*nb*nb.
template <typename T>
rmh_status launch_trtri_blocks(const T* A, int64_t lda, int64_t n, int nb, bool upper, bool trans, bool unit,
                               T* Dinv, hipStream_t stream);

// In-place B_blk <- Dinv_b * B_blk (left) or B_blk <- B_blk * Dinv_b
// (right); for left, B_blk is rows [r0, r0+kb) of the m x nrhs block of B,
// for right, B_blk is columns [r0, r0+kb).  `other` is the extent of the
// non-blocked dimension.
template <typename T>
rmh_status launch_tri_apply(bool left, const T* Dinv, int64_t ldd, int kb, T* B, int64_t ldb, int64_t r0,
                            int64_t other, hipStream_t stream);

template <typename T>
rmh_status launch_scale(T* B, int64_t ldb, int64_t rows, int64_t cols, T alpha, hipStream_t stream);

template <typename T>
rmh_status launch_gemv(bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* x,
                       int64_t incx, T beta, T* y, int64_t incy, hipStream_t stream);

template <typename T>
rmh_status launch_strided_copy(const T* src, int64_t sinc, T* dst, int64_t dinc, int64_t n, hipStream_t stream);

rmh_status launch_fill_i64(int64_t* p, int64_t n, int64_t value, hipStream_t stream);

}  // namespace rmh
