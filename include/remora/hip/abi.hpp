// remora/hip/abi.hpp -- precision dispatch (float / double overloads) onto
// the rmh_* C-ABI of libremora_hip (include/remora_hip/capi.h).
#ifndef REMORA_HIP_ABI_HPP
#define REMORA_HIP_ABI_HPP

#include <cstddef>
#include <cstdint>

#include "../../remora_hip/capi.h"

namespace remora {
namespace hip {
namespace abi {
inline rmh_status gemm(rmh_handle h, int ta, int tb, size_t m, size_t n, size_t k, double al, const double* A, size_t lda,
                       const double* B, size_t ldb, double be, double* C, size_t ldc, rmh_stream_t s) {
    return rmh_dgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
}
inline rmh_status gemm(rmh_handle h, int ta, int tb, size_t m, size_t n, size_t k, float al, const float* A, size_t lda,
                       const float* B, size_t ldb, float be, float* C, size_t ldc, rmh_stream_t s) {
    return rmh_sgemm(h, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
}
inline rmh_status syrk(rmh_handle h, int up, int tr, size_t n, size_t k, double al, const double* A, size_t lda,
                       double be, double* C, size_t ldc, rmh_stream_t s) {
    return rmh_dsyrk(h, up, tr, n, k, al, A, lda, be, C, ldc, s);
}
inline rmh_status syrk(rmh_handle h, int up, int tr, size_t n, size_t k, float al, const float* A, size_t lda,
                       float be, float* C, size_t ldc, rmh_stream_t s) {
    return rmh_ssyrk(h, up, tr, n, k, al, A, lda, be, C, ldc, s);
}
inline rmh_status trsm(rmh_handle h, int l, int up, int tr, int un, size_t m, size_t n, double al, const double* A,
                       size_t lda, double* B, size_t ldb, rmh_stream_t s) {
    return rmh_dtrsm(h, l, up, tr, un, m, n, al, A, lda, B, ldb, s);
}
inline rmh_status trsm(rmh_handle h, int l, int up, int tr, int un, size_t m, size_t n, float al, const float* A,
                       size_t lda, float* B, size_t ldb, rmh_stream_t s) {
    return rmh_strsm(h, l, up, tr, un, m, n, al, A, lda, B, ldb, s);
}
inline rmh_status trsv(rmh_handle h, int up, int tr, int un, size_t n, const double* A, size_t lda, double* x,
                       size_t incx, rmh_stream_t s) {
    return rmh_dtrsv(h, up, tr, un, n, A, lda, x, incx, s);
}
inline rmh_status trsv(rmh_handle h, int up, int tr, int un, size_t n, const float* A, size_t lda, float* x,
                       size_t incx, rmh_stream_t s) {
    return rmh_strsv(h, up, tr, un, n, A, lda, x, incx, s);
}
inline rmh_status gemv(rmh_handle h, int tr, size_t m, size_t n, double al, const double* A, size_t lda,
                       const double* x, size_t incx, double be, double* y, size_t incy, rmh_stream_t s) {
    return rmh_dgemv(h, tr, m, n, al, A, lda, x, incx, be, y, incy, s);
}
inline rmh_status gemv(rmh_handle h, int tr, size_t m, size_t n, float al, const float* A, size_t lda, const float* x,
                       size_t incx, float be, float* y, size_t incy, rmh_stream_t s) {
    return rmh_sgemv(h, tr, m, n, al, A, lda, x, incx, be, y, incy, s);
}
inline rmh_status potrf(rmh_handle h, int up, size_t n, double* A, size_t lda, int64_t* info, rmh_stream_t s) {
    return rmh_dpotrf(h, up, n, A, lda, info, s);
}
inline rmh_status potrf(rmh_handle h, int up, size_t n, float* A, size_t lda, int64_t* info, rmh_stream_t s) {
    return rmh_spotrf(h, up, n, A, lda, info, s);
}
inline rmh_status laswp(rmh_handle h, int inv, size_t m, size_t nc, double* B, size_t ldb, int rm, const int64_t* piv,
                        rmh_stream_t s) {
    return rmh_dlaswp(h, inv, m, nc, B, ldb, rm, piv, s);
}
inline rmh_status laswp(rmh_handle h, int inv, size_t m, size_t nc, float* B, size_t ldb, int rm, const int64_t* piv,
                        rmh_stream_t s) {
    return rmh_slaswp(h, inv, m, nc, B, ldb, rm, piv, s);
}
inline rmh_status pstrf(rmh_handle h, int up, size_t n, double* A, size_t lda, int64_t* piv, int64_t* rank,
                        rmh_stream_t s) {
    return rmh_dpstrf(h, up, n, A, lda, piv, rank, s);
}
inline rmh_status pstrf(rmh_handle h, int up, size_t n, float* A, size_t lda, int64_t* piv, int64_t* rank,
                        rmh_stream_t s) {
    return rmh_spstrf(h, up, n, A, lda, piv, rank, s);
}
inline rmh_status chol_update(rmh_handle h, int up, size_t n, double al, double be, double* L, size_t ldl,
                              const double* v, size_t incv, int64_t* info, rmh_stream_t s) {
    return rmh_dchol_update(h, up, n, al, be, L, ldl, v, incv, info, s);
}
inline rmh_status chol_update(rmh_handle h, int up, size_t n, float al, float be, float* L, size_t ldl,
                              const float* v, size_t incv, int64_t* info, rmh_stream_t s) {
    return rmh_schol_update(h, up, n, al, be, L, ldl, v, incv, info, s);
}
inline rmh_status trmm(rmh_handle h, int l, int up, int tr, int un, size_t m, size_t n, double al, const double* A,
                       size_t lda, double* B, size_t ldb, rmh_stream_t s) {
    return rmh_dtrmm(h, l, up, tr, un, m, n, al, A, lda, B, ldb, s);
}
inline rmh_status trmm(rmh_handle h, int l, int up, int tr, int un, size_t m, size_t n, float al, const float* A,
                       size_t lda, float* B, size_t ldb, rmh_stream_t s) {
    return rmh_strmm(h, l, up, tr, un, m, n, al, A, lda, B, ldb, s);
}
inline rmh_status trmv(rmh_handle h, int up, int tr, int un, size_t n, const double* A, size_t lda, double* x,
                       size_t incx, rmh_stream_t s) {
    return rmh_dtrmv(h, up, tr, un, n, A, lda, x, incx, s);
}
inline rmh_status trmv(rmh_handle h, int up, int tr, int un, size_t n, const float* A, size_t lda, float* x,
                       size_t incx, rmh_stream_t s) {
    return rmh_strmv(h, up, tr, un, n, A, lda, x, incx, s);
}
inline rmh_status copy_matrix(rmh_handle h, int tr, size_t m, size_t n, const double* A, size_t lda, double* B,
                              size_t ldb, rmh_stream_t s) {
    return rmh_dcopy_matrix(h, tr, m, n, A, lda, B, ldb, s);
}
inline rmh_status copy_matrix(rmh_handle h, int tr, size_t m, size_t n, const float* A, size_t lda, float* B,
                              size_t ldb, rmh_stream_t s) {
    return rmh_scopy_matrix(h, tr, m, n, A, lda, B, ldb, s);
}
}  // namespace abi
}  // namespace hip
}  // namespace remora

#endif
