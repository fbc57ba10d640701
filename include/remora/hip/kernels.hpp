// remora/hip/kernels.hpp -- kernels::{gemm, syrk, trsm, trsv, gemv, potrf}
// on hip_tag, calling the rmh_* C-ABI.
//
// These are the device-tag overloads the reference's dispatchers include
// (include/remora/kernels/{gemm,syrk,trsm,trsv,gemv,potrf}.hpp) and they keep
// their orientation rules:
//   gemm  : column-major C -> op flags from orientation mismatch; row-major C
//           computes C^T += B^T A^T (kernels/hip/gemm.hpp:45-88)
//   syrk  : row-major C flips triangle and transpose (kernels/hip/syrk.hpp:55-67)
//   trsm  : row-major A solves the transposed system with side and triangle
//           flipped; a B of the other orientation moves to the other side
//           (kernels/hip/trsm.hpp:44-88)
//   trsv  : row-major A flips triangle, right side transposes (kernels/hip/trsv.hpp:59-65)
//   gemv  : row-major A is a transposed column-major one (kernels/hip/gemv.hpp:78-100)
//   potrf : row-major storage of lower == column-major upper of the same memory.
// All operations are queued on the thread's stream of the matrix's device.
#ifndef REMORA_HIP_KERNELS_HPP
#define REMORA_HIP_KERNELS_HPP

#include <cstdint>
#include <type_traits>

#include <vector>

#include "dense.hpp"

namespace remora {
namespace hip {
namespace abi {
// precision dispatch onto the C-ABI
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
}  // namespace abi
}  // namespace hip

namespace kernels {

// ---------------------------------------------------------------- gemm --
// C += alpha * A * B (beta = 1, as the reference's glue passes)
template <class MatA, class MatB, class MatC>
void gemm(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag> const& B,
          matrix_expression<MatC, hip_tag>& C, typename MatC::value_type const& alpha) {
    typedef typename MatC::value_type T;
    static_assert(std::is_same<typename MatA::value_type, T>::value && std::is_same<typename MatB::value_type, T>::value,
                  "[gemm] Arguments do not have same element type");
    assert(A().size1() == C().size1() && B().size2() == C().size2() && A().size2() == B().size1());
    if (!MatC::orientation::is_row_major) {
        const bool ta = !std::is_same<typename MatA::orientation, typename MatC::orientation>::value;
        const bool tb = !std::is_same<typename MatB::orientation, typename MatC::orientation>::value;
        auto sa = A().raw_storage(), sb = B().raw_storage(), sc = C().raw_storage();
        hip::device& q = C().queue();
        hip::check_rmh(hip::abi::gemm(hip::get_handle(q), ta, tb, C().size1(), C().size2(), A().size2(), alpha,
                                      sa.values, sa.leading_dimension, sb.values, sb.leading_dimension, T(1),
                                      sc.values, sc.leading_dimension, hip::get_stream(q).abi()));
    } else {
        auto Ct = trans(C);
        gemm(trans(B), trans(A), Ct, alpha);
    }
}

// ---------------------------------------------------------------- syrk --
// C += alpha * A * A^T on the lower (Upper = false) or upper triangle
template <bool Upper, class MatA, class MatC>
void syrk(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatC, hip_tag>& C,
          typename MatC::value_type const& alpha) {
    typedef typename MatC::value_type T;
    static_assert(std::is_same<typename MatA::value_type, T>::value, "[syrk] Arguments do not have same element type");
    assert(A().size1() == C().size1() && C().size1() == C().size2());
    bool transA = !std::is_same<typename MatA::orientation, typename MatC::orientation>::value;
    bool up = Upper;
    if (MatC::orientation::is_row_major) {
        transA = !transA;
        up = !up;
    }
    auto sa = A().raw_storage(), sc = C().raw_storage();
    hip::device& q = C().queue();
    hip::check_rmh(hip::abi::syrk(hip::get_handle(q), up, transA, C().size1(), A().size2(), alpha, sa.values,
                                  sa.leading_dimension, T(1), sc.values, sc.leading_dimension,
                                  hip::get_stream(q).abi()));
}

// ---------------------------------------------------------------- trsm --
namespace detail {
template <bool Upper, bool Unit, bool Left, class MatA, class MatB>
void trsm_impl(MatA const& A, MatB& B, column_major) {
    typedef typename MatB::value_type T;
    const bool transA = !std::is_same<typename MatA::orientation, typename MatB::orientation>::value;
    auto sa = A.raw_storage(), sb = B.raw_storage();
    hip::device& q = B.queue();
    // a B of the other orientation is the transposed problem on the other side
    const bool left = transA ? !Left : Left;
    const std::size_t m = transA ? B.size2() : B.size1();
    const std::size_t n = transA ? B.size1() : B.size2();
    hip::check_rmh(hip::abi::trsm(hip::get_handle(q), left, Upper, transA, Unit, m, n, T(1), sa.values,
                                  sa.leading_dimension, sb.values, sb.leading_dimension, hip::get_stream(q).abi()));
}
template <bool Upper, bool Unit, bool Left, class MatA, class MatB>
void trsm_impl(MatA const& A, MatB& B, row_major) {
    auto At = trans(A);
    auto Bt = trans(B);
    trsm_impl<!Upper, Unit, !Left>(At, Bt, column_major());
}
}  // namespace detail

// Solve A X = B (Side left) or X A = B (right) in place; A triangular.
template <class Triangular, class Side, class MatA, class MatB>
void trsm(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag>& B) {
    static_assert(std::is_same<typename MatA::value_type, typename MatB::value_type>::value,
                  "[trsm] Arguments do not have same element type");
    assert(A().size1() == A().size2());
    assert(Side::is_left ? A().size2() == B().size1() : A().size2() == B().size2());
    detail::trsm_impl<Triangular::is_upper, Triangular::is_unit, Side::is_left>(A(), B(),
                                                                                typename MatA::orientation());
}

// ---------------------------------------------------------------- trsv --
template <class Triangular, class Side, class MatA, class VecV>
void trsv(matrix_expression<MatA, hip_tag> const& A, vector_expression<VecV, hip_tag>& v) {
    assert(A().size1() == A().size2() && A().size2() == v().size());
    bool transA = MatA::orientation::is_row_major;
    const bool up = transA ? !Triangular::is_upper : Triangular::is_upper;
    if (!Side::is_left) transA = !transA;
    auto sa = A().raw_storage();
    auto sv = v().raw_storage();
    hip::device& q = v().queue();
    hip::check_rmh(hip::abi::trsv(hip::get_handle(q), up, transA, Triangular::is_unit, A().size1(), sa.values,
                                  sa.leading_dimension, sv.values, sv.stride, hip::get_stream(q).abi()));
}

// ---------------------------------------------------------------- gemv --
// v += alpha * A * x
template <class MatA, class VecX, class VecV>
void gemv(matrix_expression<MatA, hip_tag> const& A, vector_expression<VecX, hip_tag> const& x,
          vector_expression<VecV, hip_tag>& v, typename VecV::value_type const& alpha) {
    typedef typename VecV::value_type T;
    assert(A().size1() == v().size() && A().size2() == x().size());
    const bool transA = MatA::orientation::is_row_major;
    std::size_t m = A().size1(), n = A().size2();
    if (transA) std::swap(m, n);
    auto sa = A().raw_storage();
    auto sx = x().raw_storage();
    auto sv = v().raw_storage();
    hip::device& q = A().queue();
    hip::check_rmh(hip::abi::gemv(hip::get_handle(q), transA, m, n, alpha, sa.values, sa.leading_dimension, sx.values,
                                  sx.stride, T(1), sv.values, sv.stride, hip::get_stream(q).abi()));
}

// --------------------------------------------------------------- potrf --
// In-place Cholesky on the device.  Returns LAPACK's info (0, or k+1 for the
// first leading minor that is not positive definite); that needs the
// stream to finish.  (The reference HIP overload returned 0 unconditionally,
// kernels/hip/potrf.hpp:139; its CPU path returns a nonzero info.)
template <class Triangular, class MatA>
std::size_t potrf(matrix_container<MatA, hip_tag>& A) {
    assert(A().size1() == A().size2());
    hip::device& q = A().queue();
    q.set_device();
    const bool up = MatA::orientation::is_row_major ? !Triangular::is_upper : Triangular::is_upper;
    auto sa = A().raw_storage();
    hip::buffer<int64_t> info(1, q);
    hip::stream& s = hip::get_stream(q);
    hip::check_rmh(hip::abi::potrf(hip::get_handle(q), up, A().size1(), sa.values, sa.leading_dimension, info.get(),
                                   s.abi()));
    int64_t h = 0;
    hip::check_hip(hipMemcpyAsync(&h, info.get(), sizeof(h), hipMemcpyDeviceToHost, s.handle()), "hipMemcpyAsync");
    s.synchronize();
    return std::size_t(h);
}

// Pivoted Cholesky on the device, kernels::pstrf<lower|upper>(A, P)
// (kernels/pstrf.hpp:58-66): A is overwritten by the factor, P (a host
// container of n indices with P(i) assignable, as the reference's
// permutation_matrix) receives the transpositions, and the rank is returned.
template <class Triangular, class MatA, class VecP>
std::size_t pstrf(matrix_container<MatA, hip_tag>& A, VecP& P) {
    assert(A().size1() == A().size2());
    const std::size_t n = A().size1();
    if (n == 0) return 0;
    hip::device& q = A().queue();
    q.set_device();
    const bool up = MatA::orientation::is_row_major ? !Triangular::is_upper : Triangular::is_upper;
    auto sa = A().raw_storage();
    hip::buffer<int64_t> piv(n, q), rank(1, q);
    hip::stream& s = hip::get_stream(q);
    hip::check_rmh(hip::abi::pstrf(hip::get_handle(q), up, n, sa.values, sa.leading_dimension, piv.get(), rank.get(),
                                   s.abi()));
    std::vector<int64_t> hp(n);
    int64_t r = 0;
    hip::check_hip(hipMemcpyAsync(hp.data(), piv.get(), n * sizeof(int64_t), hipMemcpyDeviceToHost, s.handle()),
                   "hipMemcpyAsync");
    hip::check_hip(hipMemcpyAsync(&r, rank.get(), sizeof(r), hipMemcpyDeviceToHost, s.handle()), "hipMemcpyAsync");
    s.synchronize();
    for (std::size_t i = 0; i != n; ++i) P(i) = hp[i];
    return std::size_t(r);
}

}  // namespace kernels
}  // namespace remora

#endif
