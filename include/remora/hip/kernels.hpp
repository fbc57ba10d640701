// remora/hip/kernels.hpp -- kernels::{gemm, syrk, trsm, trsv, gemv, trmm,
// trmv, potrf, pstrf} on hip_tag.
//
// These are the device-tag overloads the reference's dispatchers include
// (include/remora/kernels/{gemm,syrk,trsm,trsv,gemv,trmm,trmv,potrf}.hpp).
// The BLAS ones call hip::get_blas(queue) -- the blas_device of
// rmh_backend.hpp -- exactly as the reference's glue calls its cuBLAS-backed
// one, and they keep its orientation rules:
//   gemm  : column-major C -> op flags from orientation mismatch; row-major C
//           computes C^T += B^T A^T (kernels/hip/gemm.hpp:45-88)
//   syrk  : row-major C flips triangle and transpose (kernels/hip/syrk.hpp:55-67)
//   trsm, trmm : row-major A works on the transposed system with side and
//           triangle flipped; a B of the other orientation moves to the other
//           side (kernels/hip/trsm.hpp:44-88, kernels/hip/trmm.hpp:44-80)
//   trsv, trmv : row-major A flips triangle and transposes; the right side
//           of trsv transposes (kernels/hip/trsv.hpp:59-65, kernels/hip/trmv.hpp:57-70)
//   gemv  : row-major A is a transposed column-major one (kernels/hip/gemv.hpp:78-100)
//   potrf : row-major storage of lower == column-major upper of the same memory.
// All operations are queued on the thread's stream of the matrix's device.
#ifndef REMORA_HIP_KERNELS_HPP
#define REMORA_HIP_KERNELS_HPP

#include <cstdint>
#include <type_traits>
#include <vector>

#include "dense.hpp"
#include "rmh_backend.hpp"

namespace remora {
namespace kernels {

namespace detail {
// C = alpha * A * B + beta * C for dense device operands of any orientation
template <class MatA, class MatB, class MatC>
void gemm_beta(MatA const& A, MatB const& B, MatC& C, typename MatC::value_type alpha,
               typename MatC::value_type beta) {
    if (!MatC::orientation::is_row_major) {
        const bool ta = !std::is_same<typename MatA::orientation, typename MatC::orientation>::value;
        const bool tb = !std::is_same<typename MatB::orientation, typename MatC::orientation>::value;
        auto sa = A.raw_storage(), sb = B.raw_storage(), sc = C.raw_storage();
        hip::device& q = C.queue();
        hip::get_blas(q).gemm(ta, tb, C.size1(), C.size2(), A.size2(), alpha, sa.values, sa.leading_dimension,
                              sb.values, sb.leading_dimension, beta, sc.values, sc.leading_dimension,
                              hip::get_stream(q));
    } else {
        auto Ct = trans(C);
        auto Bt = trans(B);
        auto At = trans(A);
        gemm_beta(Bt, At, Ct, alpha, beta);
    }
}

// v = alpha * A * x + beta * v
template <class MatA, class VecX, class VecV>
void gemv_beta(MatA const& A, VecX const& x, VecV& v, typename VecV::value_type alpha,
               typename VecV::value_type beta) {
    const bool transA = MatA::orientation::is_row_major;
    std::size_t m = A.size1(), n = A.size2();
    if (transA) std::swap(m, n);
    auto sa = A.raw_storage();
    auto sx = x.raw_storage();
    auto sv = v.raw_storage();
    hip::device& q = v.queue();
    hip::get_blas(q).gemv(transA, m, n, alpha, sa.values, sa.leading_dimension, sx.values, sx.stride, beta, sv.values,
                          sv.stride, hip::get_stream(q));
}
}  // namespace detail

// ---------------------------------------------------------------- gemm --
// C += alpha * A * B (beta = 1, as the reference's glue passes)
template <class MatA, class MatB, class MatC>
void gemm(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag> const& B,
          matrix_expression<MatC, hip_tag>& C, typename MatC::value_type const& alpha) {
    typedef typename MatC::value_type T;
    static_assert(std::is_same<typename MatA::value_type, T>::value && std::is_same<typename MatB::value_type, T>::value,
                  "[gemm] Arguments do not have same element type");
    assert(A().size1() == C().size1() && B().size2() == C().size2() && A().size2() == B().size1());
    detail::gemm_beta(A(), B(), C(), alpha, T(1));
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
    hip::get_blas(q).syrk(up, transA, C().size1(), A().size2(), alpha, sa.values, sa.leading_dimension, T(1),
                          sc.values, sc.leading_dimension, hip::get_stream(q));
}

// ------------------------------------------------------- trsm and trmm --
namespace detail {
// Triangular matrix op(A) applied to B from the side Left in place: a solve
// (Solve) or a product.  With column-major B: A of the same orientation is
// used as stored; a row-major A is the transpose of a column-major matrix of
// the other triangle, and the problem is the transposed one on the other side
// -- that is, the flags passed to blas_device describe A's memory.
template <bool Solve, bool Upper, bool Unit, bool Left, class MatA, class MatB>
void tri_impl(MatA const& A, MatB& B, column_major) {
    typedef typename MatB::value_type T;
    const bool transA = !std::is_same<typename MatA::orientation, typename MatB::orientation>::value;
    auto sa = A.raw_storage(), sb = B.raw_storage();
    hip::device& q = B.queue();
    const bool up = transA ? !Upper : Upper;  // triangle of A's column-major memory
    if (Solve)
        hip::get_blas(q).trsm(Left, up, transA, Unit, B.size1(), B.size2(), T(1), sa.values, sa.leading_dimension,
                              sb.values, sb.leading_dimension, hip::get_stream(q));
    else
        hip::get_blas(q).trmm(Left, up, transA, Unit, B.size1(), B.size2(), T(1), sa.values, sa.leading_dimension,
                              sb.values, sb.leading_dimension, hip::get_stream(q));
}
// row-major B: B^T is column-major, and A B = C  <=>  B^T A^T = C^T
template <bool Solve, bool Upper, bool Unit, bool Left, class MatA, class MatB>
void tri_impl(MatA const& A, MatB& B, row_major) {
    auto At = trans(A);
    auto Bt = trans(B);
    tri_impl<Solve, !Upper, Unit, !Left>(At, Bt, column_major());
}
}  // namespace detail

// Solve A X = B (Side left) or X A = B (right) in place; A triangular.
template <class Triangular, class Side, class MatA, class MatB>
void trsm(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag>& B) {
    static_assert(std::is_same<typename MatA::value_type, typename MatB::value_type>::value,
                  "[trsm] Arguments do not have same element type");
    assert(A().size1() == A().size2());
    assert(Side::is_left ? A().size2() == B().size1() : A().size2() == B().size2());
    detail::tri_impl<true, Triangular::is_upper, Triangular::is_unit, Side::is_left>(A(), B(),
                                                                                      typename MatB::orientation());
}

// B <- A B with A triangular (kernels/hip/trmm.hpp:83-100).
template <bool Upper, bool Unit, class MatA, class MatB>
void trmm(matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag>& B) {
    static_assert(std::is_same<typename MatA::value_type, typename MatB::value_type>::value,
                  "[trmm] Arguments do not have same element type");
    assert(A().size1() == A().size2() && A().size2() == B().size1());
    detail::tri_impl<false, Upper, Unit, true>(A(), B(), typename MatB::orientation());
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
    hip::get_blas(q).trsv(up, transA, Triangular::is_unit, A().size1(), sa.values, sa.leading_dimension, sv.values,
                          sv.stride, hip::get_stream(q));
}

// ---------------------------------------------------------------- trmv --
// v <- A v with A triangular (kernels/hip/trmv.hpp:44-78)
template <bool Upper, bool Unit, class MatA, class VecV>
void trmv(matrix_expression<MatA, hip_tag> const& A, vector_expression<VecV, hip_tag>& v) {
    static_assert(std::is_same<typename MatA::value_type, typename VecV::value_type>::value,
                  "[trmv] Arguments do not have same element type");
    assert(A().size1() == A().size2() && A().size2() == v().size());
    const bool transA = MatA::orientation::is_row_major;
    const bool up = transA ? !Upper : Upper;
    auto sa = A().raw_storage();
    auto sv = v().raw_storage();
    hip::device& q = v().queue();
    hip::get_blas(q).trmv(up, transA, Unit, A().size1(), sa.values, sa.leading_dimension, sv.values, sv.stride,
                          hip::get_stream(q));
}

// ---------------------------------------------------------------- gemv --
// v += alpha * A * x
template <class MatA, class VecX, class VecV>
void gemv(matrix_expression<MatA, hip_tag> const& A, vector_expression<VecX, hip_tag> const& x,
          vector_expression<VecV, hip_tag>& v, typename VecV::value_type const& alpha) {
    typedef typename VecV::value_type T;
    assert(A().size1() == v().size() && A().size2() == x().size());
    detail::gemv_beta(A(), x(), v(), alpha, T(1));
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
    hip::stream& s = hip::get_stream(q);
    int64_t* info = hip::status_words(q);
    hip::check_rmh(hip::abi::potrf(hip::get_handle(q), up, A().size1(), sa.values, sa.leading_dimension, info,
                                   s.abi()));
    return std::size_t(hip::read_status_word(q, info));
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
    hip::buffer<int64_t> piv(n, q);
    hip::stream& s = hip::get_stream(q);
    int64_t* rank = hip::status_words(q, hip::status_rank);
    hip::check_rmh(hip::abi::pstrf(hip::get_handle(q), up, n, sa.values, sa.leading_dimension, piv.get(), rank,
                                   s.abi()));
    std::vector<int64_t> hp(n);
    hip::check_hip(hipMemcpyAsync(hp.data(), piv.get(), n * sizeof(int64_t), hipMemcpyDeviceToHost, s.handle()),
                   "hipMemcpyAsync");
    const int64_t r = hip::read_status_word(q, rank);
    for (std::size_t i = 0; i != n; ++i) P(i) = hp[i];
    return std::size_t(r);
}

}  // namespace kernels
}  // namespace remora

#endif
