// remora/hip/prod.hpp -- prod(A, B), prod(A, x), prod(x, A), their scalar
// multiples, and noalias(C) = / += / -= on hip_tag.
//
// Mirrors (same names and meaning, own code):
//   prod(A, B) -> matrix_matrix_prod, evaluated by kernels::gemm
//       ........ include/remora/matrix_expression.hpp:421-427,
//                include/remora/detail/matrix_expression_classes.hpp:940-1012
//   prod(A, x), prod(x, A) -> matrix_vector_prod, evaluated by kernels::gemv
//       ........ include/remora/matrix_expression.hpp:363-383,
//                include/remora/detail/matrix_expression_classes.hpp:760-860
//   alpha * prod(...), prod(...) * alpha, -prod(...): the product's alpha
//   noalias(C) -> noalias_proxy ....... include/remora/assignment.hpp:635-700
// Assignment writes with beta = 0 (the reference clears the target and then
// adds, matrix_expression_classes.hpp:1005-1008); += passes beta = 1, -= the
// negated alpha.  As in the reference, noalias promises that the target does
// not overlap an operand; a plain C = prod(A, B) evaluates into a temporary.
#ifndef REMORA_HIP_PROD_HPP
#define REMORA_HIP_PROD_HPP

#include <type_traits>

#include "kernels.hpp"

namespace remora {

namespace detail {
template <class M>
matrix_view<typename M::value_type, typename M::orientation, hip_tag> as_view(matrix_expression<M, hip_tag> const& m) {
    auto s = m().raw_storage();
    return {s.values, m().size1(), m().size2(), s.leading_dimension, m().queue()};
}
template <class V>
vector_view<typename V::value_type, hip_tag> as_view(vector_expression<V, hip_tag> const& v) {
    auto s = v().raw_storage();
    return {s.values, v().size(), s.stride, v().queue()};
}
}  // namespace detail

// alpha * A * B, A and B dense device matrices of any orientation
template <class T, class OA, class OB>
class matrix_matrix_prod : public matrix_lazy_expression<matrix_matrix_prod<T, OA, OB>> {
public:
    typedef T value_type;
    typedef matrix_view<T, OA, hip_tag> lhs_type;
    typedef matrix_view<T, OB, hip_tag> rhs_type;
    matrix_matrix_prod(lhs_type const& a, rhs_type const& b, T alpha) : m_lhs(a), m_rhs(b), m_alpha(alpha) {}
    std::size_t size1() const { return m_lhs.size1(); }
    std::size_t size2() const { return m_rhs.size2(); }
    hip::device& queue() const { return m_lhs.queue(); }
    lhs_type const& lhs() const { return m_lhs; }
    rhs_type const& rhs() const { return m_rhs; }
    T alpha() const { return m_alpha; }
    matrix_matrix_prod scaled(T s) const { return {m_lhs, m_rhs, m_alpha * s}; }

    template <class MatX>
    void assign_to(MatX& X) const {
        check(X);
        kernels::detail::gemm_beta(m_lhs, m_rhs, X, m_alpha, T(0));
    }
    template <class MatX>
    void plus_assign_to(MatX& X) const {
        check(X);
        kernels::detail::gemm_beta(m_lhs, m_rhs, X, m_alpha, T(1));
    }

private:
    template <class MatX>
    void check(MatX const& X) const {
        static_assert(std::is_same<typename MatX::value_type, T>::value, "[prod] target has another element type");
        if (X.size1() != size1() || X.size2() != size2())
            throw std::invalid_argument("[prod] target size does not match the product");
    }
    lhs_type m_lhs;
    rhs_type m_rhs;
    T m_alpha;
};

// alpha * A * x
template <class T, class OA>
class matrix_vector_prod : public vector_lazy_expression<matrix_vector_prod<T, OA>> {
public:
    typedef T value_type;
    typedef matrix_view<T, OA, hip_tag> matrix_type;
    typedef vector_view<T, hip_tag> vector_type;
    matrix_vector_prod(matrix_type const& A, vector_type const& x, T alpha) : m_matrix(A), m_vector(x), m_alpha(alpha) {}
    std::size_t size() const { return m_matrix.size1(); }
    hip::device& queue() const { return m_matrix.queue(); }
    T alpha() const { return m_alpha; }
    matrix_vector_prod scaled(T s) const { return {m_matrix, m_vector, m_alpha * s}; }

    template <class VecX>
    void assign_to(VecX& x) const {
        check(x);
        kernels::detail::gemv_beta(m_matrix, m_vector, x, m_alpha, T(0));
    }
    template <class VecX>
    void plus_assign_to(VecX& x) const {
        check(x);
        kernels::detail::gemv_beta(m_matrix, m_vector, x, m_alpha, T(1));
    }

private:
    template <class VecX>
    void check(VecX const& x) const {
        static_assert(std::is_same<typename VecX::value_type, T>::value, "[prod] target has another element type");
        if (x.size() != size()) throw std::invalid_argument("[prod] target size does not match the product");
    }
    matrix_type m_matrix;
    vector_type m_vector;
    T m_alpha;
};

template <class MatA, class MatB>
matrix_matrix_prod<typename MatA::value_type, typename MatA::orientation, typename MatB::orientation> prod(
    matrix_expression<MatA, hip_tag> const& A, matrix_expression<MatB, hip_tag> const& B) {
    static_assert(std::is_same<typename MatA::value_type, typename MatB::value_type>::value,
                  "[prod] arguments do not have the same element type");
    if (A().size2() != B().size1()) throw std::invalid_argument("[prod] inner sizes do not match");
    return {detail::as_view(A), detail::as_view(B), typename MatA::value_type(1)};
}

template <class MatA, class VecX>
matrix_vector_prod<typename MatA::value_type, typename MatA::orientation> prod(
    matrix_expression<MatA, hip_tag> const& A, vector_expression<VecX, hip_tag> const& x) {
    static_assert(std::is_same<typename MatA::value_type, typename VecX::value_type>::value,
                  "[prod] arguments do not have the same element type");
    if (A().size2() != x().size()) throw std::invalid_argument("[prod] sizes do not match");
    return {detail::as_view(A), detail::as_view(x), typename MatA::value_type(1)};
}

// x^T A = A^T x (matrix_expression.hpp:378-383)
template <class VecX, class MatA>
matrix_vector_prod<typename MatA::value_type, typename MatA::orientation::transposed_orientation> prod(
    vector_expression<VecX, hip_tag> const& x, matrix_expression<MatA, hip_tag> const& A) {
    return prod(trans(A), x);
}

// scalar multiples: alpha * prod, prod * alpha, -prod
template <class S, class T, class OA, class OB, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
matrix_matrix_prod<T, OA, OB> operator*(S alpha, matrix_matrix_prod<T, OA, OB> const& p) {
    return p.scaled(T(alpha));
}
template <class S, class T, class OA, class OB, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
matrix_matrix_prod<T, OA, OB> operator*(matrix_matrix_prod<T, OA, OB> const& p, S alpha) {
    return p.scaled(T(alpha));
}
template <class T, class OA, class OB>
matrix_matrix_prod<T, OA, OB> operator-(matrix_matrix_prod<T, OA, OB> const& p) {
    return p.scaled(T(-1));
}
template <class S, class T, class OA, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
matrix_vector_prod<T, OA> operator*(S alpha, matrix_vector_prod<T, OA> const& p) {
    return p.scaled(T(alpha));
}
template <class S, class T, class OA, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
matrix_vector_prod<T, OA> operator*(matrix_vector_prod<T, OA> const& p, S alpha) {
    return p.scaled(T(alpha));
}
template <class T, class OA>
matrix_vector_prod<T, OA> operator-(matrix_vector_prod<T, OA> const& p) {
    return p.scaled(T(-1));
}

// noalias(C): assignment straight into C's storage, no temporary.
template <class View>
class noalias_proxy {
public:
    explicit noalias_proxy(View const& v) : m_target(v) {}
    template <class E>
    noalias_proxy& operator=(matrix_lazy_expression<E> const& e) {
        e().assign_to(m_target);
        return *this;
    }
    template <class E>
    noalias_proxy& operator+=(matrix_lazy_expression<E> const& e) {
        e().plus_assign_to(m_target);
        return *this;
    }
    template <class E>
    noalias_proxy& operator-=(matrix_lazy_expression<E> const& e) {
        e().scaled(typename View::value_type(-1)).plus_assign_to(m_target);
        return *this;
    }
    template <class E>
    noalias_proxy& operator=(vector_lazy_expression<E> const& e) {
        e().assign_to(m_target);
        return *this;
    }
    template <class E>
    noalias_proxy& operator+=(vector_lazy_expression<E> const& e) {
        e().plus_assign_to(m_target);
        return *this;
    }
    template <class E>
    noalias_proxy& operator-=(vector_lazy_expression<E> const& e) {
        e().scaled(typename View::value_type(-1)).plus_assign_to(m_target);
        return *this;
    }

private:
    View m_target;
};

template <class M>
noalias_proxy<matrix_view<typename M::value_type, typename M::orientation, hip_tag>> noalias(
    matrix_expression<M, hip_tag>& m) {
    return noalias_proxy<matrix_view<typename M::value_type, typename M::orientation, hip_tag>>(detail::as_view(m));
}
template <class T, class O>
noalias_proxy<matrix_view<T, O, hip_tag>> noalias(matrix_view<T, O, hip_tag> const& m) {
    return noalias_proxy<matrix_view<T, O, hip_tag>>(m);
}
template <class V>
noalias_proxy<vector_view<typename V::value_type, hip_tag>> noalias(vector_expression<V, hip_tag>& v) {
    return noalias_proxy<vector_view<typename V::value_type, hip_tag>>(detail::as_view(v));
}
template <class T>
noalias_proxy<vector_view<T, hip_tag>> noalias(vector_view<T, hip_tag> const& v) {
    return noalias_proxy<vector_view<T, hip_tag>>(v);
}

}  // namespace remora

#endif
