// remora/hip/solve.hpp -- cholesky_decomposition and solve(..., symm_pos_def)
// on hip_tag.
//
// Mirrors (same names, argument meaning and operation order, own code):
//   cholesky_decomposition ............. include/remora/decompositions.hpp:70-177
//     ctor/decompose: copy A, potrf<lower>  (:81-91; the info is ignored, as there)
//     solve(B, left):  trsm<lower,left>(L), trsm<upper,left>(L^T)  (:101-105)
//     solve(B, right): trsm<upper,right>(L^T), trsm<lower,right>(L) (:106-110)
//     solve(b):        trsv<lower,left>(L), trsv<upper,left>(L^T)   (:111-115)
//   solver_traits<MatA, symm_pos_def> ... include/remora/decompositions.hpp:602-609
//   solve(A, B, symm_pos_def(), left|right) and the lazy
//   matrix_matrix_solve::assign_to ..... include/remora/solve.hpp:110-182, 459-482
//   solve(A, b, symm_pos_def(), left|right) for a vector and the lazy
//   matrix_vector_solve::assign_to ..... include/remora/solve.hpp:40-107, 459-469
//   inv(A, symm_pos_def()) and prod(inv(A), B) -> solve ... include/remora/solve.hpp:184-248, 432-452, 485-493
#ifndef REMORA_HIP_SOLVE_HPP
#define REMORA_HIP_SOLVE_HPP

#include <stdexcept>

#include "prod.hpp"

namespace remora {

template <class MatrixStorage>
class cholesky_decomposition {
    typedef typename MatrixStorage::value_type value_type;

public:
    typedef hip_tag device_type;
    cholesky_decomposition() {}
    template <class E>
    explicit cholesky_decomposition(matrix_expression<E, hip_tag> const& e) : m_cholesky(e) {
        kernels::potrf<lower>(m_cholesky);
    }
    template <class E>
    void decompose(matrix_expression<E, hip_tag> const& e) {
        m_cholesky = e;
        kernels::potrf<lower>(m_cholesky);
    }
    MatrixStorage const& lower_factor() const { return m_cholesky; }
    auto upper_factor() const -> decltype(trans(std::declval<MatrixStorage const&>())) { return trans(m_cholesky); }

    template <class MatB>
    void solve(matrix_expression<MatB, hip_tag>& B, left) const {
        kernels::trsm<lower, left>(lower_factor(), B);
        kernels::trsm<upper, left>(upper_factor(), B);
    }
    template <class MatB>
    void solve(matrix_expression<MatB, hip_tag>& B, right) const {
        kernels::trsm<upper, right>(upper_factor(), B);
        kernels::trsm<lower, right>(lower_factor(), B);
    }
    template <class VecB, bool Left>
    void solve(vector_expression<VecB, hip_tag>& b, system_tag<Left>) const {
        kernels::trsv<lower, left>(lower_factor(), b);
        kernels::trsv<upper, left>(upper_factor(), b);
    }

    // Rank-one update A <- alpha A + beta v v^T of the factorised matrix, in
    // O(n^2) on the device (decompositions.hpp:132-169).  Throws
    // std::invalid_argument, as the reference does, when the update would
    // make the matrix indefinite; the factor is then left unchanged.
    template <class VecV>
    void update(value_type alpha, value_type beta, vector_expression<VecV, hip_tag> const& v) {
        hip::device& q = m_cholesky.queue();
        q.set_device();
        auto sl = m_cholesky.raw_storage();
        auto sv = v().raw_storage();
        // row-major storage is the transposed column-major matrix: its lower
        // factor is stored as an upper one
        const bool up = MatrixStorage::orientation::is_row_major;
        int64_t* info = hip::status_words(q, hip::status_update);
        hip::check_rmh(hip::abi::chol_update(hip::get_handle(q), up, m_cholesky.size1(), alpha, beta, sl.values,
                                             sl.leading_dimension, sv.values, sv.stride, info,
                                             hip::get_stream(q).abi()));
        if (hip::read_status_word(q, info) != 0)
            throw std::invalid_argument(
                "[cholesky_decomposition::update] update makes matrix indefinite, no update available");
    }

private:
    MatrixStorage m_cholesky;
};

// matrix_temporary<MatA> of a hip view or container: an owning matrix of the
// same element type and orientation.
template <class M>
struct matrix_temporary {
    typedef matrix<typename M::value_type, typename M::orientation, hip_tag> type;
};

// Solver for symmetric positive semi-definite systems on the pivoted
// Cholesky factorisation P^T A P = L L^T of rank r
// (symm_pos_semi_definite_solver, include/remora/decompositions.hpp:320-396):
// at full rank two triangular solves; below it the least-squares solution
// X = P L (L^T L)^-2 L^T P^T B with L the first r columns, (L^T L) factored
// once.  The row interchanges run on the device (rmh_?laswp).
template <class MatrixStorage>
class symm_pos_semi_definite_solver {
    typedef typename MatrixStorage::value_type value_type;

public:
    typedef hip_tag device_type;
    template <class E>
    explicit symm_pos_semi_definite_solver(matrix_expression<E, hip_tag> const& e)
        : m_factor(e), m_piv(std::max<std::size_t>(e().size1(), 1), e().queue()) {
        const std::size_t n = m_factor.size1();
        hip::device& q = m_factor.queue();
        q.set_device();
        hip::stream& s = hip::get_stream(q);
        auto sa = m_factor.raw_storage();
        int64_t* rank = hip::status_words(q, hip::status_rank);
        hip::check_rmh(hip::abi::pstrf(hip::get_handle(q), MatrixStorage::orientation::is_row_major, n, sa.values,
                                       sa.leading_dimension, m_piv.get(), rank, s.abi()));
        m_rank = std::size_t(hip::read_status_word(q, rank));
        if (m_rank == n || m_rank == 0) return;
        auto L = subrange(m_factor, 0, n, 0, m_rank);
        MatrixStorage LtL(m_rank, m_rank, q);
        clear(LtL);
        kernels::gemm(trans(L), L, LtL, value_type(1));
        m_cholesky.decompose(LtL);
    }
    std::size_t rank() const { return m_rank; }

    template <class MatB>
    void solve(matrix_expression<MatB, hip_tag>& B, left) const {
        const std::size_t n = m_factor.size1();
        swap_rows(B, false);
        if (m_rank == 0) {
            clear(B());
        } else if (m_rank == n) {
            kernels::trsm<lower, left>(m_factor, B);
            auto Ut = trans(m_factor);
            kernels::trsm<upper, left>(Ut, B);
        } else {
            auto L = subrange(m_factor, 0, n, 0, m_rank);
            MatrixStorage Z(m_rank, B().size2(), m_factor.queue());
            clear(Z);
            kernels::gemm(trans(L), B(), Z, value_type(1));
            m_cholesky.solve(Z, left());
            m_cholesky.solve(Z, left());
            clear(B());
            kernels::gemm(L, Z, B, value_type(1));
        }
        swap_rows(B, true);
    }
    template <class MatB>
    void solve(matrix_expression<MatB, hip_tag>& B, right) const {
        auto Bt = trans(B);  // X A = B  <=>  A X^T = B^T (A symmetric)
        solve(Bt, left());
    }

private:
    template <class M>
    void swap_rows(M& B, bool inverse) const {
        auto sb = B.raw_storage();
        hip::device& q = m_factor.queue();
        hip::check_rmh(hip::abi::laswp(hip::get_handle(q), inverse, B.size1(), B.size2(), sb.values,
                                       sb.leading_dimension, M::orientation::is_row_major, m_piv.get(),
                                       hip::get_stream(q).abi()));
    }
    template <class M>
    void swap_rows(matrix_expression<M, hip_tag>& B, bool inverse) const {
        swap_rows(B(), inverse);
    }
    template <class M>
    static void clear(M& X) {
        if (X.size1() == 0 || X.size2() == 0) return;
        auto sx = X.raw_storage();
        const bool rm = M::orientation::is_row_major;
        const std::size_t lines = rm ? X.size1() : X.size2(), len = rm ? X.size2() : X.size1();
        hip::check_hip(hipMemset2DAsync(sx.values, sx.leading_dimension * sizeof(value_type), 0,
                                        len * sizeof(value_type), lines, hip::get_stream(X.queue()).handle()),
                       "hipMemset2DAsync");
    }

    MatrixStorage m_factor;
    hip::buffer<int64_t> m_piv;
    std::size_t m_rank = 0;
    cholesky_decomposition<MatrixStorage> m_cholesky;
};

template <class MatA, class SolverType>
struct solver;
template <class MatA>
struct solver<MatA, symm_pos_def> : public cholesky_decomposition<typename matrix_temporary<MatA>::type> {
    template <class M>
    solver(M const& m, symm_pos_def = symm_pos_def())
        : cholesky_decomposition<typename matrix_temporary<MatA>::type>(m) {}
};
template <class MatA>
struct solver<MatA, symm_semi_pos_def>
    : public symm_pos_semi_definite_solver<typename matrix_temporary<MatA>::type> {
    template <class M>
    solver(M const& m, symm_semi_pos_def = symm_semi_pos_def())
        : symm_pos_semi_definite_solver<typename matrix_temporary<MatA>::type>(m) {}
};

// Lazy X = A^-1 B (left) or B A^-1 (right); evaluated on assignment to a
// device matrix: X <- B, factor a copy of A, solve in place
// (matrix_matrix_solve::assign_to, solve.hpp:165-170).  The operands are
// held by reference: evaluate the expression within the statement that
// makes it, as the reference's expression templates require.
template <class MatA, class MatB, class Side, class Tag = symm_pos_def>
class matrix_matrix_solve : public matrix_lazy_expression<matrix_matrix_solve<MatA, MatB, Side, Tag>> {
public:
    typedef typename MatB::value_type value_type;
    typedef typename MatB::orientation orientation;
    matrix_matrix_solve(MatA const& A, MatB const& B) : m_A(A), m_B(B) {}
    std::size_t size1() const { return m_B.size1(); }
    std::size_t size2() const { return m_B.size2(); }
    hip::device& queue() const { return m_B.queue(); }
    template <class MatX>
    void assign_to(MatX& X) const {
        detail::copy_dense(X, m_B);
        solver<MatA, Tag> alg(m_A);
        alg.solve(X, Side());
    }
    template <class MatX>
    void plus_assign_to(MatX&) const {
        static_assert(sizeof(MatX) == 0, "[solve] += of a solve is not provided on hip_tag; assign it to a matrix");
    }
    matrix_matrix_solve scaled(value_type) const {
        static_assert(sizeof(MatA) == 0, "[solve] -= of a solve is not provided on hip_tag; assign it to a matrix");
        return *this;
    }

private:
    MatA const& m_A;
    MatB const& m_B;
};

// Lazy x = A^-1 b for a vector b (matrix_vector_solve::assign_to,
// solve.hpp:90-101): x <- b, factor a copy of A, two triangular solves
// (cholesky_decomposition::solve(b), decompositions.hpp:111-115).
template <class MatA, class VecB, class Side>
class matrix_vector_solve : public vector_lazy_expression<matrix_vector_solve<MatA, VecB, Side>> {
public:
    typedef typename VecB::value_type value_type;
    matrix_vector_solve(MatA const& A, VecB const& b) : m_A(A), m_b(b) {}
    std::size_t size() const { return m_b.size(); }
    hip::device& queue() const { return m_b.queue(); }
    template <class VecX>
    void assign_to(VecX& x) const {
        if (x.size() != size()) throw std::invalid_argument("[solve] target size does not match");
        auto sb = m_b.raw_storage();
        auto sx = x.raw_storage();
        detail::copy_vector(sx.values, sx.stride, sb.values, sb.stride, size(), x.queue());
        solver<MatA, symm_pos_def> alg(m_A);
        alg.solve(x, Side());
    }
    template <class VecX>
    void plus_assign_to(VecX&) const {
        static_assert(sizeof(VecX) == 0, "[solve] += of a solve is not provided on hip_tag; assign it to a vector");
    }
    matrix_vector_solve scaled(value_type) const {
        static_assert(sizeof(MatA) == 0, "[solve] -= of a solve is not provided on hip_tag; assign it to a vector");
        return *this;
    }

private:
    MatA const& m_A;
    VecB const& m_b;
};

template <class MatA, class VecB, class Side>
matrix_vector_solve<MatA, VecB, Side> solve(matrix_expression<MatA, hip_tag> const& A,
                                            vector_expression<VecB, hip_tag> const& b, symm_pos_def, Side) {
    assert(A().size1() == A().size2() && A().size2() == b().size());
    return matrix_vector_solve<MatA, VecB, Side>(A(), b());
}

template <class MatA, class MatB, class Side>
matrix_matrix_solve<MatA, MatB, Side> solve(matrix_expression<MatA, hip_tag> const& A,
                                            matrix_expression<MatB, hip_tag> const& B, symm_pos_def, Side) {
    assert(A().size1() == A().size2());
    assert(Side::is_left ? A().size2() == B().size1() : A().size1() == B().size2());
    return matrix_matrix_solve<MatA, MatB, Side>(A(), B());
}

template <class MatA, class MatB, class Side>
matrix_matrix_solve<MatA, MatB, Side, symm_semi_pos_def> solve(matrix_expression<MatA, hip_tag> const& A,
                                                               matrix_expression<MatB, hip_tag> const& B,
                                                               symm_semi_pos_def, Side) {
    assert(A().size1() == A().size2());
    assert(Side::is_left ? A().size2() == B().size1() : A().size1() == B().size2());
    return matrix_matrix_solve<MatA, MatB, Side, symm_semi_pos_def>(A(), B());
}

// inv(A, symm_pos_def()) / inv(A, symm_semi_pos_def()) are only usable inside
// prod(), which rewrites to a solve
template <class MatA, class Tag = symm_pos_def>
struct matrix_inverse {
    MatA const& A;
};
template <class MatA>
matrix_inverse<MatA, symm_pos_def> inv(matrix_expression<MatA, hip_tag> const& A, symm_pos_def) {
    return {A()};
}
template <class MatA>
matrix_inverse<MatA, symm_semi_pos_def> inv(matrix_expression<MatA, hip_tag> const& A, symm_semi_pos_def) {
    return {A()};
}
template <class MatA, class Tag, class MatB>
matrix_matrix_solve<MatA, MatB, left, Tag> prod(matrix_inverse<MatA, Tag> const& Ainv,
                                                matrix_expression<MatB, hip_tag> const& B) {
    return matrix_matrix_solve<MatA, MatB, left, Tag>(Ainv.A, B());
}
template <class MatB, class MatA, class Tag>
matrix_matrix_solve<MatA, MatB, right, Tag> prod(matrix_expression<MatB, hip_tag> const& B,
                                                 matrix_inverse<MatA, Tag> const& Ainv) {
    return matrix_matrix_solve<MatA, MatB, right, Tag>(Ainv.A, B());
}
// prod(inv(A), b) and prod(b, inv(A)) for a vector b (solve.hpp:432-452)
template <class MatA, class VecB>
matrix_vector_solve<MatA, VecB, left> prod(matrix_inverse<MatA, symm_pos_def> const& Ainv,
                                           vector_expression<VecB, hip_tag> const& b) {
    return matrix_vector_solve<MatA, VecB, left>(Ainv.A, b());
}
template <class VecB, class MatA>
matrix_vector_solve<MatA, VecB, right> prod(vector_expression<VecB, hip_tag> const& b,
                                            matrix_inverse<MatA, symm_pos_def> const& Ainv) {
    return matrix_vector_solve<MatA, VecB, right>(Ainv.A, b());
}

}  // namespace remora

#endif
