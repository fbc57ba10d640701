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
//   inv(A, symm_pos_def()) and prod(inv(A), B) -> solve ... include/remora/solve.hpp:184-248, 432-452, 485-493
#ifndef REMORA_HIP_SOLVE_HPP
#define REMORA_HIP_SOLVE_HPP

#include <stdexcept>

#include "kernels.hpp"

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
        hip::buffer<int64_t> info(1, q);
        hip::stream& s = hip::get_stream(q);
        hip::check_rmh(hip::abi::chol_update(hip::get_handle(q), up, m_cholesky.size1(), alpha, beta, sl.values,
                                             sl.leading_dimension, sv.values, sv.stride, info.get(), s.abi()));
        int64_t h = 0;
        hip::check_hip(hipMemcpyAsync(&h, info.get(), sizeof(h), hipMemcpyDeviceToHost, s.handle()), "hipMemcpyAsync");
        s.synchronize();
        if (h != 0)
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

template <class MatA, class SolverType>
struct solver;
template <class MatA>
struct solver<MatA, symm_pos_def> : public cholesky_decomposition<typename matrix_temporary<MatA>::type> {
    template <class M>
    solver(M const& m, symm_pos_def = symm_pos_def())
        : cholesky_decomposition<typename matrix_temporary<MatA>::type>(m) {}
};

// Lazy X = A^-1 B (left) or B A^-1 (right); evaluated on assignment to a
// device matrix: X <- B, factor a copy of A, solve in place.
template <class MatA, class MatB, class Side>
class matrix_matrix_solve {
public:
    typedef typename MatB::value_type value_type;
    typedef typename MatB::orientation orientation;
    matrix_matrix_solve(MatA const& A, MatB const& B) : m_A(A), m_B(B) {}
    std::size_t size1() const { return m_B.size1(); }
    std::size_t size2() const { return m_B.size2(); }
    template <class MatX>
    void assign_to(MatX& X) const {
        X = m_B;
        solver<MatA, symm_pos_def> alg(m_A);
        alg.solve(X, Side());
    }
    template <class O>
    operator matrix<value_type, O, hip_tag>() const {
        matrix<value_type, O, hip_tag> X(m_B.queue());
        assign_to(X);
        return X;
    }

private:
    MatA const& m_A;
    MatB const& m_B;
};

template <class MatA, class MatB, class Side>
matrix_matrix_solve<MatA, MatB, Side> solve(matrix_expression<MatA, hip_tag> const& A,
                                            matrix_expression<MatB, hip_tag> const& B, symm_pos_def, Side) {
    assert(A().size1() == A().size2());
    assert(Side::is_left ? A().size2() == B().size1() : A().size1() == B().size2());
    return matrix_matrix_solve<MatA, MatB, Side>(A(), B());
}

// inv(A, symm_pos_def()) is only usable inside prod(), which rewrites to a solve
template <class MatA>
struct matrix_inverse {
    MatA const& A;
};
template <class MatA>
matrix_inverse<MatA> inv(matrix_expression<MatA, hip_tag> const& A, symm_pos_def) {
    return {A()};
}
template <class MatA, class MatB>
matrix_matrix_solve<MatA, MatB, left> prod(matrix_inverse<MatA> const& Ainv, matrix_expression<MatB, hip_tag> const& B) {
    return matrix_matrix_solve<MatA, MatB, left>(Ainv.A, B());
}
template <class MatB, class MatA>
matrix_matrix_solve<MatA, MatB, right> prod(matrix_expression<MatB, hip_tag> const& B, matrix_inverse<MatA> const& Ainv) {
    return matrix_matrix_solve<MatA, MatB, right>(Ainv.A, B());
}

}  // namespace remora

#endif
