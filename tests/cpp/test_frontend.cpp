// C++ front-end tests on hip_tag, restating the reference's HIP test
// programs against the header-only front-end:
//   Test/hip_cholesky.cpp:36-65          potrf<lower|upper> on row/column-major == CPU potrf
//   Test/hip_syrk.cpp:17-98              float syrk closed form
//   Test/hip_triangular_solve.cpp:20-193 float trsm / trsv residuals
//   Test/hip_prod.cpp:31-261             float prod(A,x), prod(x,A), prod(A,B) with noalias
//                                        = / += / -=, all orientations
//   hip::blas_device (cublas_backend.hpp:121-431): every method, float and double
//   Test/symm_solve.cpp:89-139           solve(A, B, symm_pos_def(), left|right) on the device
// The CPU reference for potrf is the oracle's restatement of Remora's CPU
// kernels (liboracle.so).  Prints one line per check group; exit status 0
// iff every check passed.
#include <cmath>
#include <cstdio>
#include <string>

#include "remora/hip/solve.hpp"

extern "C" size_t orc_dpotrf(double* A, size_t n, size_t lda, int row_major, int upper);
extern "C" size_t orc_dpstrf(double* A, size_t m, size_t lda, int row_major, int upper, int64_t* P);

using namespace remora;

static int g_fail = 0, g_checks = 0;
#define CHECK_CLOSE_PCT(a, b, pct)                                                                   \
    do {                                                                                             \
        ++g_checks;                                                                                  \
        double a_ = (a), b_ = (b);                                                                   \
        double d_ = std::fabs(a_ - b_), s_ = std::fmax(std::fabs(a_), std::fabs(b_));                \
        if (!(d_ <= (pct) / 100.0 * s_)) {                                                           \
            if (g_fail < 20) std::printf("  FAIL %s:%d %.17g vs %.17g\n", __FILE__, __LINE__, a_, b_); \
            ++g_fail;                                                                                \
        }                                                                                            \
    } while (0)
#define CHECK_SMALL(a, tol)                                                                          \
    do {                                                                                             \
        ++g_checks;                                                                                  \
        double a_ = (a);                                                                             \
        if (!(std::fabs(a_) <= (tol))) {                                                             \
            if (g_fail < 20) std::printf("  FAIL %s:%d |%.17g| > %g\n", __FILE__, __LINE__, a_, double(tol)); \
            ++g_fail;                                                                                \
        }                                                                                            \
    } while (0)

// createSymm with the HIP test's diagonal (Test/hip_cholesky.cpp:14-33) or the CPU one
template <class O>
matrix<double, O, cpu_tag> create_symm(std::size_t n, bool hip_variant) {
    matrix<double, row_major, cpu_tag> R(n, n, 0.0);
    for (std::size_t i = 0; i != n; ++i) {
        for (std::size_t j = 0; j < i; ++j) R(i, j) = 0.2 / std::abs((int)i - (int)j);
        R(i, i) = (hip_variant ? 1.0 : 0.0) + 0.5 / n * i + 1;
    }
    matrix<double, O, cpu_tag> A(n, n, 0.0);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != n; ++j) {
            double s = 0;
            for (std::size_t k = 0; k != n; ++k) s += R(i, k) * R(j, k);
            A(i, j) = s;
        }
    return A;
}

template <class O>
void potrf_test() {
    const std::size_t n = 123;
    auto A = create_symm<O>(n, true);
    auto dlow = copy_to_device(A, hip_tag());
    auto dup = copy_to_device(A, hip_tag());
    std::size_t i1 = kernels::potrf<lower>(dlow);
    std::size_t i2 = kernels::potrf<upper>(dup);
    CHECK_SMALL(double(i1 + i2), 0);
    auto low = copy_to_cpu(dlow);
    auto up = copy_to_cpu(dup);
    auto low_ref = A, up_ref = A;
    const bool rm = O::is_row_major;
    orc_dpotrf(low_ref.data(), n, n, rm, 0);
    orc_dpotrf(up_ref.data(), n, n, rm, 1);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            CHECK_CLOSE_PCT(low(i, j), low_ref(i, j), 1e-12);
            CHECK_CLOSE_PCT(up(i, j), up_ref(i, j), 1e-12);
        }
}

template <class O>
void syrk_test() {
    const std::size_t dims = 936, K = 1039;
    matrix<float, O, cpu_tag> arg(dims, K);
    for (std::size_t i = 0; i != dims; ++i)
        for (std::size_t j = 0; j != K; ++j) arg(i, j) = float((1.0 / dims) * i + 0.2 / K * j + 1);
    auto darg = copy_to_device(arg, hip_tag());
    for (int upper_case = 0; upper_case < 2; ++upper_case) {
        matrix<float, O, cpu_tag> init(dims, dims, 3.0f);
        auto dres = copy_to_device(init, hip_tag());
        if (upper_case) kernels::syrk<true>(darg, dres, 2.0f);
        else kernels::syrk<false>(darg, dres, 2.0f);
        auto res = copy_to_cpu(dres);
        for (std::size_t i = 0; i < dims; i += 7)
            for (std::size_t j = 0; j < dims; j += 5) {
                bool touched = upper_case ? j >= i : j <= i;
                if (!touched) {
                    CHECK_CLOSE_PCT(res(i, j), 3.0, 1e-10);
                } else {
                    double s = 0;
                    for (std::size_t k = 0; k != K; ++k) s += double(arg(i, k)) * double(arg(j, k));
                    CHECK_CLOSE_PCT(res(i, j), 2.0 * s + 3.0, 1e-3);  // FP32 sum of 1039 terms vs exact
                }
            }
    }
}

// Test/hip_prod.cpp:113-261: noalias(C) += -2 * prod(A, B) (C = 1.5) and
// noalias(C) = -2.0 * prod(A, B), every orientation of A, B and C, against
// the textbook sum; plus C = prod(A, B), -= and kernels::gemm directly.
template <class OA, class OB, class OC>
void prod_test() {
    const std::size_t rows = 50, cols = 80, middle = 33;
    matrix<float, OA, cpu_tag> a(rows, middle);
    matrix<float, OB, cpu_tag> b(middle, cols);
    for (std::size_t i = 0; i != rows; ++i)
        for (std::size_t j = 0; j != middle; ++j) a(i, j) = float(0.1 * i + 0.02 * j);
    for (std::size_t i = 0; i != middle; ++i)
        for (std::size_t j = 0; j != cols; ++j) b(i, j) = float(0.05 * i - 0.01 * j + 1);
    matrix<float, OC, cpu_tag> c(rows, cols, 1.5f);
    auto da = copy_to_device(a, hip_tag());
    auto db = copy_to_device(b, hip_tag());
    auto check = [&](matrix<float, OC, hip_tag> const& dres, double factor, double init) {
        auto res = copy_to_cpu(dres);
        for (std::size_t i = 0; i != rows; ++i)
            for (std::size_t j = 0; j != cols; ++j) {
                double s = init;
                for (std::size_t k = 0; k != middle; ++k) s += factor * double(a(i, k)) * double(b(k, j));
                CHECK_CLOSE_PCT(res(i, j), s, 1e-2);
            }
    };
    {
        auto dc = copy_to_device(c, hip_tag());
        noalias(dc) += -2 * prod(da, db);
        check(dc, -2.0, 1.5);
    }
    {
        auto dc = copy_to_device(c, hip_tag());
        noalias(dc) = -2.0 * prod(da, db);
        check(dc, -2.0, 0.0);
    }
    {
        auto dc = copy_to_device(c, hip_tag());
        noalias(dc) -= prod(da, db) * 0.5;
        check(dc, -0.5, 1.5);
    }
    {
        matrix<float, OC, hip_tag> dc = prod(da, db);
        check(dc, 1.0, 0.0);
    }
    {
        auto dc = copy_to_device(c, hip_tag());
        kernels::gemm(da, db, dc, -2.0f);
        check(dc, -2.0, 1.5);
    }
}

// Test/hip_prod.cpp:31-106: noalias(r) += -2 * prod(A, x) and prod(x, A^T)
// with r = 1.5, and noalias(r) = ..., A 50 x 40 in both orientations.
template <class O>
void prod_vector_test() {
    const std::size_t rows = 50, columns = 40;
    matrix<float, O, cpu_tag> a(rows, columns), at(columns, rows);
    for (std::size_t i = 0; i != rows; ++i)
        for (std::size_t j = 0; j != columns; ++j) at(j, i) = a(i, j) = float(0.01 * i * columns + 0.1 * j);
    vector<float, cpu_tag> x(columns);
    for (std::size_t j = 0; j != columns; ++j) x(j) = float(0.1 * j + 0.1);
    auto da = copy_to_device(a, hip_tag());
    auto dat = copy_to_device(at, hip_tag());
    auto dx = copy_to_device(x, hip_tag());
    auto check = [&](vector<float, hip_tag> const& dr, double init) {
        auto r = copy_to_cpu(dr);
        for (std::size_t i = 0; i != rows; ++i) {
            double s = init;
            for (std::size_t k = 0; k != columns; ++k) s += -2.0 * double(a(i, k)) * double(x(k));
            CHECK_CLOSE_PCT(r(i), s, 1e-2);
        }
    };
    for (int plus = 1; plus >= 0; --plus) {
        vector<float, hip_tag> r1(rows, 1.5f), r2(rows, 1.5f);
        if (plus) {
            noalias(r1) += -2 * prod(da, dx);
            noalias(r2) += -2 * prod(dx, dat);
        } else {
            noalias(r1) = -2 * prod(da, dx);
            noalias(r2) = -2 * prod(dx, dat);
        }
        check(r1, plus ? 1.5 : 0.0);
        check(r2, plus ? 1.5 : 0.0);
    }
    vector<float, hip_tag> r3 = prod(da, dx);
    auto h3 = copy_to_cpu(r3);
    for (std::size_t i = 0; i != rows; ++i) {
        double s = 0;
        for (std::size_t k = 0; k != columns; ++k) s += double(a(i, k)) * double(x(k));
        CHECK_CLOSE_PCT(h3(i), s, 1e-2);
    }
}

template <class O>
void trsm_test() {
    // Test/hip_triangular_solve.cpp: lower A of size 139, B 139 x 238 (float)
    const std::size_t n = 139, k = 238;
    matrix<float, row_major, cpu_tag> A(n, n, 1.0f);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j <= i; ++j) A(i, j) = float(0.1 / n * i - 0.05 / (i + 1.0) * j + (i == j ? 1.0 : 0.0));
    matrix<float, O, cpu_tag> B(n, k);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) B(i, j) = float((0.1 / n) * i + 0.1 * j / k);
    auto dA = copy_to_device(A, hip_tag());
    {   // left lower: A X = B
        auto dB = copy_to_device(B, hip_tag());
        kernels::trsm<lower, left>(dA, dB);
        auto X = copy_to_cpu(dB);
        double err = 0;
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != k; ++j) {
                double s = 0;
                for (std::size_t t = 0; t <= i; ++t) s += double(A(i, t)) * double(X(t, j));
                err = std::fmax(err, std::fabs(s - B(i, j)));
            }
        CHECK_SMALL(err, 1e-5);
    }
    {   // left upper via the transposed view: A^T X = B
        auto dB = copy_to_device(B, hip_tag());
        kernels::trsm<upper, left>(trans(dA), dB);
        auto X = copy_to_cpu(dB);
        double err = 0;
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != k; ++j) {
                double s = 0;
                for (std::size_t t = i; t < n; ++t) s += double(A(t, i)) * double(X(t, j));
                err = std::fmax(err, std::fabs(s - B(i, j)));
            }
        CHECK_SMALL(err, 1e-5);
    }
    {   // right lower: X A = C with C k x n
        matrix<float, O, cpu_tag> Cm(k, n);
        for (std::size_t i = 0; i != k; ++i)
            for (std::size_t j = 0; j != n; ++j) Cm(i, j) = B(j, i);
        auto dC = copy_to_device(Cm, hip_tag());
        kernels::trsm<lower, right>(dA, dC);
        auto X = copy_to_cpu(dC);
        double err = 0;
        for (std::size_t i = 0; i != k; ++i)
            for (std::size_t j = 0; j != n; ++j) {
                double s = 0;
                for (std::size_t t = j; t < n; ++t) s += double(X(i, t)) * double(A(t, j));
                err = std::fmax(err, std::fabs(s - Cm(i, j)));
            }
        CHECK_SMALL(err, 1e-5);
    }
    {   // trsv, lower, left
        matrix<float, row_major, cpu_tag> b(n, 1);
        for (std::size_t i = 0; i != n; ++i) b(i, 0) = float((0.1 / n) * i);
        vector<float, hip_tag> dv(n);
        hip::check_hip(hipMemcpy(dv.raw_storage().values, b.data(), n * sizeof(float), hipMemcpyHostToDevice));
        kernels::trsv<lower, left>(dA, dv);
        hip::get_stream(dv.queue()).synchronize();
        std::vector<float> x(n);
        hip::check_hip(hipMemcpy(x.data(), dv.raw_storage().values, n * sizeof(float), hipMemcpyDeviceToHost));
        double err = 0;
        for (std::size_t i = 0; i != n; ++i) {
            double s = 0;
            for (std::size_t t = 0; t <= i; ++t) s += double(A(i, t)) * x[t];
            err = std::fmax(err, std::fabs(s - b(i, 0)));
        }
        CHECK_SMALL(err, 1e-5);
    }
}

template <class O>
void symm_solve_test() {
    const std::size_t n = 128, k = 151;
    auto A = create_symm<row_major>(n, false);
    matrix<double, O, cpu_tag> B(n, k), Bright(k, n);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) Bright(j, i) = B(i, j) = 0.1 * j + (1.0 / n) * i;
    auto dA = copy_to_device(A, hip_tag());
    auto dB = copy_to_device(B, hip_tag());
    auto dBr = copy_to_device(Bright, hip_tag());
    matrix<double, O, hip_tag> dX = solve(dA, dB, symm_pos_def(), left());
    matrix<double, O, hip_tag> dXp = prod(inv(dA, symm_pos_def()), dB);
    matrix<double, O, hip_tag> dXr = solve(dA, dBr, symm_pos_def(), right());
    auto X = copy_to_cpu(dX), Xp = copy_to_cpu(dXp), Xr = copy_to_cpu(dXr);
    double err = 0, errr = 0, same = 0;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) {
            double s = 0, sr = 0;
            for (std::size_t t = 0; t != n; ++t) {
                s += A(i, t) * X(t, j);
                sr += Xr(j, t) * A(t, i);
            }
            err = std::fmax(err, std::fabs(s - B(i, j)));
            errr = std::fmax(errr, std::fabs(sr - Bright(j, i)));
            same = std::fmax(same, std::fabs(X(i, j) - Xp(i, j)));
        }
    CHECK_SMALL(err, 1e-12);
    CHECK_SMALL(errr, 1e-12);
    CHECK_SMALL(same, 1e-15);
    // cholesky_decomposition object, vector solve
    cholesky_decomposition<matrix<double, row_major, hip_tag>> dec(dA);
    vector<double, hip_tag> db(n);
    std::vector<double> hb(n), hx(n);
    for (std::size_t i = 0; i != n; ++i) hb[i] = (1.0 / n) * i;
    hip::check_hip(hipMemcpy(db.raw_storage().values, hb.data(), n * sizeof(double), hipMemcpyHostToDevice));
    dec.solve(db, left());
    hip::get_stream(db.queue()).synchronize();
    hip::check_hip(hipMemcpy(hx.data(), db.raw_storage().values, n * sizeof(double), hipMemcpyDeviceToHost));
    double verr = 0;
    for (std::size_t i = 0; i != n; ++i) {
        double s = 0;
        for (std::size_t t = 0; t != n; ++t) s += A(i, t) * hx[t];
        verr = std::fmax(verr, std::fabs(s - hb[i]));
    }
    CHECK_SMALL(verr, 1e-12);
}

// Test/cholesky.cpp:146-167: n = 123, v_i = 0.1/123*i - 0.5, alpha 0.1,
// beta 0.5; dec.update(alpha, beta, v) equals the factor of the updated
// matrix to 1e-11 percent.  Then a downdate that makes the matrix indefinite
// throws std::invalid_argument.
template <class O>
static void update_test() {
    const std::size_t n = 123;
    auto A = create_symm<O>(n, false);
    matrix<double, O, cpu_tag> Au(n, n);
    std::vector<double> v(n);
    for (std::size_t i = 0; i != n; ++i) v[i] = 0.1 / n * i - 0.5;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != n; ++j) Au(i, j) = 0.1 * A(i, j) + 0.5 * v[i] * v[j];
    auto dA = copy_to_device(A, hip_tag());
    auto dAu = copy_to_device(Au, hip_tag());
    vector<double, hip_tag> dv(n);
    hip::check_hip(hipMemcpy(dv.raw_storage().values, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
    cholesky_decomposition<matrix<double, O, hip_tag>> dec(dA), decUpdate(dAu);
    dec.update(0.1, 0.5, dv);
    auto L = copy_to_cpu(dec.lower_factor()), R = copy_to_cpu(decUpdate.lower_factor());
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j <= i; ++j) CHECK_CLOSE_PCT(L(i, j), R(i, j), 1e-11);
    bool threw = false;
    try {
        dec.update(1.0, -1.0e6, dv);
    } catch (std::invalid_argument const&) {
        threw = true;
    }
    ++g_checks;
    if (!threw) {
        std::printf("  FAIL %s:%d indefinite downdate did not throw\n", __FILE__, __LINE__);
        ++g_fail;
    }
    auto L2 = copy_to_cpu(dec.lower_factor());
    double same = 0;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != n; ++j) same = std::fmax(same, std::fabs(L2(i, j) - L(i, j)));
    CHECK_SMALL(same, 0.0);
}

// kernels::pstrf<lower|upper> on the device equals the oracle's restatement of
// the reference (kernels/default/pstrf.hpp:40-160) on createSymm(123): same
// rank and transpositions, factor to 1e-12 percent (Test/cholesky.cpp:84-111).
struct index_vector {
    std::vector<int64_t> v;
    explicit index_vector(std::size_t n) : v(n) {}
    int64_t& operator()(std::size_t i) { return v[i]; }
};

template <class O>
static void pstrf_test() {
    const std::size_t n = 123;
    auto A = create_symm<O>(n, false);
    const bool rm = O::is_row_major;
    for (int upper = 0; upper < 2; ++upper) {
        auto dA = copy_to_device(A, hip_tag());
        index_vector P(n);
        std::size_t r = upper ? kernels::pstrf<remora::upper>(dA, P) : kernels::pstrf<remora::lower>(dA, P);
        auto F = copy_to_cpu(dA);
        auto ref = A;
        std::vector<int64_t> Pr(n);
        for (std::size_t i = 0; i != n; ++i) Pr[i] = (int64_t)i;
        std::size_t rr = orc_dpstrf(ref.data(), n, n, rm, upper, Pr.data());
        CHECK_SMALL(double(r) - double(rr), 0.0);
        CHECK_SMALL(double(r) - double(n), 0.0);
        for (std::size_t i = 0; i != n; ++i) CHECK_SMALL(double(P.v[i] - Pr[i]), 0.0);
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != n; ++j) CHECK_CLOSE_PCT(F(i, j), ref(i, j), 1e-12);
    }
}

// solve(A, B, symm_semi_pos_def(), left|right) (Test/symm_solve.cpp:190-230):
// full-rank createSymm(128), B(i,j) = 0.1 j + i/128 with 151 columns: A X = B
// to 1e-12 and solve == prod(inv(A, symm_semi_pos_def()), B).  Then a rank-50
// consistent system A = R R^T, B = A Y, whose least-squares solution
// satisfies A X = B.
template <class O>
static void semi_solve_test() {
    const std::size_t n = 128, k = 151;
    auto A = create_symm<row_major>(n, false);
    matrix<double, O, cpu_tag> B(n, k), Bright(k, n);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) Bright(j, i) = B(i, j) = 0.1 * j + (1.0 / n) * i;
    auto dA = copy_to_device(A, hip_tag());
    auto dB = copy_to_device(B, hip_tag());
    auto dBr = copy_to_device(Bright, hip_tag());
    matrix<double, O, hip_tag> dX = solve(dA, dB, symm_semi_pos_def(), left());
    matrix<double, O, hip_tag> dXp = prod(inv(dA, symm_semi_pos_def()), dB);
    matrix<double, O, hip_tag> dXr = solve(dA, dBr, symm_semi_pos_def(), right());
    auto X = copy_to_cpu(dX), Xp = copy_to_cpu(dXp), Xr = copy_to_cpu(dXr);
    double err = 0, errr = 0, same = 0;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) {
            double s = 0, sr = 0;
            for (std::size_t t = 0; t != n; ++t) {
                s += A(i, t) * X(t, j);
                sr += Xr(j, t) * A(t, i);
            }
            err = std::fmax(err, std::fabs(s - B(i, j)));
            errr = std::fmax(errr, std::fabs(sr - Bright(j, i)));
            same = std::fmax(same, std::fabs(X(i, j) - Xp(i, j)));
        }
    CHECK_SMALL(err, 1e-12);
    CHECK_SMALL(errr, 1e-12);
    CHECK_SMALL(same, 1e-15);
    // rank 50, A stored in orientation O: the pivoted factor runs on
    // pstrf(lower) of either storage.  B = A Y is consistent; the solver
    // returns X = P L (L^T L)^-2 L^T P^T B = A^+ B, so A X = B and the
    // columns of X lie in range(A) = range(R) (they are orthogonal to A's
    // null space: any X + z with A z = 0 would pass the first check only).
    // The same on the right side, X A = B^T.
    const std::size_t r = 50, m = 3;
    matrix<double, O, cpu_tag> Ad(n, n);
    matrix<double, O, cpu_tag> Bd(n, m), BdT(m, n);
    matrix<double, row_major, cpu_tag> Y(n, m);
    std::vector<double> R(n * r);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t t = 0; t != r; ++t) {  // pseudo-random entries in [-0.5, 0.5): rank r
            const double h = std::sin(12.9898 * i + 78.233 * t + 0.1) * 43758.5453;
            R[i * r + t] = h - std::floor(h) - 0.5;
        }
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != n; ++j) {
            double s = 0;
            for (std::size_t t = 0; t != r; ++t) s += R[i * r + t] * R[j * r + t];
            Ad(i, j) = s;
        }
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != m; ++j) Y(i, j) = std::cos(0.11 * i + 0.7 * j);
    double bmax = 0;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != m; ++j) {
            double s = 0;
            for (std::size_t t = 0; t != n; ++t) s += Ad(i, t) * Y(t, j);
            BdT(j, i) = Bd(i, j) = s;
            bmax = std::fmax(bmax, std::fabs(s));
        }
    // host Cholesky of R^T R (r x r), for the projection onto range(R)
    std::vector<double> G(r * r);
    for (std::size_t a = 0; a != r; ++a)
        for (std::size_t b2 = 0; b2 != r; ++b2) {
            double s = 0;
            for (std::size_t i = 0; i != n; ++i) s += R[i * r + a] * R[i * r + b2];
            G[a * r + b2] = s;
        }
    for (std::size_t j = 0; j != r; ++j) {
        for (std::size_t t = 0; t != j; ++t) G[j * r + j] -= G[j * r + t] * G[j * r + t];
        G[j * r + j] = std::sqrt(G[j * r + j]);
        for (std::size_t i = j + 1; i != r; ++i) {
            for (std::size_t t = 0; t != j; ++t) G[i * r + j] -= G[i * r + t] * G[j * r + t];
            G[i * r + j] /= G[j * r + j];
        }
    }
    // || x - R (R^T R)^-1 R^T x || / ||x||
    auto off_range = [&](std::vector<double> const& x) {
        std::vector<double> c(r);
        for (std::size_t a = 0; a != r; ++a) {
            double s = 0;
            for (std::size_t i = 0; i != n; ++i) s += R[i * r + a] * x[i];
            c[a] = s;
        }
        for (std::size_t a = 0; a != r; ++a) {
            for (std::size_t t = 0; t != a; ++t) c[a] -= G[a * r + t] * c[t];
            c[a] /= G[a * r + a];
        }
        for (std::size_t a = r; a-- > 0;) {
            for (std::size_t t = a + 1; t != r; ++t) c[a] -= G[t * r + a] * c[t];
            c[a] /= G[a * r + a];
        }
        double num = 0, den = 0;
        for (std::size_t i = 0; i != n; ++i) {
            double p = 0;
            for (std::size_t a = 0; a != r; ++a) p += R[i * r + a] * c[a];
            num += (x[i] - p) * (x[i] - p);
            den += x[i] * x[i];
        }
        return std::sqrt(num / den);
    };
    auto dAd = copy_to_device(Ad, hip_tag());
    auto dBd = copy_to_device(Bd, hip_tag());
    auto dBdT = copy_to_device(BdT, hip_tag());
    solver<matrix<double, O, hip_tag>, symm_semi_pos_def> alg(dAd);
    CHECK_SMALL(double(alg.rank()) - double(r), 0.0);
    matrix<double, O, hip_tag> dXd = solve(dAd, dBd, symm_semi_pos_def(), left());
    matrix<double, O, hip_tag> dXdr = solve(dAd, dBdT, symm_semi_pos_def(), right());
    auto Xd = copy_to_cpu(dXd), Xdr = copy_to_cpu(dXdr);
    double errd = 0, errdr = 0, offr = 0, sides = 0;
    for (std::size_t j = 0; j != m; ++j) {
        std::vector<double> col(n);
        for (std::size_t i = 0; i != n; ++i) {
            double s = 0, sr = 0;
            for (std::size_t t = 0; t != n; ++t) {
                s += Ad(i, t) * Xd(t, j);
                sr += Xdr(j, t) * Ad(t, i);
            }
            errd = std::fmax(errd, std::fabs(s - Bd(i, j)));
            errdr = std::fmax(errdr, std::fabs(sr - BdT(j, i)));
            col[i] = Xd(i, j);
            sides = std::fmax(sides, std::fabs(Xd(i, j) - Xdr(j, i)));
        }
        offr = std::fmax(offr, off_range(col));
    }
    CHECK_SMALL(errd / bmax, 1e-10);
    CHECK_SMALL(errdr / bmax, 1e-10);
    CHECK_SMALL(offr, 1e-10);
    CHECK_SMALL(sides, 1e-12);
}

// copy_to_device / copy_to_cpu round trip exact, above the staging threshold
// with a ragged last chunk (3000 x 3001 doubles = 72 MB: two 64 MB chunks),
// and below it.
template <class O>
static void transport_test() {
    for (std::size_t n : {3000, 300}) {
        matrix<double, O, cpu_tag> A(n, n + 1);
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != n + 1; ++j) A(i, j) = 0.25 * i - 1.0 / (j + 1);
        auto dA = copy_to_device(A, hip_tag());
        auto B = copy_to_cpu(dA);
        double diff = 0;
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != n + 1; ++j) diff = std::fmax(diff, std::fabs(B(i, j) - A(i, j)));
        CHECK_SMALL(diff, 0.0);
    }
}

// Every method of hip::blas_device (the cublas_backend.hpp:121-431 method
// set, over the rmh_* ABI), called directly as the reference's glue calls it,
// in float and double, against host arithmetic: gemm, gemv, trmm, trmv,
// trsm, trsv, syrk.  Column-major host operands with padded leading
// dimensions; small integer-valued entries where the result is exact.
template <class T>
static void blas_device_test() {
    hip::device& d = hip::device_manager::instance().default_device();
    hip::blas_device& blas = hip::get_blas(d);
    hip::stream& s = hip::get_stream(d);
    const double tol = sizeof(T) == 8 ? 1e-13 : 1e-5;
    const std::size_t m = 37, n = 29, k = 23, ld = 41;
    auto val = [](std::size_t i, std::size_t j, int seed) { return T(int((i * 7 + j * 13 + seed * 5) % 9) - 4); };
    auto upload = [&](std::vector<T> const& h) {
        T* p = nullptr;
        hip::check_hip(hipMalloc(&p, h.size() * sizeof(T)), "hipMalloc");
        hip::check_hip(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        return p;
    };
    auto download = [&](T* p, std::vector<T>& h) {
        s.synchronize();
        hip::check_hip(hipMemcpy(h.data(), p, h.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        hip::check_hip(hipFree(p), "hipFree");
    };
    // A (ld x 64 storage), B, C
    std::vector<T> A(ld * 64), B(ld * 64), C(ld * 64);
    for (std::size_t j = 0; j != 64; ++j)
        for (std::size_t i = 0; i != ld; ++i) {
            A[i + j * ld] = val(i, j, 1);
            B[i + j * ld] = val(i, j, 2);
            C[i + j * ld] = val(i, j, 3);
        }
    auto at = [&](std::vector<T> const& M, std::size_t i, std::size_t j) { return double(M[i + j * ld]); };
    // gemm: C = 2 op(A) op(B) - C, all four transposes
    for (int ta = 0; ta < 2; ++ta)
        for (int tb = 0; tb < 2; ++tb) {
            T *dA = upload(A), *dB = upload(B), *dC = upload(C);
            blas.gemm(ta, tb, m, n, k, T(2), dA, ld, dB, ld, T(-1), dC, ld, s);
            std::vector<T> R(C.size());
            download(dC, R);
            hip::check_hip(hipFree(dA), "hipFree");
            hip::check_hip(hipFree(dB), "hipFree");
            double err = 0;
            for (std::size_t i = 0; i != m; ++i)
                for (std::size_t j = 0; j != n; ++j) {
                    double acc = -at(C, i, j);
                    for (std::size_t t = 0; t != k; ++t)
                        acc += 2 * (ta ? at(A, t, i) : at(A, i, t)) * (tb ? at(B, j, t) : at(B, t, j));
                    err = std::fmax(err, std::fabs(double(R[i + j * ld]) - acc));
                }
            CHECK_SMALL(err, 0.0);
        }
    // gemv: v = -1 op(A) x + 2 v, strides 2 and 3
    for (int ta = 0; ta < 2; ++ta) {
        std::vector<T> x(3 * 64), v(3 * 64);
        for (std::size_t i = 0; i != x.size(); ++i) {
            x[i] = val(i, 0, 4);
            v[i] = val(i, 1, 5);
        }
        T *dA = upload(A), *dx = upload(x), *dv = upload(v);
        blas.gemv(ta, m, n, T(-1), dA, ld, dx, 2, T(2), dv, 3, s);
        std::vector<T> R(v.size());
        download(dv, R);
        hip::check_hip(hipFree(dA), "hipFree");
        hip::check_hip(hipFree(dx), "hipFree");
        const std::size_t len = ta ? n : m, inner = ta ? m : n;
        double err = 0;
        for (std::size_t i = 0; i != len; ++i) {
            double acc = 2 * double(v[3 * i]);
            for (std::size_t t = 0; t != inner; ++t) acc -= (ta ? at(A, t, i) : at(A, i, t)) * double(x[2 * t]);
            err = std::fmax(err, std::fabs(double(R[3 * i]) - acc));
        }
        CHECK_SMALL(err, 0.0);
    }
    // triangular operand: op(T) of A's upper or lower triangle, optionally unit
    auto tri = [&](int up, int unit, std::size_t i, std::size_t j) {
        if (i == j) return unit ? 1.0 : at(A, i, i);
        return (up ? i < j : i > j) ? at(A, i, j) : 0.0;
    };
    // well-conditioned triangle for the solves: off-diagonal a/64 (at most
    // 1/16, so the solution grows by at most (1 + 1/16)^37 < 10 even with a
    // unit diagonal), diagonal 10 + |a|
    std::vector<T> As(A.size());
    for (std::size_t e = 0; e != A.size(); ++e) As[e] = A[e] / T(64);
    for (std::size_t i = 0; i != ld; ++i) As[i + i * ld] = T(10 + std::fabs(double(A[i + i * ld])));
    auto tris = [&](int up, int unit, std::size_t i, std::size_t j) {
        if (i == j) return unit ? 1.0 : double(As[i + i * ld]);
        return (up ? i < j : i > j) ? double(As[i + j * ld]) : 0.0;
    };
    for (int left = 0; left < 2; ++left)
        for (int up = 0; up < 2; ++up)
            for (int tr = 0; tr < 2; ++tr)
                for (int unit = 0; unit < 2; ++unit) {
                    const std::size_t na = left ? m : n;
                    auto op = [&](auto f, std::size_t i, std::size_t j) { return tr ? f(up, unit, j, i) : f(up, unit, i, j); };
                    // trmm: B = 3 op(T) B (left) or 3 B op(T)
                    {
                        T *dA = upload(A), *dB = upload(B);
                        blas.trmm(left, up, tr, unit, m, n, T(3), dA, ld, dB, ld, s);
                        std::vector<T> R(B.size());
                        download(dB, R);
                        hip::check_hip(hipFree(dA), "hipFree");
                        double err = 0;
                        for (std::size_t i = 0; i != m; ++i)
                            for (std::size_t j = 0; j != n; ++j) {
                                double acc = 0;
                                if (left)
                                    for (std::size_t t = 0; t != na; ++t) acc += op(tri, i, t) * at(B, t, j);
                                else
                                    for (std::size_t t = 0; t != na; ++t) acc += at(B, i, t) * op(tri, t, j);
                                err = std::fmax(err, std::fabs(double(R[i + j * ld]) - 3 * acc));
                            }
                        CHECK_SMALL(err, 0.0);
                    }
                    // trsm: op(T) X = 0.5 B (left) or X op(T) = 0.5 B: residual
                    {
                        T *dA = upload(As), *dB = upload(B);
                        blas.trsm(left, up, tr, unit, m, n, T(0.5), dA, ld, dB, ld, s);
                        std::vector<T> X(B.size());
                        download(dB, X);
                        hip::check_hip(hipFree(dA), "hipFree");
                        double err = 0;
                        for (std::size_t i = 0; i != m; ++i)
                            for (std::size_t j = 0; j != n; ++j) {
                                double acc = 0;
                                if (left)
                                    for (std::size_t t = 0; t != na; ++t) acc += op(tris, i, t) * double(X[t + j * ld]);
                                else
                                    for (std::size_t t = 0; t != na; ++t) acc += double(X[i + t * ld]) * op(tris, t, j);
                                err = std::fmax(err, std::fabs(acc - 0.5 * at(B, i, j)));
                            }
                        CHECK_SMALL(err, 100 * tol);
                    }
                    if (!left) continue;
                    // trmv / trsv on a vector of stride 2, order m
                    std::vector<T> x(2 * m);
                    for (std::size_t i = 0; i != x.size(); ++i) x[i] = val(i, 3, 6);
                    {
                        T *dA = upload(A), *dx = upload(x);
                        blas.trmv(up, tr, unit, m, dA, ld, dx, 2, s);
                        std::vector<T> R(x.size());
                        download(dx, R);
                        hip::check_hip(hipFree(dA), "hipFree");
                        double err = 0;
                        for (std::size_t i = 0; i != m; ++i) {
                            double acc = 0;
                            for (std::size_t t = 0; t != m; ++t) acc += op(tri, i, t) * double(x[2 * t]);
                            err = std::fmax(err, std::fabs(double(R[2 * i]) - acc));
                        }
                        CHECK_SMALL(err, 0.0);
                    }
                    {
                        T *dA = upload(As), *dx = upload(x);
                        blas.trsv(up, tr, unit, m, dA, ld, dx, 2, s);
                        std::vector<T> X(x.size());
                        download(dx, X);
                        hip::check_hip(hipFree(dA), "hipFree");
                        double err = 0;
                        for (std::size_t i = 0; i != m; ++i) {
                            double acc = 0;
                            for (std::size_t t = 0; t != m; ++t) acc += op(tris, i, t) * double(X[2 * t]);
                            err = std::fmax(err, std::fabs(acc - double(x[2 * i])));
                        }
                        CHECK_SMALL(err, 100 * tol);
                    }
                }
    // syrk: C = 2 op(A) op(A)^T - C on one triangle; the other untouched
    for (int up = 0; up < 2; ++up)
        for (int tr = 0; tr < 2; ++tr) {
            T *dA = upload(A), *dC = upload(C);
            blas.syrk(up, tr, n, k, T(2), dA, ld, T(-1), dC, ld, s);
            std::vector<T> R(C.size());
            download(dC, R);
            hip::check_hip(hipFree(dA), "hipFree");
            double err = 0;
            for (std::size_t i = 0; i != n; ++i)
                for (std::size_t j = 0; j != n; ++j) {
                    double want = at(C, i, j);
                    if (up ? i <= j : i >= j) {
                        want = -want;
                        for (std::size_t t = 0; t != k; ++t)
                            want += 2 * (tr ? at(A, t, i) * at(A, t, j) : at(A, i, t) * at(A, j, t));
                    }
                    err = std::fmax(err, std::fabs(double(R[i + j * ld]) - want));
                }
            CHECK_SMALL(err, 0.0);
        }
}

// kernels::trmm<Upper, Unit> / kernels::trmv on both orientations: B <- A B and
// v <- A v against host arithmetic (exact on small integers).
template <class OA, class OB>
static void trmm_trmv_test() {
    const std::size_t n = 45, k = 17;
    matrix<double, OA, cpu_tag> a(n, n);
    matrix<double, OB, cpu_tag> b(n, k);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != n; ++j) a(i, j) = double(int((i * 5 + j * 3) % 7) - 3);
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != k; ++j) b(i, j) = double(int((i * 2 + j * 11) % 5) - 2);
    auto da = copy_to_device(a, hip_tag());
    for (int up = 0; up < 2; ++up) {
        auto db = copy_to_device(b, hip_tag());
        if (up) kernels::trmm<true, false>(da, db);
        else kernels::trmm<false, true>(da, db);
        auto r = copy_to_cpu(db);
        double err = 0;
        for (std::size_t i = 0; i != n; ++i)
            for (std::size_t j = 0; j != k; ++j) {
                double acc = 0;
                for (std::size_t t = 0; t != n; ++t) {
                    double aij = (t == i) ? (up ? a(i, i) : 1.0) : ((up ? t > i : t < i) ? a(i, t) : 0.0);
                    acc += aij * b(t, j);
                }
                err = std::fmax(err, std::fabs(r(i, j) - acc));
            }
        CHECK_SMALL(err, 0.0);
        vector<double, cpu_tag> x(n);
        for (std::size_t i = 0; i != n; ++i) x(i) = double(int(i % 7) - 3);
        auto dx = copy_to_device(x, hip_tag());
        if (up) kernels::trmv<true, false>(da, dx);
        else kernels::trmv<false, false>(da, dx);
        auto rx = copy_to_cpu(dx);
        double errv = 0;
        for (std::size_t i = 0; i != n; ++i) {
            double acc = 0;
            for (std::size_t t = 0; t != n; ++t)
                if (up ? t >= i : t <= i) acc += a(i, t) * x(t);
            errv = std::fmax(errv, std::fabs(rx(i) - acc));
        }
        CHECK_SMALL(errv, 0.0);
    }
}

// solve(A, b, symm_pos_def(), left|right) for a vector (Test/symm_solve.cpp:
// 89-139 with one right-hand side): A x = b to 1e-12, solve == prod(inv(A), b),
// and both sides give the same x (A symmetric).
template <class O>
static void vector_solve_test() {
    const std::size_t n = 128;
    auto A = create_symm<O>(n, false);
    vector<double, cpu_tag> b(n);
    for (std::size_t i = 0; i != n; ++i) b(i) = 0.1 * 17 + (1.0 / n) * i;
    auto dA = copy_to_device(A, hip_tag());
    auto db = copy_to_device(b, hip_tag());
    vector<double, hip_tag> dx = solve(dA, db, symm_pos_def(), left());
    vector<double, hip_tag> dxr = solve(dA, db, symm_pos_def(), right());
    vector<double, hip_tag> dxp = prod(inv(dA, symm_pos_def()), db);
    auto x = copy_to_cpu(dx), xr = copy_to_cpu(dxr), xp = copy_to_cpu(dxp);
    double err = 0, same = 0;
    for (std::size_t i = 0; i != n; ++i) {
        double s = 0;
        for (std::size_t t = 0; t != n; ++t) s += A(i, t) * x(t);
        err = std::fmax(err, std::fabs(s - b(i)));
        same = std::fmax(same, std::fmax(std::fabs(x(i) - xp(i)), std::fabs(x(i) - xr(i))));
    }
    CHECK_SMALL(err, 1e-12);
    CHECK_SMALL(same, 0.0);
    // the residual as a device expression: r = b - A x
    vector<double, hip_tag> dr = db;
    noalias(dr) -= prod(dA, dx);
    auto r = copy_to_cpu(dr);
    double rmax = 0;
    for (std::size_t i = 0; i != n; ++i) rmax = std::fmax(rmax, std::fabs(r(i)));
    CHECK_SMALL(rmax, 1e-12);
}

// Assignment between orientations reads only the view's elements: an offset
// subrange that ends at the last row and column of its matrix (so the rest of
// its last line is past the allocation) assigned into a matrix of the other
// orientation, and the reverse; exact.
template <class O>
static void assign_subrange_test() {
    typedef typename O::transposed_orientation OT;
    const std::size_t n1 = 300, n2 = 211;
    matrix<double, O, cpu_tag> a(n1, n2);
    for (std::size_t i = 0; i != n1; ++i)
        for (std::size_t j = 0; j != n2; ++j) a(i, j) = i * 1000.0 + j;
    auto da = copy_to_device(a, hip_tag());
    auto view = subrange(da, 17, n1, 5, n2);
    matrix<double, OT, hip_tag> db(view);
    matrix<double, O, hip_tag> dc(trans(db));  // back, transposed
    auto b = copy_to_cpu(db);
    auto c = copy_to_cpu(dc);
    double err = 0;
    for (std::size_t i = 0; i != n1 - 17; ++i)
        for (std::size_t j = 0; j != n2 - 5; ++j) {
            err = std::fmax(err, std::fabs(b(i, j) - a(i + 17, j + 5)));
            err = std::fmax(err, std::fabs(c(j, i) - a(i + 17, j + 5)));
        }
    CHECK_SMALL(err, 0.0);
}

static void group(const char* name, void (*fn)()) {
    int before = g_fail, cbefore = g_checks;
    try {
        fn();
    } catch (std::exception const& e) {
        std::printf("  EXCEPTION %s\n", e.what());
        ++g_fail;
    }
    std::printf("%-32s %s (%d checks)\n", name, g_fail == before ? "ok" : "FAILED", g_checks - cbefore);
}

int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    group("semi-definite solve row_major", semi_solve_test<row_major>);
    group("semi-definite solve column_major", semi_solve_test<column_major>);
    group("pstrf row_major", pstrf_test<row_major>);
    group("pstrf column_major", pstrf_test<column_major>);
    group("transport row_major", transport_test<row_major>);
    group("transport column_major", transport_test<column_major>);
    group("update row_major", update_test<row_major>);
    group("update column_major", update_test<column_major>);
    group("potrf row_major", potrf_test<row_major>);
    group("potrf column_major", potrf_test<column_major>);
    group("syrk row_major", syrk_test<row_major>);
    group("syrk column_major", syrk_test<column_major>);
    group("prod rrr", prod_test<row_major, row_major, row_major>);
    group("prod rrc", prod_test<row_major, row_major, column_major>);
    group("prod rcr", prod_test<row_major, column_major, row_major>);
    group("prod rcc", prod_test<row_major, column_major, column_major>);
    group("prod crr", prod_test<column_major, row_major, row_major>);
    group("prod crc", prod_test<column_major, row_major, column_major>);
    group("prod ccr", prod_test<column_major, column_major, row_major>);
    group("prod ccc", prod_test<column_major, column_major, column_major>);
    group("prod vector row_major", prod_vector_test<row_major>);
    group("prod vector column_major", prod_vector_test<column_major>);
    group("blas_device double", blas_device_test<double>);
    group("blas_device float", blas_device_test<float>);
    group("trmm/trmv rr", trmm_trmv_test<row_major, row_major>);
    group("trmm/trmv rc", trmm_trmv_test<row_major, column_major>);
    group("trmm/trmv cr", trmm_trmv_test<column_major, row_major>);
    group("trmm/trmv cc", trmm_trmv_test<column_major, column_major>);
    group("vector solve row_major", vector_solve_test<row_major>);
    group("vector solve column_major", vector_solve_test<column_major>);
    group("assign subrange row_major", assign_subrange_test<row_major>);
    group("assign subrange column_major", assign_subrange_test<column_major>);
    group("trsm row_major B", trsm_test<row_major>);
    group("trsm column_major B", trsm_test<column_major>);
    group("symm solve row_major", symm_solve_test<row_major>);
    group("symm solve column_major", symm_solve_test<column_major>);
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
