// C++ front-end tests on hip_tag, restating the reference's HIP test
// programs against the header-only front-end:
//   Test/hip_cholesky.cpp:36-65          potrf<lower|upper> on row/column-major == CPU potrf
//   Test/hip_syrk.cpp:17-98              float syrk closed form
//   Test/hip_triangular_solve.cpp:20-193 float trsm / trsv residuals
//   Test/hip_prod.cpp:113-261            float gemm, all orientations
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
    auto dc = copy_to_device(c, hip_tag());
    kernels::gemm(da, db, dc, -2.0f);
    auto res = copy_to_cpu(dc);
    for (std::size_t i = 0; i != rows; ++i)
        for (std::size_t j = 0; j != cols; ++j) {
            double s = 1.5;
            for (std::size_t k = 0; k != middle; ++k) s += -2.0 * double(a(i, k)) * double(b(k, j));
            CHECK_CLOSE_PCT(res(i, j), s, 1e-2);
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
    // rank 50
    const std::size_t r = 50, m = 3;
    matrix<double, row_major, cpu_tag> Ad(n, n), Y(n, m);
    matrix<double, O, cpu_tag> Bd(n, m);
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
            Bd(i, j) = s;
            bmax = std::fmax(bmax, std::fabs(s));
        }
    auto dAd = copy_to_device(Ad, hip_tag());
    auto dBd = copy_to_device(Bd, hip_tag());
    solver<matrix<double, row_major, hip_tag>, symm_semi_pos_def> alg(dAd);
    CHECK_SMALL(double(alg.rank()) - double(r), 0.0);
    matrix<double, O, hip_tag> dXd = solve(dAd, dBd, symm_semi_pos_def(), left());
    auto Xd = copy_to_cpu(dXd);
    double errd = 0;
    for (std::size_t i = 0; i != n; ++i)
        for (std::size_t j = 0; j != m; ++j) {
            double s = 0;
            for (std::size_t t = 0; t != n; ++t) s += Ad(i, t) * Xd(t, j);
            errd = std::fmax(errd, std::fabs(s - Bd(i, j)));
        }
    CHECK_SMALL(errd / bmax, 1e-10);
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
    group("gemm rrr", prod_test<row_major, row_major, row_major>);
    group("gemm rcr", prod_test<row_major, column_major, row_major>);
    group("gemm crc", prod_test<column_major, row_major, column_major>);
    group("gemm ccc", prod_test<column_major, column_major, column_major>);
    group("gemm rrc", prod_test<row_major, row_major, column_major>);
    group("gemm ccr", prod_test<column_major, column_major, row_major>);
    group("trsm row_major B", trsm_test<row_major>);
    group("trsm column_major B", trsm_test<column_major>);
    group("symm solve row_major", symm_solve_test<row_major>);
    group("symm solve column_major", symm_solve_test<column_major>);
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
