/*
 * oracle/remora_cpu.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Remora's default (CPU) dense kernels on the
 * Cholesky path.  It is the checker for the HIP library and the CPU baseline
 * timed beside it; nothing in the product links or calls it.
 *
 * Followed, algorithm by algorithm (paths relative to the Remora tree):
 *   potrf_block (3 leaf variants) ...... include/remora/kernels/default/potrf.hpp:42-103
 *   potrf_recursive (block 32) ......... include/remora/kernels/default/potrf.hpp:106-145
 *   trsm_block<32,16> (4 variants) ..... include/remora/kernels/default/trsm.hpp:45-225
 *   trsm_recursive (block 32) .......... include/remora/kernels/default/trsm.hpp:227-275
 *   syrk_impl (MC=3*mr*nr, EC=1024) .... include/remora/kernels/default/syrk.hpp:42-130
 *   dense_gemm (mc128/kc512/nc1020) .... include/remora/kernels/default/dense_gemm.hpp:52-141
 *   gemm orientation dispatch .......... include/remora/kernels/gemm.hpp:47-106
 *   ugemm / mgemm / pack_A / pack_B .... include/remora/kernels/default/mgemm.hpp:50-187
 *   vector width (32 B with AVX) ....... include/remora/kernels/default/simd.hpp:49-72
 *   cholesky_decomposition::solve ...... include/remora/decompositions.hpp:101-115
 *   cholesky_decomposition::update ..... include/remora/decompositions.hpp:132-169
 *   pstrf (pivoted, block 20) .......... include/remora/kernels/default/pstrf.hpp:40-160
 *   trsv_impl (4 variants + right) ..... include/remora/kernels/default/trsv.hpp:44-183
 *   gemv_impl (row/column-major) ....... include/remora/kernels/default/gemv.hpp:43-93
 *   dot (dense) ........................ include/remora/kernels/default/dot.hpp:39-52
 *   cholesky_decomposition::solve(b) ... include/remora/decompositions.hpp:111-115
 *
 * Matrices are strided views: element (i,j) lives at p[i*s1 + j*s2] and the
 * view carries Remora's orientation tag (row_major: s2 is the unit stride).
 * trans() swaps sizes, strides and the tag, exactly as Remora's trans()
 * proxy does, so every orientation dispatch of the reference is reproduced.
 *
 * Build: gcc -O3 -march=x86-64-v3 -DREMORA_USE_SIMD (oracle/build.py).  GCC's
 * default -ffp-contract=fast contracts a*b+c into FMA as it does for the
 * reference compiled with the same compiler.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- views -- */
typedef struct {
    double *p;
    size_t m, n;      /* size1, size2 */
    ptrdiff_t s1, s2; /* strides */
    int row_major;    /* orientation tag */
} view;

static inline double *at(view v, size_t i, size_t j) {
    return v.p + (ptrdiff_t)i * v.s1 + (ptrdiff_t)j * v.s2;
}
static inline view trans(view v) {
    view t = {v.p, v.n, v.m, v.s2, v.s1, !v.row_major};
    return t;
}
static inline view subrange(view v, size_t r0, size_t r1, size_t c0, size_t c1) {
    view s = {at(v, r0, c0), r1 - r0, c1 - c0, v.s1, v.s2, v.row_major};
    return s;
}
static view make_view(double *p, size_t m, size_t n, size_t ld, int row_major) {
    view v;
    v.p = p; v.m = m; v.n = n; v.row_major = row_major;
    if (row_major) { v.s1 = (ptrdiff_t)ld; v.s2 = 1; }
    else { v.s1 = 1; v.s2 = (ptrdiff_t)ld; }
    return v;
}
static inline size_t min_sz(size_t a, size_t b) { return a < b ? a : b; }

/* error channel for the trsm singularity exception of the reference */
static int g_trsm_singular = 0;

/* ------------------------------------------------------------- simd.hpp -- */
#ifdef __AVX__
#define ORC_VECTOR_LENGTH 32
#else
#define ORC_VECTOR_LENGTH 16
#endif
#define ORC_MAX_VEC (ORC_VECTOR_LENGTH / 8) /* block<double>::max_vector_elements */
#ifdef REMORA_USE_SIMD
#define ORC_VEC ORC_MAX_VEC
typedef double vx __attribute__((vector_size(ORC_VECTOR_LENGTH)));
#else
#define ORC_VEC 1
typedef double vx;
#endif

/* block sizes (dense_gemm.hpp:52-60, syrk.hpp:42-49) */
#define GEMM_MR 4
#define GEMM_NR (3 * ORC_MAX_VEC)
#define GEMM_MC 128
#define GEMM_KC 512
#define GEMM_NC ((1024 / GEMM_NR) * GEMM_NR)
#define SYRK_MR 4
#define SYRK_NR (SYRK_MR * ORC_MAX_VEC)
#define SYRK_MC (3 * SYRK_MR * SYRK_NR)
#define SYRK_EC 1024
#define MAX_NR 32

static void *aligned_buf(size_t count) {
    void *p = NULL;
    size_t bytes = (count * sizeof(double) + 63) / 64 * 64;
    if (bytes == 0) bytes = 64;
    if (posix_memalign(&p, 64, bytes) != 0) return NULL;
    return p;
}

/* ----------------------------------------------------------- mgemm.hpp -- */
/* ugemm: MR x NR register tile as outer products (mgemm.hpp:50-99). */
static void ugemm(size_t mr, size_t nr, size_t kc, double alpha, const double *A,
                  const double *B, double *C, ptrdiff_t stride1, ptrdiff_t stride2) {
    const size_t vecNR = nr / ORC_VEC;
    vx P[4 * MAX_NR / ORC_VEC];
    memset(P, 0, sizeof(P));
    const vx *b = (const vx *)B;
#ifdef ORC_MUTANT_UGEMM_DROP_LAST_K
    if (kc > 1) kc -= 1; /* planted mistake: one term of the sum dropped */
#endif
    for (size_t l = 0; l < kc; ++l) {
        for (size_t i = 0; i < mr; ++i)
            for (size_t j = 0; j < vecNR; ++j)
                P[i * vecNR + j] += A[i] * b[j];
        A += mr;
        b += vecNR;
    }
    if (alpha != 1.0) {
        for (size_t i = 0; i < mr; ++i)
            for (size_t j = 0; j < vecNR; ++j)
                P[i * vecNR + j] *= alpha;
    }
    const double *p = (const double *)P;
    for (size_t i = 0; i < mr; ++i)
        for (size_t j = 0; j < nr; ++j)
            C[(ptrdiff_t)i * stride1 + (ptrdiff_t)j * stride2] += p[i * nr + j];
}

/* mgemm: macro kernel over packed blocks, partial tiles through a temporary
 * (mgemm.hpp:103-143). */
static void mgemm(size_t MR, size_t NR, size_t mc, size_t nc, size_t kc, double alpha,
                  const double *A, const double *B, double *C, ptrdiff_t stride1,
                  ptrdiff_t stride2) {
    const size_t mp = (mc + MR - 1) / MR;
    const size_t np = (nc + NR - 1) / NR;
    for (size_t j = 0; j < np; ++j) {
        const size_t nr = min_sz(NR, nc - j * NR);
        for (size_t i = 0; i < mp; ++i) {
            const size_t mr = min_sz(MR, mc - i * MR);
            double *Cb = C + (ptrdiff_t)(i * MR) * stride1 + (ptrdiff_t)(j * NR) * stride2;
            if (mr == MR && nr == NR) {
                ugemm(MR, NR, kc, alpha, &A[i * kc * MR], &B[j * kc * NR], Cb, stride1, stride2);
            } else {
                double tmp[4 * MAX_NR] __attribute__((aligned(64)));
                for (size_t t = 0; t < MR * NR; ++t) tmp[t] = 0.0;
                ugemm(MR, NR, kc, alpha, &A[i * kc * MR], &B[j * kc * NR], tmp, (ptrdiff_t)NR, 1);
                for (size_t i0 = 0; i0 < mr; ++i0)
                    for (size_t j0 = 0; j0 < nr; ++j0)
                        Cb[(ptrdiff_t)i0 * stride1 + (ptrdiff_t)j0 * stride2] += tmp[i0 * NR + j0];
            }
        }
    }
}

/* pack_A_dense: MR-interleaved row stripes, zero padded (mgemm.hpp:147-166) */
static void pack_A(view A, double *p, size_t MR) {
    const size_t mc = A.m, kc = A.n;
    const size_t mp = (mc + MR - 1) / MR;
    size_t nu = 0;
    for (size_t l = 0; l < mp; ++l)
        for (size_t j = 0; j < kc; ++j)
            for (size_t i = l * MR; i < l * MR + MR; ++i, ++nu)
                p[nu] = (i < mc) ? *at(A, i, j) : 0.0;
}

/* pack_B_dense: NR-interleaved column stripes, zero padded (mgemm.hpp:169-187) */
static void pack_B(view B, double *p, size_t NR) {
    const size_t kc = B.m, nc = B.n;
    const size_t np = (nc + NR - 1) / NR;
    size_t nu = 0;
    for (size_t l = 0; l < np; ++l)
        for (size_t i = 0; i < kc; ++i)
            for (size_t j = l * NR; j < l * NR + NR; ++j, ++nu)
                p[nu] = (j < nc) ? *at(B, i, j) : 0.0;
}

/* ------------------------------------------------------ dense_gemm.hpp -- */
/* m += alpha * e1 * e2, m row-major (dense_gemm.hpp:83-141) */
static void dense_gemm(view e1, view e2, view m, double alpha) {
    const size_t M = m.m, N = m.n, K = e1.n;
    const size_t mb = (M + GEMM_MC - 1) / GEMM_MC;
    const size_t nb = (N + GEMM_NC - 1) / GEMM_NC;
    const size_t kb = (K + GEMM_KC - 1) / GEMM_KC;
    double *A = aligned_buf((size_t)GEMM_MC * GEMM_KC);
    double *B = aligned_buf((size_t)GEMM_NC * GEMM_KC);
    const ptrdiff_t ldc = m.s1;
    for (size_t j = 0; j < nb; ++j) {
        const size_t nc = min_sz(GEMM_NC, N - j * GEMM_NC);
        for (size_t l = 0; l < kb; ++l) {
            const size_t kc = min_sz(GEMM_KC, K - l * GEMM_KC);
            pack_B(subrange(e2, l * GEMM_KC, l * GEMM_KC + kc, j * GEMM_NC, j * GEMM_NC + nc), B, GEMM_NR);
            for (size_t i = 0; i < mb; ++i) {
                const size_t mc = min_sz(GEMM_MC, M - i * GEMM_MC);
                pack_A(subrange(e1, i * GEMM_MC, i * GEMM_MC + mc, l * GEMM_KC, l * GEMM_KC + kc), A, GEMM_MR);
                mgemm(GEMM_MR, GEMM_NR, mc, nc, kc, alpha, A, B,
                      m.p + (ptrdiff_t)(i * GEMM_MC) * ldc + (ptrdiff_t)(j * GEMM_NC), ldc, 1);
            }
        }
    }
    free(A);
    free(B);
}

/* kernels::gemm dispatch: a column-major target is transposed to row-major
 * via m^T += e2^T e1^T (gemm.hpp:47-72). */
static void kernels_gemm(view e1, view e2, view m, double alpha) {
    if (m.m == 0 || m.n == 0) return;
    if (m.row_major) dense_gemm(e1, e2, m, alpha);
    else dense_gemm(trans(e2), trans(e1), trans(m), alpha);
}

/* ------------------------------------------------------------ syrk.hpp -- */
/* m += alpha * e * e^T on one triangle (syrk.hpp:50-130) */
static void syrk_impl(view e, view m, double alpha, int upper) {
    const size_t MC = SYRK_MC, EC = SYRK_EC;
    double *E_left = aligned_buf(MC * EC);
    double *E_right = aligned_buf(MC * EC);
    double *Md = aligned_buf(MC * MC);
    const size_t M = e.m, K = e.n;
    const size_t Mb = (M + MC - 1) / MC;
    const size_t Eb = (K + EC - 1) / EC;
    const ptrdiff_t stride1 = m.s1, stride2 = m.s2;
    for (size_t k = 0; k < Eb; ++k) {
        const size_t kc = min_sz(EC, K - k * EC);
        for (size_t i = 0; i < Mb; ++i) {
            const size_t mc = min_sz(MC, M - i * MC);
            pack_A(subrange(e, i * MC, i * MC + mc, k * EC, k * EC + kc), E_left, SYRK_MR);
            const size_t start_j = upper ? i : 0;
            const size_t end_j = upper ? Mb : i + 1;
            for (size_t j = start_j; j < end_j; ++j) {
                const size_t mc2 = min_sz(MC, M - j * MC);
                pack_B(trans(subrange(e, j * MC, j * MC + mc2, k * EC, k * EC + kc)), E_right, SYRK_NR);
                if (i == j) {
                    for (size_t i0 = 0; i0 != mc; ++i0)
                        for (size_t j0 = 0; j0 != mc2; ++j0) Md[i0 * MC + j0] = 0.0;
                    mgemm(SYRK_MR, SYRK_NR, mc, mc2, kc, alpha, E_left, E_right, Md, (ptrdiff_t)MC, 1);
                    double *Mdiag = m.p + (ptrdiff_t)(i * MC) * stride1 + (ptrdiff_t)(j * MC) * stride2;
                    for (size_t i0 = 0; i0 != mc; ++i0) {
#ifdef ORC_MUTANT_SYRK_WRONG_TRIANGLE
                        const size_t s = !upper ? i0 : 0; /* planted: other triangle */
                        const size_t e0 = !upper ? mc2 : i0 + 1;
#else
                        const size_t s = upper ? i0 : 0;
                        const size_t e0 = upper ? mc2 : i0 + 1;
#endif
                        for (size_t j0 = s; j0 < e0; ++j0)
                            Mdiag[(ptrdiff_t)i0 * stride1 + (ptrdiff_t)j0 * stride2] += Md[i0 * MC + j0];
                    }
                } else {
                    mgemm(SYRK_MR, SYRK_NR, mc, mc2, kc, alpha, E_left, E_right,
                          m.p + (ptrdiff_t)(i * MC) * stride1 + (ptrdiff_t)(j * MC) * stride2, stride1, stride2);
                }
            }
        }
    }
    free(E_left);
    free(E_right);
    free(Md);
}

/* ------------------------------------------------------------ trsm.hpp -- */
#define TRSM_B1 32
#define TRSM_B2 16

/* trsm_block<32,16,Unit> for the four (lower|upper) x (row|column-major B)
 * variants (trsm.hpp:45-225). */
static void trsm_block(view A, view B, int upper, int unit) {
    const size_t size = A.m;
    double blockA[TRSM_B1][TRSM_B1];
    for (size_t i = 0; i != size; ++i) {
#ifdef ORC_MUTANT_TRSM_TRANSPOSED_A
        /* planted mistake: the triangle is read transposed */
        if (upper) { for (size_t j = i; j != size; ++j) blockA[i][j] = *at(A, j, i); }
        else { for (size_t j = 0; j <= i; ++j) blockA[i][j] = *at(A, j, i); }
#else
        if (upper) { for (size_t j = i; j != size; ++j) blockA[i][j] = *at(A, i, j); }
        else { for (size_t j = 0; j <= i; ++j) blockA[i][j] = *at(A, i, j); }
#endif
    }
    if (B.row_major) {
        double blockB[TRSM_B2][TRSM_B1];
        const size_t numBlocks = (B.n + TRSM_B2 - 1) / TRSM_B2;
        for (size_t bi = 0; bi != numBlocks; ++bi) {
            const size_t startB = bi * TRSM_B2;
            const size_t cur = min_sz(TRSM_B2, B.n - startB);
            for (size_t i = 0; i != size; ++i)
                for (size_t k = 0; k != cur; ++k) blockB[k][i] = *at(B, i, startB + k);
            for (size_t k = 0; k != cur; ++k) {
                if (!upper) {
                    for (size_t i = 0; i != size; ++i) {
                        for (size_t j = 0; j != i; ++j) blockB[k][i] -= blockA[i][j] * blockB[k][j];
                        if (!unit) {
                            if (blockA[i][i] == 0.0) { g_trsm_singular = 1; return; }
                            blockB[k][i] /= blockA[i][i];
                        }
                    }
                } else {
                    for (size_t nn = 0; nn != size; ++nn) {
                        const size_t i = size - nn - 1;
                        for (size_t j = i + 1; j != size; ++j) blockB[k][i] -= blockA[i][j] * blockB[k][j];
                        if (!unit) {
                            if (blockA[i][i] == 0.0) { g_trsm_singular = 1; return; }
                            blockB[k][i] /= blockA[i][i];
                        }
                    }
                }
            }
            for (size_t i = 0; i != size; ++i)
                for (size_t k = 0; k != cur; ++k) *at(B, i, startB + k) = blockB[k][i];
        }
    } else {
        for (size_t k = 0; k != B.n; ++k) {
            if (!upper) {
                for (size_t i = 0; i != size; ++i) {
                    for (size_t j = 0; j != i; ++j) *at(B, i, k) -= blockA[i][j] * *at(B, j, k);
                    if (!unit) {
                        if (blockA[i][i] == 0.0) { g_trsm_singular = 1; return; }
                        *at(B, i, k) /= blockA[i][i];
                    }
                }
            } else {
                for (size_t nn = 0; nn != size; ++nn) {
                    const size_t i = size - nn - 1;
                    for (size_t j = i + 1; j != size; ++j) *at(B, i, k) -= blockA[i][j] * *at(B, j, k);
                    if (!unit) {
                        if (blockA[i][i] == 0.0) { g_trsm_singular = 1; return; }
                        *at(B, i, k) /= blockA[i][i];
                    }
                }
            }
        }
    }
}

/* trsm_recursive, left side (trsm.hpp:227-261) */
static void trsm_recursive_left(view Afull, view Bfull, size_t start, size_t end, int upper, int unit) {
    const size_t num_rhs = Bfull.n;
    view A = subrange(Afull, start, end, start, end);
    view B = subrange(Bfull, start, end, 0, num_rhs);
    if (A.m <= TRSM_B1) {
        trsm_block(A, B, upper, unit);
        return;
    }
    const size_t size = A.m;
    const size_t numBlocks = (A.m + TRSM_B1 - 1) / TRSM_B1;
    const size_t split = numBlocks / 2 * TRSM_B1;
    view Bfront = subrange(B, 0, split, 0, num_rhs);
    view Bback = subrange(B, split, size, 0, num_rhs);
    if (upper) {
        trsm_recursive_left(Afull, Bfull, start + split, end, upper, unit);
        if (g_trsm_singular) return;
        kernels_gemm(subrange(A, 0, split, split, size), Bback, Bfront, -1.0);
        trsm_recursive_left(Afull, Bfull, start, start + split, upper, unit);
    } else {
        trsm_recursive_left(Afull, Bfull, start, start + split, upper, unit);
        if (g_trsm_singular) return;
        kernels_gemm(subrange(A, split, size, 0, split), Bfront, Bback, -1.0);
        trsm_recursive_left(Afull, Bfull, start + split, end, upper, unit);
    }
}

/* kernels::trsm<Triangular,Side>: the right side transposes both operands
 * and flips the triangle (trsm.hpp:263-286). */
static void kernels_trsm(view A, view B, int upper, int unit, int left) {
    if (left) trsm_recursive_left(A, B, 0, A.m, upper, unit);
    else trsm_recursive_left(trans(A), trans(B), 0, A.m, !upper, unit);
}

/* ----------------------------------------------------------- potrf.hpp -- */
/* row-major lower leaf: left-looking jik with a double accumulator
 * (potrf.hpp:42-64) */
static size_t potrf_block_rm_lower(view A) {
    const size_t m = A.m;
    for (size_t j = 0; j < m; j++) {
        for (size_t i = j; i < m; i++) {
            double s = *at(A, i, j);
#ifdef ORC_MUTANT_POTRF_SIGN
            for (size_t k = 0; k < j; k++) s += *at(A, i, k) * *at(A, j, k); /* planted sign error */
#else
            for (size_t k = 0; k < j; k++) s -= *at(A, i, k) * *at(A, j, k);
#endif
            if (i == j) {
                if (s <= 0) return i + 1;
                *at(A, i, j) = sqrt(s);
            } else {
                *at(A, i, j) = s / *at(A, j, j);
            }
        }
    }
    return 0;
}

/* row-major upper leaf: right-looking rank-1 updates; only a negative pivot
 * is reported (potrf.hpp:67-92) */
static size_t potrf_block_rm_upper(view A) {
    const size_t m = A.m;
    for (size_t i = 0; i < m; i++) {
        double *Aii = at(A, i, i);
        if (*Aii < 0) return i + 1;
        *Aii = sqrt(*Aii);
        for (size_t j = i + 1; j < m; ++j) *at(A, i, j) /= *Aii;
        for (size_t k = i + 1; k < m; k++)
            for (size_t j = k; j < m; ++j) *at(A, k, j) -= *at(A, i, k) * *at(A, i, j);
    }
    return 0;
}

/* leaf dispatcher: column-major goes through the transposed view
 * (potrf.hpp:94-103) */
static size_t potrf_block(view A, int upper) {
    if (A.row_major) return upper ? potrf_block_rm_upper(A) : potrf_block_rm_lower(A);
    return potrf_block(trans(A), !upper);
}

/* potrf_recursive, lower (potrf.hpp:106-134) */
static size_t potrf_recursive_lower(view Afull, size_t start, size_t end) {
    const size_t block_size = 32;
    view A = subrange(Afull, start, end, start, end);
    const size_t size = A.m;
    if (size <= block_size) return potrf_block(A, 0);
    const size_t numBlocks = (A.m + block_size - 1) / block_size;
    const size_t split = numBlocks / 2 * block_size;
    size_t result = potrf_recursive_lower(Afull, start, start + split);
    if (result) return result;
    view Aul = subrange(A, 0, split, 0, split);
    view All = subrange(A, split, size, 0, split);
    view Alr = subrange(A, split, size, split, size);
    kernels_trsm(trans(Aul), All, /*upper=*/1, /*unit=*/0, /*left=*/0);
    syrk_impl(All, Alr, -1.0, /*upper=*/0);
    return potrf_recursive_lower(Afull, start + split, end);
}

/* As in the reference, a failing pivot is reported relative to the 32-wide
 * leaf that found it (potrf.hpp:55-56 return i+1 of the leaf view and the
 * recursion passes it up unchanged, :125-133): nonzero means "not positive
 * definite", the value is not a global column index. */
static size_t kernels_potrf(view A, int upper) {
    if (upper) A = trans(A); /* potrf.hpp:136-145 */
    return potrf_recursive_lower(A, 0, A.m);
}

/* ========================================================== C exports == */
/* All matrices are dense with a leading dimension `ld`; `row_major` selects
 * Remora's orientation tag.  Returns are LAPACK-style info for potrf and 0/1
 * (1 = singular triangular matrix, the reference's invalid_argument) for
 * trsm-based routines. */

size_t orc_dpotrf(double *A, size_t n, size_t lda, int row_major, int upper) {
    return kernels_potrf(make_view(A, n, n, lda, row_major), upper);
}

int orc_dtrsm(const double *A, size_t n, size_t lda, int a_row_major, double *B, size_t m1, size_t m2,
              size_t ldb, int b_row_major, int upper, int unit, int left) {
    g_trsm_singular = 0;
    kernels_trsm(make_view((double *)A, n, n, lda, a_row_major), make_view(B, m1, m2, ldb, b_row_major),
                 upper, unit, left);
    return g_trsm_singular;
}

void orc_dsyrk(const double *E, size_t n, size_t k, size_t lde, int e_row_major, double *M, size_t ldm,
               int m_row_major, double alpha, int upper) {
    syrk_impl(make_view((double *)E, n, k, lde, e_row_major), make_view(M, n, n, ldm, m_row_major), alpha,
              upper);
}

/* M(m x n) += alpha * E1(m x k) * E2(k x n) */
void orc_dgemm(const double *E1, size_t lde1, int e1_row_major, const double *E2, size_t lde2,
               int e2_row_major, double *M, size_t ldm, int m_row_major, size_t m, size_t n, size_t k,
               double alpha) {
    kernels_gemm(make_view((double *)E1, m, k, lde1, e1_row_major),
                 make_view((double *)E2, k, n, lde2, e2_row_major), make_view(M, m, n, ldm, m_row_major),
                 alpha);
}

/* cholesky_decomposition(A) followed by solve(B, left|right)
 * (decompositions.hpp:81-110).  A is overwritten by its lower factor (the
 * strict upper triangle keeps A's values), B by the solution. */
int orc_dposv(double *A, size_t n, size_t lda, int a_row_major, double *B, size_t m1, size_t m2, size_t ldb,
              int b_row_major, int left, size_t *info) {
    view L = make_view(A, n, n, lda, a_row_major);
    view Bv = make_view(B, m1, m2, ldb, b_row_major);
    *info = kernels_potrf(L, 0);
    g_trsm_singular = 0;
    if (left) {
        kernels_trsm(L, Bv, 0, 0, 1);
        if (!g_trsm_singular) kernels_trsm(trans(L), Bv, 1, 0, 1);
    } else {
        kernels_trsm(trans(L), Bv, 1, 0, 0);
        if (!g_trsm_singular) kernels_trsm(L, Bv, 0, 0, 0);
    }
    return g_trsm_singular;
}

/* cholesky_decomposition::update(alpha, beta, v) (decompositions.hpp:132-169):
 * the lower factor L of A becomes that of alpha*A + beta*v*v^T, column by
 * column, in the reference's order of operations.  beta == 0 scales the
 * whole stored matrix by sqrt(alpha), as the reference does.  Returns 0, or
 * j+1 when column j would make the matrix indefinite (the reference throws
 * std::invalid_argument there, after updating columns 0..j-1). */
size_t orc_dchol_update(double *Lp, size_t n, size_t ldl, int row_major, double alpha, double beta,
                        const double *v, size_t incv) {
    view L = make_view(Lp, n, n, ldl, row_major);
    if (beta == 0) {
        const double s = sqrt(alpha);
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j) *at(L, i, j) *= s;
        return 0;
    }
    double *temp = (double *)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; ++i) temp[i] = v[i * incv];
    double beta_prime = 1;
    const double a = sqrt(alpha);
    size_t info = 0;
    for (size_t j = 0; j != n; ++j) {
        const double Ljj = a * *at(L, j, j);
        const double dj = Ljj * Ljj;
        const double wj = temp[j];
        const double swj2 = beta * wj * wj;
        const double gamma = dj * beta_prime + swj2;
        const double x = dj + swj2 / beta_prime;
        if (x <= 0.0) {
            info = j + 1;
            break;
        }
        const double nLjj = sqrt(x);
        *at(L, j, j) = nLjj;
        beta_prime += swj2 / dj;
        if (j + 1 < n) {
#ifdef ORC_MUTANT_UPDATE_UNSCALED_TEMP
            for (size_t i = j + 1; i < n; ++i) temp[i] -= (wj / Ljj) * *at(L, i, j);
            for (size_t i = j + 1; i < n; ++i) *at(L, i, j) *= a;
#else
            for (size_t i = j + 1; i < n; ++i) *at(L, i, j) *= a;
            for (size_t i = j + 1; i < n; ++i) temp[i] -= (wj / Ljj) * *at(L, i, j);
#endif
            if (gamma == 0) continue;
            for (size_t i = j + 1; i < n; ++i) *at(L, i, j) *= nLjj / Ljj;
            const double f = nLjj * beta * wj / gamma;
            for (size_t i = j + 1; i < n; ++i) *at(L, i, j) += f * temp[i];
        }
    }
    free(temp);
    return info;
}

/* kernels::pstrf<lower>(A, P) (kernels/default/pstrf.hpp:40-148; upper is
 * the same on trans(A), :150-158): pivoted Cholesky in blocks of 20 columns.
 * Per column the pivot is the first largest running diagonal value; rows and
 * then columns of the whole matrix are swapped; the column takes the block's
 * deferred updates (a GEMV) and is scaled; the row right of the diagonal is
 * cleared.  After each block the whole trailing square takes -= L21 L21^T.
 * Stops when the pivot drops below n^2 * eps * max|A_ii|, clearing the
 * trailing square.  P (identity on entry, as permutation_matrix(n) is) gets
 * P[j] = pivot row for every swap.  Returns the rank. */
size_t orc_dpstrf(double *Ap, size_t m, size_t lda, int row_major, int upper, int64_t *P) {
    view A = make_view(Ap, m, m, lda, row_major);
    if (upper) A = trans(A);
    const size_t block_size = 20;
    double *pv = (double *)malloc((m ? m : 1) * sizeof(double));
    double max_diag = m ? *at(A, 0, 0) : 0.0;
    for (size_t i = 1; i < m; ++i) max_diag = fmax(max_diag, fabs(*at(A, i, i)));
    const double epsilon = (double)m * (double)m * 2.220446049250313e-16 * max_diag;
    for (size_t k = 0; k < m; k += block_size) {
        const size_t cur = min_sz(m - k, block_size);
        view Ak = subrange(A, k, m, k, m);
        double *piv = pv + k;
        const size_t mk = m - k;
        for (size_t j = 0; j != cur; ++j) {
            if (j == 0) {
                for (size_t i = 0; i != mk; ++i) piv[i] = *at(Ak, i, i);
            } else {
                for (size_t i = j; i != mk; ++i) piv[i] -= *at(Ak, i, j - 1) * *at(Ak, i, j - 1);
            }
            size_t pivot = j;
#ifdef ORC_MUTANT_PSTRF_LAST_MAX
            for (size_t i = j + 1; i < mk; ++i)
                if (piv[i] >= piv[pivot]) pivot = i;
#else
            for (size_t i = j + 1; i < mk; ++i)
                if (piv[pivot] < piv[i]) pivot = i;
#endif
            if (pivot != j) {
                P[k + j] = (int64_t)(pivot + k);
                for (size_t t = 0; t < m; ++t) {
                    double x = *at(A, k + j, t);
                    *at(A, k + j, t) = *at(A, k + pivot, t);
                    *at(A, k + pivot, t) = x;
                }
                for (size_t t = 0; t < m; ++t) {
                    double x = *at(A, t, k + j);
                    *at(A, t, k + j) = *at(A, t, k + pivot);
                    *at(A, t, k + pivot) = x;
                }
                double x = piv[j];
                piv[j] = piv[pivot];
                piv[pivot] = x;
            }
            const double pivotValue = piv[j];
            if (pivotValue < epsilon) {
                for (size_t r = j; r < mk; ++r)
                    for (size_t c = j; c < mk; ++c) *at(Ak, r, c) = 0.0;
                free(pv);
                return k + j;
            }
            *at(Ak, j, j) = sqrt(pivotValue);
            if (j > 0) {
                for (size_t r = j + 1; r < mk; ++r) {
                    double acc = 0.0;
                    for (size_t t = 0; t < j; ++t) acc += *at(Ak, r, t) * *at(Ak, j, t);
                    *at(Ak, r, j) -= acc;
                }
            }
            for (size_t r = j + 1; r < mk; ++r) *at(Ak, r, j) /= *at(Ak, j, j);
            for (size_t c = j + 1; c < mk; ++c) *at(Ak, j, c) = 0.0;
        }
        if (k + cur < m) {
            view LL = subrange(Ak, block_size, mk, 0, block_size);
            view LR = subrange(Ak, block_size, mk, block_size, mk);
            kernels_gemm(LL, trans(LL), LR, -1.0);
        }
    }
    free(pv);
    return m;
}

/* ============================================================ trsv.hpp == */
/* A dense vector of Remora is a strided array: element i at x[i*inc]. */
static int g_trsv_singular = 0;

/* dot(v1, v2) of default/dot.hpp:39-52: result starts at 0 and takes
 * v1_k * v2_k in increasing k. */
static double dot_strided(const double *v1, ptrdiff_t inc1, const double *v2, ptrdiff_t inc2, size_t len) {
    double r = 0.0;
    for (size_t k = 0; k < len; ++k) r += v1[(ptrdiff_t)k * inc1] * v2[(ptrdiff_t)k * inc2];
    return r;
}

/* lower, column-major storage: b_n /= A_nn, then the column below takes
 * b[n+1:] = b[n+1:] + (-b_n) * A[n+1:, n], skipped when b_n == 0
 * (default/trsv.hpp:44-68; multiply_and_add is x + scalar * y,
 * cpu/traits.hpp:131-140). */
static void trsv_lower_cm(view A, double *b, ptrdiff_t inc, int unit) {
    const size_t size = A.m;
    for (size_t n = 0; n != size; ++n) {
        if (!unit) {
            if (*at(A, n, n) == 0.0) { g_trsv_singular = 1; return; }
            b[(ptrdiff_t)n * inc] /= *at(A, n, n);
        }
        const double bn = b[(ptrdiff_t)n * inc];
        if (bn != 0.0) {
            const double s = -bn;
            for (size_t i = n + 1; i < size; ++i) b[(ptrdiff_t)i * inc] = b[(ptrdiff_t)i * inc] + s * *at(A, i, n);
        }
    }
}

/* lower, row-major storage: b_n -= dot(b[0:n], A[n, 0:n]), then / A_nn
 * (default/trsv.hpp:70-96). */
static void trsv_lower_rm(view A, double *b, ptrdiff_t inc, int unit) {
    const size_t size = A.m;
    for (size_t n = 0; n < size; ++n) {
#ifdef ORC_MUTANT_TRSV_SKIP_FIRST
        double value = n > 0 ? dot_strided(b + inc, inc, at(A, n, 1), A.s2, n - 1) : 0.0; /* planted: drops k = 0 */
#else
        double value = dot_strided(b, inc, at(A, n, 0), A.s2, n);
#endif
        b[(ptrdiff_t)n * inc] -= value;
        if (!unit) {
            if (*at(A, n, n) == 0.0) { g_trsv_singular = 1; return; }
            b[(ptrdiff_t)n * inc] /= *at(A, n, n);
        }
    }
}

/* upper, column-major storage: back to front, b_n /= A_nn, then
 * b[0:n] = b[0:n] + (-b_n) * A[0:n, n], skipped when b_n == 0
 * (default/trsv.hpp:98-124). */
static void trsv_upper_cm(view A, double *b, ptrdiff_t inc, int unit) {
    const size_t size = A.m;
    for (size_t i = 0; i < size; ++i) {
        const size_t n = size - i - 1;
        if (!unit) {
            if (*at(A, n, n) == 0.0) { g_trsv_singular = 1; return; }
            b[(ptrdiff_t)n * inc] /= *at(A, n, n);
        }
        const double bn = b[(ptrdiff_t)n * inc];
        if (bn != 0.0) {
            const double s = -bn;
            for (size_t r = 0; r < n; ++r) b[(ptrdiff_t)r * inc] = b[(ptrdiff_t)r * inc] + s * *at(A, r, n);
        }
    }
}

/* upper, row-major storage: back to front, b_n -= dot(b[n+1:], A[n, n+1:]),
 * then / A_nn (default/trsv.hpp:125-151). */
static void trsv_upper_rm(view A, double *b, ptrdiff_t inc, int unit) {
    const size_t size = A.m;
    for (size_t i = 0; i < size; ++i) {
        const size_t n = size - i - 1;
        double value = dot_strided(b + (ptrdiff_t)(n + 1) * inc, inc, at(A, n, n + 1), A.s2, size - n - 1);
        b[(ptrdiff_t)n * inc] -= value;
        if (!unit) {
            if (*at(A, n, n) == 0.0) { g_trsv_singular = 1; return; }
            b[(ptrdiff_t)n * inc] /= *at(A, n, n);
        }
    }
}

/* trsv<Triangular, Side>(A, b) (kernels/trsv.hpp:53-63 -> default/trsv.hpp:154-183):
 * dispatch on the orientation tag; the right side solves b A = b as
 * A^T x = b through trans(A), which flips the triangle and the tag. */
static void kernels_trsv(view A, double *b, ptrdiff_t inc, int upper, int unit, int left) {
    if (!left) {
        A = trans(A);
        upper = !upper;
    }
    if (upper) {
        if (A.row_major) trsv_upper_rm(A, b, inc, unit);
        else trsv_upper_cm(A, b, inc, unit);
    } else {
        if (A.row_major) trsv_lower_rm(A, b, inc, unit);
        else trsv_lower_cm(A, b, inc, unit);
    }
}

/* ============================================================ gemv.hpp == */
/* result += alpha * A * x (kernels/gemv.hpp:58-72 -> default/gemv.hpp:43-93):
 *   row-major A:    value = dot(row(A, i), x); if value != 0,
 *                   result_i += alpha * value;
 *   column-major A: for each j, result = result + (alpha * x_j) * column(A, j). */
static void kernels_gemv(view A, const double *x, ptrdiff_t incx, double *y, ptrdiff_t incy, double alpha) {
    if (A.row_major) {
        for (size_t i = 0; i != A.m; ++i) {
            double value = dot_strided(at(A, i, 0), A.s2, x, incx, A.n);
            if (value != 0.0) y[(ptrdiff_t)i * incy] += alpha * value;
        }
    } else {
        for (size_t j = 0; j != A.n; ++j) {
#ifdef ORC_MUTANT_GEMV_NO_ALPHA
            const double s = x[(ptrdiff_t)j * incx]; /* planted: alpha dropped */
#else
            const double s = alpha * x[(ptrdiff_t)j * incx];
#endif
            for (size_t i = 0; i != A.m; ++i) y[(ptrdiff_t)i * incy] = y[(ptrdiff_t)i * incy] + s * *at(A, i, j);
        }
    }
}

/* Solve op-free A x = b (left) or x A = b (right) in place; A n x n
 * triangular.  Returns 1 where the reference throws "[TRSV] Matrix is
 * singular!" (a zero diagonal, non-unit only); b then holds the partial
 * result. */
int orc_dtrsv(const double *A, size_t n, size_t lda, int a_row_major, double *b, ptrdiff_t incb, int upper,
              int unit, int left) {
    g_trsv_singular = 0;
    if (n == 0) return 0;
    kernels_trsv(make_view((double *)A, n, n, lda, a_row_major), b, incb, upper, unit, left);
    return g_trsv_singular;
}

/* y(m) += alpha * A(m x n) * x(n) */
void orc_dgemv(const double *A, size_t m, size_t n, size_t lda, int a_row_major, const double *x, ptrdiff_t incx,
               double *y, ptrdiff_t incy, double alpha) {
    if (m == 0 || n == 0) return;
    kernels_gemv(make_view((double *)A, m, n, lda, a_row_major), x, incx, y, incy, alpha);
}

/* cholesky_decomposition(A) followed by solve(b) for a vector
 * (decompositions.hpp:81-91, 111-115): potrf<lower>, then
 * trsv<lower,left>(L) and trsv<upper,left>(trans(L)), for either side (A is
 * symmetric).  A is overwritten by its factor, b by the solution. */
int orc_dposv_vec(double *A, size_t n, size_t lda, int a_row_major, double *b, ptrdiff_t incb, size_t *info) {
    view L = make_view(A, n, n, lda, a_row_major);
    *info = kernels_potrf(L, 0);
    g_trsv_singular = 0;
    if (n == 0) return 0;
    kernels_trsv(L, b, incb, 0, 0, 1);
    if (!g_trsv_singular) kernels_trsv(trans(L), b, incb, 1, 0, 1);
    return g_trsv_singular;
}

/* The build's SIMD width, for tests that pin the block constants. */
int orc_vector_length(void) { return ORC_VECTOR_LENGTH; }
int orc_simd_enabled(void) {
#ifdef REMORA_USE_SIMD
    return 1;
#else
    return 0;
#endif
}
