/*
 * elem_oracle.c -- CPU restatement of Elemental's lower HPD Cholesky
 * factor-and-solve path.  TEST INFRASTRUCTURE ONLY: it is the checker that
 * the HIP library is compared against, and the CPU baseline that bench.py
 * times beside the GPU.  Nothing in the product library links, loads or
 * calls this file.
 *
 * What it restates (all file:line citations are into aterrel/Elemental):
 *   eo_lvar3unb        cholesky::LVar3Unb     include/elemental/lapack-like/Cholesky/LVar3.hpp:19-49
 *   eo_lvar3           cholesky::LVar3        include/elemental/lapack-like/Cholesky/LVar3.hpp:83-123
 *                      block clamp            include/elemental/core/partition_impl.hpp:403-424
 *   eo_solve_after     cholesky::SolveAfter   include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75
 *   eo_hpd_solve       HPDSolve               include/elemental/lapack-like/HPDSolve.hpp:18-28
 *   eo_uvar3unb        cholesky::UVar3Unb     include/elemental/lapack-like/Cholesky/UVar3.hpp:21-49
 *   eo_uvar3           cholesky::UVar3        include/elemental/lapack-like/Cholesky/UVar3.hpp:85-125
 *   eo_solve_after_upper  SolveAfter(UPPER)   include/elemental/lapack-like/Cholesky/SolveAfter.hpp:33-43,56-72
 *   eo_trrk_tn         internal::TrrkTN       src/blas-like/Trrk/Local.hpp:733-775 (leaf :452-490)
 *   eo_local_trrk_tn   LocalTrrk (1x1 grid)   src/blas-like/Trrk/Local.hpp:931-976 (leaf :492-536)
 *   eo_scale_trapezoid ScaleTrapezoid         include/elemental/blas-like/level1/ScaleTrapezoid.hpp:15-84
 *   eo_axpy_triangle   AxpyTriangle           include/elemental/blas-like/level1/AxpyTriangle.hpp:16-38
 *
 * The reference hands its level-3/level-2 arithmetic to an external,
 * unpinned Fortran BLAS (CMakeLists.txt:281-315; call sites
 * src/core/imports/blas.cpp:901-904 dtrsv, :930-940 dgemm, :1043-1050 dsyrk,
 * :1203-1210 dtrsm).  The survey container resolved it to Intel MKL 2021.4.
 * This file restates the published algorithm of that dependency: the Netlib
 * reference BLAS (the column-oriented loops of dtrsm/dsyrk/dgemm/dtrsv as
 * published with LAPACK 3.x), written here as eo_dtrsm/eo_dsyrk/eo_dgemm/
 * eo_dtrsv.  For the CPU baseline at full size the same blocked loop can
 * instead call an optimised host BLAS bound at run time (eo_bind_blas);
 * that is what the reference itself does.
 *
 * Build with -ffp-contract=off so every multiply and add rounds separately,
 * as in the Fortran reference loops.
 *
 * Parity pins (tests/test_oracle.py): integer-exact known answers
 * (A = L0*L0^T with integer L0 and power-of-two diagonal, and the Pascal
 * matrix whose factor is the binomial triangle), LAPACK dpotrf/dpotrs and
 * the host BLAS via scipy, the splitmix64 published output sequence, and
 * the invariants L*L^T = A, diag(L) > 0 and backward error.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

#define AT(A, i, j, ld) ((A)[(i64)(i) + (i64)(j) * (i64)(ld)])

/* ------------------------------------------------------------------ */
/* Portable seeded generator: splitmix64 (Steele, Lea, Flood 2014) as a
 * counter-based stream, and Box-Muller normals.  The same definitions are
 * used by the HIP library's generator and by the Python tests, so inputs
 * are identical on CPU and GPU up to libm rounding of log/cos.            */

uint64_t eo_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* i-th output of the stream started at `seed` (counter form). */
static uint64_t sm64_at(uint64_t seed, uint64_t i) {
    uint64_t s = seed + i * 0x9E3779B97F4A7C15ULL;
    return eo_splitmix64(&s);
}

/* uniform in (0,1]: 53 high bits, offset by one ulp so log() is finite */
static double u01(uint64_t x) { return ((double)(x >> 11) + 1.0) * 0x1.0p-53; }

/* out[t] for t in [0,count): standard normal number (offset + t) of the
 * stream `seed`.  Pair p = (offset+t)/2 uses outputs 2p and 2p+1. */
void eo_fill_normal(double *out, i64 count, uint64_t seed, i64 offset) {
    const double two_pi = 6.283185307179586476925286766559;
    for (i64 t = 0; t < count; ++t) {
        i64 idx = offset + t;
        uint64_t p = (uint64_t)(idx >> 1);
        double u1 = u01(sm64_at(seed, 2 * p));
        double u2 = u01(sm64_at(seed, 2 * p + 1));
        double r = sqrt(-2.0 * log(u1));
        out[t] = (idx & 1) ? r * sin(two_pi * u2) : r * cos(two_pi * u2);
    }
}

/* ------------------------------------------------------------------ */
/* Netlib reference BLAS restatements (the published algorithms).       */

static int lsame(char a, char b) {
    if (a >= 'a' && a <= 'z') a = (char)(a - 32);
    if (b >= 'a' && b <= 'z') b = (char)(b - 32);
    return a == b;
}

/* C := alpha*op(A)*op(B) + beta*C, op in {N,T}. */
void eo_dgemm(char ta, char tb, int m, int n, int k, double alpha,
              const double *A, int lda, const double *B, int ldb,
              double beta, double *C, int ldc) {
    int nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    if (alpha == 0.0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i)
                AT(C, i, j, ldc) = (beta == 0.0) ? 0.0 : beta * AT(C, i, j, ldc);
        return;
    }
    if (notb) {
        if (nota) {
            for (int j = 0; j < n; ++j) {
                if (beta == 0.0) for (int i = 0; i < m; ++i) AT(C, i, j, ldc) = 0.0;
                else if (beta != 1.0) for (int i = 0; i < m; ++i) AT(C, i, j, ldc) = beta * AT(C, i, j, ldc);
                for (int l = 0; l < k; ++l) {
                    double temp = alpha * AT(B, l, j, ldb);
                    for (int i = 0; i < m; ++i) AT(C, i, j, ldc) += temp * AT(A, i, l, lda);
                }
            }
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < m; ++i) {
                    double temp = 0.0;
                    for (int l = 0; l < k; ++l) temp += AT(A, l, i, lda) * AT(B, l, j, ldb);
                    AT(C, i, j, ldc) = (beta == 0.0) ? alpha * temp : alpha * temp + beta * AT(C, i, j, ldc);
                }
        }
    } else {
        if (nota) {
            for (int j = 0; j < n; ++j) {
                if (beta == 0.0) for (int i = 0; i < m; ++i) AT(C, i, j, ldc) = 0.0;
                else if (beta != 1.0) for (int i = 0; i < m; ++i) AT(C, i, j, ldc) = beta * AT(C, i, j, ldc);
                for (int l = 0; l < k; ++l) {
                    double temp = alpha * AT(B, j, l, ldb);
                    for (int i = 0; i < m; ++i) AT(C, i, j, ldc) += temp * AT(A, i, l, lda);
                }
            }
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < m; ++i) {
                    double temp = 0.0;
                    for (int l = 0; l < k; ++l) temp += AT(A, l, i, lda) * AT(B, j, l, ldb);
                    AT(C, i, j, ldc) = (beta == 0.0) ? alpha * temp : alpha * temp + beta * AT(C, i, j, ldc);
                }
        }
    }
}

/* SYRK on one triangle of C (uplo 'L' or 'U'): C := alpha*A*A^T + beta*C
 * (trans N, A n x k) or C := alpha*A^T*A + beta*C (trans T, A k x n).  The
 * other triangle is untouched. */
void eo_dsyrk(char uplo, char trans, int n, int k, double alpha,
              const double *A, int lda, double beta, double *C, int ldc) {
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    if (lsame(uplo, 'U')) {  /* upper triangle (Netlib dsyrk's UPLO = 'U' loops) */
        if (lsame(trans, 'N')) {
            for (int j = 0; j < n; ++j) {
                if (beta == 0.0) for (int i = 0; i <= j; ++i) AT(C, i, j, ldc) = 0.0;
                else if (beta != 1.0) for (int i = 0; i <= j; ++i) AT(C, i, j, ldc) = beta * AT(C, i, j, ldc);
                if (alpha == 0.0) continue;
                for (int l = 0; l < k; ++l) {
                    if (AT(A, j, l, lda) != 0.0) {
                        double temp = alpha * AT(A, j, l, lda);
                        for (int i = 0; i <= j; ++i) AT(C, i, j, ldc) += temp * AT(A, i, l, lda);
                    }
                }
            }
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i <= j; ++i) {
                    double temp = 0.0;
                    for (int l = 0; l < k; ++l) temp += AT(A, l, i, lda) * AT(A, l, j, lda);
                    AT(C, i, j, ldc) = (beta == 0.0) ? alpha * temp : alpha * temp + beta * AT(C, i, j, ldc);
                }
        }
        return;
    }
    if (lsame(trans, 'N')) {
        for (int j = 0; j < n; ++j) {
            if (beta == 0.0) for (int i = j; i < n; ++i) AT(C, i, j, ldc) = 0.0;
            else if (beta != 1.0) for (int i = j; i < n; ++i) AT(C, i, j, ldc) = beta * AT(C, i, j, ldc);
            if (alpha == 0.0) continue;
            for (int l = 0; l < k; ++l) {
                if (AT(A, j, l, lda) != 0.0) {
                    double temp = alpha * AT(A, j, l, lda);
                    for (int i = j; i < n; ++i) AT(C, i, j, ldc) += temp * AT(A, i, l, lda);
                }
            }
        }
    } else {
        for (int j = 0; j < n; ++j)
            for (int i = j; i < n; ++i) {
                double temp = 0.0;
                for (int l = 0; l < k; ++l) temp += AT(A, l, i, lda) * AT(A, l, j, lda);
                AT(C, i, j, ldc) = (beta == 0.0) ? alpha * temp : alpha * temp + beta * AT(C, i, j, ldc);
            }
    }
}

/* Triangular TRSM (Netlib dtrsm loops), lower or upper, non-unit or unit
 * diagonal.
 *   side L, trans N : B := alpha*inv(A)*B
 *   side L, trans T : B := alpha*inv(A^T)*B
 *   side R, trans T : B := alpha*B*inv(A^T)
 *   side R, trans N : B := alpha*B*inv(A)
 * A is m x m (side L) or n x n (side R). */
static void dtrsm_upper(int left, int notrans, int nounit, int m, int n, double alpha,
                        const double *A, int lda, double *B, int ldb) {
    if (left) {
        if (notrans) { /* B := alpha*inv(A)*B */
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0) for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = alpha * AT(B, i, j, ldb);
                for (int k = m - 1; k >= 0; --k) {
                    if (AT(B, k, j, ldb) != 0.0) {
                        if (nounit) AT(B, k, j, ldb) = AT(B, k, j, ldb) / AT(A, k, k, lda);
                        for (int i = 0; i < k; ++i) AT(B, i, j, ldb) -= AT(B, k, j, ldb) * AT(A, i, k, lda);
                    }
                }
            }
        } else { /* B := alpha*inv(A^T)*B */
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < m; ++i) {
                    double temp = alpha * AT(B, i, j, ldb);
                    for (int k = 0; k < i; ++k) temp -= AT(A, k, i, lda) * AT(B, k, j, ldb);
                    if (nounit) temp = temp / AT(A, i, i, lda);
                    AT(B, i, j, ldb) = temp;
                }
        }
    } else {
        if (notrans) { /* B := alpha*B*inv(A) */
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0) for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = alpha * AT(B, i, j, ldb);
                for (int k = 0; k < j; ++k)
                    if (AT(A, k, j, lda) != 0.0)
                        for (int i = 0; i < m; ++i) AT(B, i, j, ldb) -= AT(A, k, j, lda) * AT(B, i, k, ldb);
                if (nounit) {
                    double temp = 1.0 / AT(A, j, j, lda);
                    for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = temp * AT(B, i, j, ldb);
                }
            }
        } else { /* B := alpha*B*inv(A^T) */
            for (int k = n - 1; k >= 0; --k) {
                if (nounit) {
                    double temp = 1.0 / AT(A, k, k, lda);
                    for (int i = 0; i < m; ++i) AT(B, i, k, ldb) = temp * AT(B, i, k, ldb);
                }
                for (int j = 0; j < k; ++j) {
                    if (AT(A, j, k, lda) != 0.0) {
                        double temp = AT(A, j, k, lda);
                        for (int i = 0; i < m; ++i) AT(B, i, j, ldb) -= temp * AT(B, i, k, ldb);
                    }
                }
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) AT(B, i, k, ldb) = alpha * AT(B, i, k, ldb);
            }
        }
    }
}

void eo_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
              double alpha, const double *A, int lda, double *B, int ldb) {
    int nounit = lsame(diag, 'N');
    if (m == 0 || n == 0) return;
    if (alpha == 0.0) {
        for (int j = 0; j < n; ++j) for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = 0.0;
        return;
    }
    if (lsame(uplo, 'U')) {
        dtrsm_upper(lsame(side, 'L'), lsame(transa, 'N'), nounit, m, n, alpha, A, lda, B, ldb);
        return;
    }
    if (lsame(side, 'L')) {
        if (lsame(transa, 'N')) {
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0) for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = alpha * AT(B, i, j, ldb);
                for (int k = 0; k < m; ++k) {
                    if (AT(B, k, j, ldb) != 0.0) {
                        if (nounit) AT(B, k, j, ldb) = AT(B, k, j, ldb) / AT(A, k, k, lda);
                        for (int i = k + 1; i < m; ++i)
                            AT(B, i, j, ldb) -= AT(B, k, j, ldb) * AT(A, i, k, lda);
                    }
                }
            }
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = m - 1; i >= 0; --i) {
                    double temp = alpha * AT(B, i, j, ldb);
                    for (int k = i + 1; k < m; ++k) temp -= AT(A, k, i, lda) * AT(B, k, j, ldb);
                    if (nounit) temp = temp / AT(A, i, i, lda);
                    AT(B, i, j, ldb) = temp;
                }
        }
    } else {
        if (lsame(transa, 'N')) {
            for (int j = n - 1; j >= 0; --j) {
                if (alpha != 1.0) for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = alpha * AT(B, i, j, ldb);
                for (int k = j + 1; k < n; ++k)
                    if (AT(A, k, j, lda) != 0.0)
                        for (int i = 0; i < m; ++i) AT(B, i, j, ldb) -= AT(A, k, j, lda) * AT(B, i, k, ldb);
                if (nounit) {
                    double temp = 1.0 / AT(A, j, j, lda);
                    for (int i = 0; i < m; ++i) AT(B, i, j, ldb) = temp * AT(B, i, j, ldb);
                }
            }
        } else {
            for (int k = 0; k < n; ++k) {
                if (nounit) {
                    double temp = 1.0 / AT(A, k, k, lda);
                    for (int i = 0; i < m; ++i) AT(B, i, k, ldb) = temp * AT(B, i, k, ldb);
                }
                for (int j = k + 1; j < n; ++j) {
                    if (AT(A, j, k, lda) != 0.0) {
                        double temp = AT(A, j, k, lda);
                        for (int i = 0; i < m; ++i) AT(B, i, j, ldb) -= temp * AT(B, i, k, ldb);
                    }
                }
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) AT(B, i, k, ldb) = alpha * AT(B, i, k, ldb);
            }
        }
    }
}

/* Triangular TRSV (Netlib dtrsv loops) with unit stride, lower or upper,
 * trans N or T. */
void eo_dtrsv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x) {
    int nounit = lsame(diag, 'N');
    if (lsame(uplo, 'U')) {
        if (lsame(trans, 'N')) {
            for (int j = n - 1; j >= 0; --j) {
                if (x[j] != 0.0) {
                    if (nounit) x[j] = x[j] / AT(A, j, j, lda);
                    double temp = x[j];
                    for (int i = j - 1; i >= 0; --i) x[i] -= temp * AT(A, i, j, lda);
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                double temp = x[j];
                for (int i = 0; i < j; ++i) temp -= AT(A, i, j, lda) * x[i];
                if (nounit) temp = temp / AT(A, j, j, lda);
                x[j] = temp;
            }
        }
        return;
    }
    if (lsame(trans, 'N')) {
        for (int j = 0; j < n; ++j) {
            if (x[j] != 0.0) {
                if (nounit) x[j] = x[j] / AT(A, j, j, lda);
                double temp = x[j];
                for (int i = j + 1; i < n; ++i) x[i] -= temp * AT(A, i, j, lda);
            }
        }
    } else {
        for (int j = n - 1; j >= 0; --j) {
            double temp = x[j];
            for (int i = n - 1; i >= j + 1; --i) temp -= AT(A, i, j, lda) * x[i];
            if (nounit) temp = temp / AT(A, j, j, lda);
            x[j] = temp;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Optional optimised host BLAS, bound at run time (CPU baseline).      */

typedef void (*fp_dtrsm)(const char *, const char *, const char *, const char *,
                         const int *, const int *, const double *, const double *,
                         const int *, double *, const int *);
typedef void (*fp_dsyrk)(const char *, const char *, const int *, const int *,
                         const double *, const double *, const int *,
                         const double *, double *, const int *);
typedef void (*fp_dtrsv)(const char *, const char *, const char *, const int *,
                         const double *, const int *, double *, const int *);
typedef void (*fp_setnt)(int);

static fp_dtrsm g_dtrsm;
static fp_dsyrk g_dsyrk;
static fp_dtrsv g_dtrsv;

/* Bind dtrsm/dsyrk/dtrsv from a shared BLAS whose Fortran symbols carry
 * `prefix` (e.g. "scipy_" for scipy's OpenBLAS, "" for MKL).  Returns 0 on
 * success.  `setnt_name` optionally names a set-num-threads function. */
int eo_bind_blas(const char *path, const char *prefix, const char *setnt_name, int nthreads) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    char name[128];
    snprintf(name, sizeof name, "%sdtrsm_", prefix);
    g_dtrsm = (fp_dtrsm)dlsym(h, name);
    snprintf(name, sizeof name, "%sdsyrk_", prefix);
    g_dsyrk = (fp_dsyrk)dlsym(h, name);
    snprintf(name, sizeof name, "%sdtrsv_", prefix);
    g_dtrsv = (fp_dtrsv)dlsym(h, name);
    if (!g_dtrsm || !g_dsyrk || !g_dtrsv) return 2;
    if (setnt_name && setnt_name[0]) {
        fp_setnt f = (fp_setnt)dlsym(h, setnt_name);
        if (!f) return 3;
        f(nthreads);
    }
    return 0;
}

int eo_blas_bound(void) { return g_dtrsm && g_dsyrk && g_dtrsv; }

static void x_dtrsm(int use_host, char side, char uplo, char ta, char diag, int m, int n,
                    double alpha, const double *A, int lda, double *B, int ldb) {
    if (use_host) g_dtrsm(&side, &uplo, &ta, &diag, &m, &n, &alpha, A, &lda, B, &ldb);
    else eo_dtrsm(side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}
static void x_dsyrk(int use_host, char uplo, char trans, int n, int k, double alpha,
                    const double *A, int lda, double beta, double *C, int ldc) {
    if (use_host) g_dsyrk(&uplo, &trans, &n, &k, &alpha, A, &lda, &beta, C, &ldc);
    else eo_dsyrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}
static void x_dtrsv(int use_host, char uplo, char trans, char diag, int n,
                    const double *A, int lda, double *x) {
    int one = 1;
    if (use_host) g_dtrsv(&uplo, &trans, &diag, &n, A, &lda, x, &one);
    else eo_dtrsv(uplo, trans, diag, n, A, lda, x);
}

/* ------------------------------------------------------------------ */
/* Elemental's algorithms.                                              */

/* LVar3Unb.  Returns 0, or j+1 for the first column whose pivot is <= 0
 * (the reference throws logic_error("A was not numerically HPD") there,
 * leaving columns < j factored).  A NaN pivot is not detected, exactly as
 * in the reference (NaN <= 0 is false). */
int eo_lvar3unb(int n, double *A, int lda) {
    for (int j = 0; j < n; ++j) {
        double alpha = AT(A, j, j, lda);
        if (alpha <= 0.0) return j + 1;
        alpha = sqrt(alpha);
        AT(A, j, j, lda) = alpha;
        for (int k = j + 1; k < n; ++k) AT(A, k, j, lda) /= alpha;
        for (int k = j + 1; k < n; ++k)
            for (int i = k; i < n; ++i)
                AT(A, i, k, lda) -= AT(A, i, j, lda) * AT(A, k, j, lda);
    }
    return 0;
}

/* LVar3: blocked right-looking lower Cholesky with block size nb
 * (Elemental's default Blocksize() is 128).  Each step clamps the diagonal
 * block to the remaining height.  Returns 0 or the global failing column+1. */
static int lvar3_impl(int use_host, int n, double *A, int lda, int nb) {
    if (nb < 1) nb = 1;
    for (int k = 0; k < n; k += nb) {
        int b = (n - k < nb) ? n - k : nb;
        int m2 = n - k - b;
        double *A11 = &AT(A, k, k, lda);
        int info = eo_lvar3unb(b, A11, lda);
        if (info) return k + info;
        if (m2 > 0) {
            double *A21 = &AT(A, k + b, k, lda);
            double *A22 = &AT(A, k + b, k + b, lda);
            x_dtrsm(use_host, 'R', 'L', 'T', 'N', m2, b, 1.0, A11, lda, A21, lda);
            x_dsyrk(use_host, 'L', 'N', m2, b, -1.0, A21, lda, 1.0, A22, lda);
        }
    }
    return 0;
}

int eo_lvar3(int n, double *A, int lda, int nb) { return lvar3_impl(0, n, A, lda, nb); }
int eo_lvar3_hostblas(int n, double *A, int lda, int nb) { return lvar3_impl(1, n, A, lda, nb); }

/* cholesky::SolveAfter(LOWER, NORMAL, L, B): B := inv(L^T)*inv(L)*B.
 * One column -> two Trsv; otherwise two Trsm.  For real data TRANSPOSE
 * orientation only conjugates, which is the identity. */
static void solve_after_impl(int use_host, int n, int k, const double *L, int ldl, double *B, int ldb) {
    if (k == 1) {
        x_dtrsv(use_host, 'L', 'N', 'N', n, L, ldl, B);
        x_dtrsv(use_host, 'L', 'T', 'N', n, L, ldl, B);
    } else {
        x_dtrsm(use_host, 'L', 'L', 'N', 'N', n, k, 1.0, L, ldl, B, ldb);
        x_dtrsm(use_host, 'L', 'L', 'T', 'N', n, k, 1.0, L, ldl, B, ldb);
    }
}

void eo_solve_after(int n, int k, const double *L, int ldl, double *B, int ldb) {
    solve_after_impl(0, n, k, L, ldl, B, ldb);
}
void eo_solve_after_hostblas(int n, int k, const double *L, int ldl, double *B, int ldb) {
    solve_after_impl(1, n, k, L, ldl, B, ldb);
}

int eo_hpd_solve(int n, int k, double *A, int lda, double *B, int ldb, int nb) {
    int info = eo_lvar3(n, A, lda, nb);
    if (info) return info;
    eo_solve_after(n, k, A, lda, B, ldb);
    return 0;
}

int eo_hpd_solve_hostblas(int n, int k, double *A, int lda, double *B, int ldb, int nb) {
    int info = eo_lvar3_hostblas(n, A, lda, nb);
    if (info) return info;
    eo_solve_after_hostblas(n, k, A, lda, B, ldb);
    return 0;
}

/* UVar3Unb (Cholesky/UVar3.hpp:21-49): the upper factor U with A = U^T U
 * overwrites the upper triangle row by row.  Returns 0 or the failing
 * column + 1 (the reference throws logic_error there). */
int eo_uvar3unb(int n, double *A, int lda) {
    for (int j = 0; j < n; ++j) {
        double alpha = AT(A, j, j, lda);
        if (alpha <= 0.0) return j + 1;
        alpha = sqrt(alpha);
        AT(A, j, j, lda) = alpha;
        for (int k = j + 1; k < n; ++k) AT(A, j, k, lda) /= alpha;
        for (int k = j + 1; k < n; ++k)
            for (int i = j + 1; i <= k; ++i)
                AT(A, i, k, lda) -= AT(A, j, i, lda) * AT(A, j, k, lda);
    }
    return 0;
}

/* UVar3 (Cholesky/UVar3.hpp:85-125): per step UVar3Unb(A11);
 * Trsm(LEFT, UPPER, ADJOINT, NON_UNIT, 1, A11, A12);
 * Herk(UPPER, ADJOINT, -1, A12, 1, A22). */
int eo_uvar3(int n, double *A, int lda, int nb) {
    if (nb < 1) nb = 1;
    for (int k = 0; k < n; k += nb) {
        int b = (n - k < nb) ? n - k : nb;
        int m2 = n - k - b;
        double *A11 = &AT(A, k, k, lda);
        int info = eo_uvar3unb(b, A11, lda);
        if (info) return k + info;
        if (m2 > 0) {
            double *A12 = &AT(A, k, k + b, lda);
            double *A22 = &AT(A, k + b, k + b, lda);
            eo_dtrsm('L', 'U', 'T', 'N', b, m2, 1.0, A11, lda, A12, lda);
            eo_dsyrk('U', 'T', m2, b, -1.0, A12, lda, 1.0, A22, lda);
        }
    }
    return 0;
}

/* SolveAfter(UPPER, NORMAL, U, B) (SolveAfter.hpp:33-43, 56-72):
 * B := inv(U) inv(U^T) B; two Trsv for one column, two Trsm otherwise. */
void eo_solve_after_upper(int n, int k, const double *U, int ldu, double *B, int ldb) {
    if (k == 1) {
        eo_dtrsv('U', 'T', 'N', n, U, ldu, B);
        eo_dtrsv('U', 'N', 'N', n, U, ldu, B);
    } else {
        eo_dtrsm('L', 'U', 'T', 'N', n, k, 1.0, U, ldu, B, ldb);
        eo_dtrsm('L', 'U', 'N', 'N', n, k, 1.0, U, ldu, B, ldb);
    }
}

int eo_hpd_solve_upper(int n, int k, double *A, int lda, double *B, int ldb, int nb) {
    int info = eo_uvar3(n, A, lda, nb);
    if (info) return info;
    eo_solve_after_upper(n, k, A, lda, B, ldb);
    return 0;
}

/* ScaleTrapezoid(alpha, LOWER, A, offset=0, side=LEFT). */
void eo_scale_trapezoid_lower(double alpha, int m, int n, double *A, int lda) {
    for (int j = 0; j < n; ++j) {
        int nz = j > 0 ? j : 0;
        for (int i = nz; i < m; ++i) AT(A, i, j, lda) *= alpha;
    }
}

/* AxpyTriangle(LOWER, alpha, X, Y): per column j, Y(j:n,j) += alpha*X(j:n,j)
 * (a daxpy per column; daxpy is y := y + alpha*x elementwise). */
void eo_axpy_triangle_lower(int n, double alpha, const double *X, int ldx, double *Y, int ldy) {
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) AT(Y, i, j, ldy) += alpha * AT(X, i, j, ldx);
}

/* Leaf of TrrkTN: lower trapezoid of C := alpha*A^T*B + beta*C,
 * A and B are k x n, C is n x n.  Elemental's leaf scales the trapezoid by
 * beta, computes the off-diagonal quarter with one GEMM and each diagonal
 * quarter as a full square GEMM into a temporary that is then
 * triangle-added (Trrk/Local.hpp:452-490 and 492-536). */
static void trrk_tn_leaf(int k, int n, double alpha, const double *A, int lda,
                         const double *B, int ldb, double beta, double *C, int ldc) {
    int half = n / 2;
    eo_scale_trapezoid_lower(beta, n, n, C, ldc);
    /* CBL += alpha * AR^T * BL   (AR = A(:,half:n), BL = B(:,0:half)) */
    eo_dgemm('T', 'N', n - half, half, k, alpha, &AT(A, 0, half, lda), lda,
             B, ldb, 1.0, &AT(C, half, 0, ldc), ldc);
    /* DTL := alpha*AL^T*BL ; CTL_lower += DTL_lower */
    if (half > 0) {
        double *D = (double *)malloc(sizeof(double) * (size_t)half * (size_t)half);
        eo_dgemm('T', 'N', half, half, k, alpha, A, lda, B, ldb, 0.0, D, half);
        eo_axpy_triangle_lower(half, 1.0, D, half, C, ldc);
        free(D);
    }
    int h2 = n - half;
    if (h2 > 0) {
        double *D = (double *)malloc(sizeof(double) * (size_t)h2 * (size_t)h2);
        eo_dgemm('T', 'N', h2, h2, k, alpha, &AT(A, 0, half, lda), lda,
                 &AT(B, 0, half, ldb), ldb, 0.0, D, h2);
        eo_axpy_triangle_lower(h2, 1.0, D, h2, &AT(C, half, half, ldc), ldc);
        free(D);
    }
}

/* Recursive halving of TrrkTN / LocalTrrk with leaf threshold `bsize`
 * (LocalTrrkBlocksize<double>() = 64 by default, global.cpp:48-51; the
 * 1x1 grid makes g.Width() = 1). */
void eo_trrk_tn(int k, int n, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int bsize) {
    if (n < bsize) {
        trrk_tn_leaf(k, n, alpha, A, lda, B, ldb, beta, C, ldc);
        return;
    }
    int half = n / 2;
    eo_dgemm('T', 'N', n - half, half, k, alpha, &AT(A, 0, half, lda), lda,
             B, ldb, beta, &AT(C, half, 0, ldc), ldc);
    eo_trrk_tn(k, half, alpha, A, lda, B, ldb, beta, C, ldc, bsize);
    eo_trrk_tn(k, n - half, alpha, &AT(A, 0, half, lda), lda, &AT(B, 0, half, ldb), ldb,
               beta, &AT(C, half, half, ldc), ldc, bsize);
}

/* ------------------------------------------------------------------ */
/* Norms used by the checks.                                             */

/* R := B - A*X with A symmetric from its lower triangle (n x n), X, B n x k. */
void eo_sym_lower_residual(int n, int k, const double *A, int lda, const double *X, int ldx,
                           const double *B, int ldb, double *R, int ldr) {
    for (int c = 0; c < k; ++c) {
        for (int i = 0; i < n; ++i) AT(R, i, c, ldr) = AT(B, i, c, ldb);
        for (int j = 0; j < n; ++j) {
            double xj = AT(X, j, c, ldx);
            AT(R, j, c, ldr) -= AT(A, j, j, lda) * xj;
            for (int i = j + 1; i < n; ++i) {
                double a = AT(A, i, j, lda);
                AT(R, i, c, ldr) -= a * xj;
                AT(R, j, c, ldr) -= a * AT(X, i, c, ldx);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Triangular inverse and HPD inverse: the callers of Cholesky that invert. */

/* triangular_inverse::LVar3Unb (include/elemental/lapack-like/
 * TriangularInverse/LVar3.hpp:19-49) on the lower triangle: per column j,
 * row j left of the diagonal /= -lambda, the rank-1 Geru update of the block
 * below-left of it (Netlib dger order: column by column, temp = 1*y[k],
 * columns with y[k] == 0 skipped), column j below the diagonal /= lambda and
 * the pivot inverted.  unit: lambda = 1 and the diagonal is left alone. */
void eo_trtri_lvar3unb(int unit, int n, double *L, int ldl) {
    for (int j = 0; j < n; ++j) {
        const double lambda = unit ? 1.0 : AT(L, j, j, ldl);
        for (int k = 0; k < j; ++k) AT(L, j, k, ldl) /= -lambda;
        for (int k = 0; k < j; ++k) {
            const double temp = 1.0 * AT(L, j, k, ldl);
            if (temp != 0.0)
                for (int i = j + 1; i < n; ++i) AT(L, i, k, ldl) += AT(L, i, j, ldl) * temp;
        }
        if (!unit) {
            for (int k = j + 1; k < n; ++k) AT(L, k, j, ldl) /= lambda;
            AT(L, j, j, ldl) = 1.0 / AT(L, j, j, ldl);
        }
    }
}

/* triangular_inverse::LVar3 (TriangularInverse/LVar3.hpp:52-95), lower: per
 * diagonal block, Trsm(LEFT, LOWER, NORMAL, diag, -1, L11, L10),
 * Gemm(NORMAL, NORMAL, 1, L21, L10, 1, L20), Trsm(RIGHT, LOWER, NORMAL,
 * diag, 1, L11, L21), LVar3Unb(diag, L11). */
void eo_trtri_lvar3(int unit, int n, double *L, int ldl, int nb) {
    const char diag = unit ? 'U' : 'N';
    if (nb < 1) nb = 1;
    for (int k = 0; k < n; k += nb) {
        const int b = (n - k < nb) ? n - k : nb, m = n - k - b;
        double *L10 = &AT(L, k, 0, ldl), *L11 = &AT(L, k, k, ldl);
        double *L20 = &AT(L, k + b, 0, ldl), *L21 = &AT(L, k + b, k, ldl);
        if (k > 0) eo_dtrsm('L', 'L', 'N', diag, b, k, -1.0, L11, ldl, L10, ldl);
        if (m > 0 && k > 0) eo_dgemm('N', 'N', m, k, b, 1.0, L21, ldl, L10, ldl, 1.0, L20, ldl);
        if (m > 0) eo_dtrsm('R', 'L', 'N', diag, m, b, 1.0, L11, ldl, L21, ldl);
        eo_trtri_lvar3unb(unit, b, L11, ldl);
    }
}

/* internal::TrtrmmLUnblocked(TRANSPOSE) (include/elemental/blas-like/level3/
 * Trtrmm/Unblocked.hpp:15-57): the lower triangle of L := L^T L. */
void eo_trtrmm_lunb(int n, double *L, int ldl) {
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < j; ++k) {
            const double gamma = AT(L, j, k, ldl);
            for (int i = k; i < j; ++i) AT(L, i, k, ldl) += AT(L, j, i, ldl) * gamma;
        }
        const double lambda = AT(L, j, j, ldl);
        for (int k = 0; k < j; ++k) AT(L, j, k, ldl) *= lambda;
        AT(L, j, j, ldl) = lambda * lambda;
    }
}

/* hpd_inverse::CholeskyLVar2 (include/elemental/lapack-like/Inverse/
 * CholeskyLVar2.hpp:29-75) on the lower triangle with block size nb.  The
 * calls on the diagonal block use the reference's global Blocksize() (128),
 * so for nb <= 128 they are single-block: Cholesky -> LVar3Unb,
 * TriangularInverse -> LVar3Unb, Trtrmm -> TrtrmmLUnblocked.  Returns 0, or
 * the global failing column + 1 (the reference throws there). */
int eo_hpd_inverse_lvar2(int n, double *A, int lda, int nb) {
    if (nb < 1) nb = 1;
    for (int k = 0; k < n; k += nb) {
        const int b = (n - k < nb) ? n - k : nb, m = n - k - b;
        double *A00 = A, *A10 = &AT(A, k, 0, lda), *A11 = &AT(A, k, k, lda);
        double *A20 = &AT(A, k + b, 0, lda), *A21 = &AT(A, k + b, k, lda), *A22 = &AT(A, k + b, k + b, lda);
        const int info = eo_lvar3(b, A11, lda, 128);
        if (info) return k + info;
        if (k > 0) eo_dtrsm('L', 'L', 'N', 'N', b, k, 1.0, A11, lda, A10, lda);
        if (m > 0) eo_dtrsm('R', 'L', 'T', 'N', m, b, 1.0, A11, lda, A21, lda);
        if (k > 0) eo_dsyrk('L', 'T', k, b, 1.0, A10, lda, 1.0, A00, lda);
        if (k > 0 && m > 0) eo_dgemm('N', 'N', m, k, b, -1.0, A21, lda, A10, lda, 1.0, A20, lda);
        if (m > 0) eo_dsyrk('L', 'N', m, b, -1.0, A21, lda, 1.0, A22, lda);
        if (k > 0) eo_dtrsm('L', 'L', 'T', 'N', b, k, 1.0, A11, lda, A10, lda);
        if (m > 0) eo_dtrsm('R', 'L', 'N', 'N', m, b, -1.0, A11, lda, A21, lda);
        eo_trtri_lvar3(0, b, A11, lda, 128);
        eo_trtrmm_lunb(b, A11, lda);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Two-sided reductions of the generalized-definite eigensolver.        */

/* internal::TwoSidedTrsmLUnb (include/elemental/blas-like/level3/
 * TwoSidedTrsm.hpp:14-60), variant 4, non-unit: the lower triangle of
 * A := inv(L) A inv(L)^T, with the reference BLAS orders of Geru (column by
 * column, temp = alpha*y[k], zero y skipped) and Her2 (dsyr2 lower: column
 * by column, temp1 = alpha*y[c], temp2 = alpha*x[c]). */
void eo_two_sided_trsm_lunb(int n, double *A, int lda, const double *L, int ldl) {
    for (int j = 0; j < n; ++j) {
        const int h = n - (j + 1);
        const double lambda11 = AT(L, j, j, ldl);
        for (int k = 0; k < j; ++k) AT(A, j, k, lda) /= lambda11;
        for (int k = 0; k < j; ++k) { /* A20 -= l21 a10 */
            const double temp = -1.0 * AT(A, j, k, lda);
            if (temp != 0.0)
                for (int i = 0; i < h; ++i) AT(A, j + 1 + i, k, lda) += AT(L, j + 1 + i, j, ldl) * temp;
        }
        AT(A, j, j, lda) /= lambda11 * lambda11;
        const double alpha11 = AT(A, j, j, lda);
        for (int k = 0; k < h; ++k) AT(A, j + 1 + k, j, lda) /= lambda11;
        for (int k = 0; k < h; ++k) AT(A, j + 1 + k, j, lda) -= (alpha11 / 2) * AT(L, j + 1 + k, j, ldl);
        for (int c = 0; c < h; ++c) { /* A22 -= l21 a21^T + a21 l21^T (lower) */
            const double xc = AT(L, j + 1 + c, j, ldl), yc = AT(A, j + 1 + c, j, lda);
            if (xc != 0.0 || yc != 0.0) {
                const double temp1 = -1.0 * yc, temp2 = -1.0 * xc;
                for (int i = c; i < h; ++i)
                    AT(A, j + 1 + i, j + 1 + c, lda) +=
                        AT(L, j + 1 + i, j, ldl) * temp1 + AT(A, j + 1 + i, j, lda) * temp2;
            }
        }
        for (int k = 0; k < h; ++k) AT(A, j + 1 + k, j, lda) -= (alpha11 / 2) * AT(L, j + 1 + k, j, ldl);
    }
}

/* internal::TwoSidedTrmmLUnb (include/elemental/blas-like/level3/
 * TwoSidedTrmm.hpp:14-64), variant 4, non-unit: the lower triangle of
 * A := L^T A L. */
void eo_two_sided_trmm_lunb(int n, double *A, int lda, const double *L, int ldl) {
    for (int j = 0; j < n; ++j) {
        const int h = n - (j + 1);
        const double alpha11 = AT(A, j, j, lda);
        const double lambda11 = AT(L, j, j, ldl);
        for (int k = 0; k < j; ++k) AT(A, j, k, lda) += (alpha11 / 2) * AT(L, j, k, ldl);
        for (int c = 0; c < j; ++c) { /* A00 += a10^T l10 + l10^T a10 (lower), x = a10, y = l10 */
            const double xc = AT(A, j, c, lda), yc = AT(L, j, c, ldl);
            if (xc != 0.0 || yc != 0.0) {
                const double temp1 = 1.0 * yc, temp2 = 1.0 * xc;
                for (int i = c; i < j; ++i) AT(A, i, c, lda) += AT(A, j, i, lda) * temp1 + AT(L, j, i, ldl) * temp2;
            }
        }
        for (int k = 0; k < j; ++k) AT(A, j, k, lda) += (alpha11 / 2) * AT(L, j, k, ldl);
        for (int k = 0; k < j; ++k) AT(A, j, k, lda) *= lambda11;
        AT(A, j, j, lda) *= lambda11 * lambda11;
        for (int k = 0; k < j; ++k) { /* A20 += a21 l10 */
            const double temp = 1.0 * AT(L, j, k, ldl);
            if (temp != 0.0)
                for (int i = 0; i < h; ++i) AT(A, j + 1 + i, k, lda) += AT(A, j + 1 + i, j, lda) * temp;
        }
        for (int k = 0; k < h; ++k) AT(A, j + 1 + k, j, lda) *= lambda11;
    }
}
