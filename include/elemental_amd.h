/*
 * elemental_amd.h -- C ABI of the MI355X (gfx950) HPD Cholesky factor-and-solve
 * library.  It is the drop-in boundary for Elemental's lower Cholesky /
 * HPDSolve path on one GPU's local matrix.
 *
 * Conventions
 *   - Matrices are column-major with a leading dimension in elements, exactly
 *     like elem::Matrix<F> (include/elemental/core/matrix.hpp:16-183, element
 *     (i,j) at i + j*ldim).  All matrix pointers are DEVICE pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Calls are asynchronous on that stream unless documented otherwise.
 *   - Return value: ELEMAMD_OK (0) when the work was enqueued (or done), a
 *     negative ELEMAMD_ERR_* code on an invalid argument or a HIP failure.
 *   - Factorizations report the reference's non-HPD condition through a
 *     device int `dinfo`: 0 on success, otherwise the 1-based global column
 *     whose pivot was <= 0 -- the column at which the reference's
 *     cholesky::LVar3Unb throws std::logic_error("A was not numerically HPD")
 *     (include/elemental/lapack-like/Cholesky/LVar3.hpp:37-38).  Work after a
 *     failing block is skipped, so A is left as the reference leaves it.  As
 *     in the reference, a NaN pivot is not detected.
 *   - Character arguments follow the Fortran BLAS conventions the reference
 *     forwards to (src/core/imports/blas.cpp): uplo 'L'/'U', trans 'N'/'T'/'C'
 *     ('C' == 'T' for real data), side 'L'/'R', diag 'N'/'U'.
 *   - Every entry point is re-entrant: no global mutable state; block size,
 *     precision and refinement parameters are per call.
 */
#ifndef ELEMENTAL_AMD_H
#define ELEMENTAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ELEMAMD_OK 0
#define ELEMAMD_ERR_ARG (-1)
#define ELEMAMD_ERR_HIP (-2)
#define ELEMAMD_ERR_NOT_CONVERGED (-3)
#define ELEMAMD_ERR_WORKSPACE (-4)

/* Trailing-update arithmetic of the low-precision factor. */
#define ELEMAMD_TRAILING_F64 0
#define ELEMAMD_TRAILING_F32 1
#define ELEMAMD_TRAILING_BF16X3 2
#define ELEMAMD_TRAILING_BF16X6 3

/* Library version string. */
const char *elemamd_version(void);

/* Synchronize `stream` and return the value of the device int *dinfo
 * (a negative ELEMAMD_ERR_* if the copy fails). */
int elemamd_info_read(const int *dinfo, void *stream);

/* ------------------------------------------------------------------ */
/* Local kernels of the blocked factorization.                        */

/* Unblocked lower Cholesky of an n x n block, n <= 128, in place.
 * Replaces cholesky::LVar3Unb (include/elemental/lapack-like/Cholesky/LVar3.hpp:19-49),
 * reached from LocalCholesky (include/elemental/lapack-like/Cholesky.hpp:29-37).
 * Bit-identical to the reference loop in IEEE double.  *dinfo must be 0 on
 * entry (a nonzero value makes the call a no-op, as after an earlier failure). */
int elemamd_potrf_unb_f64(int n, double *A, int64_t lda, int *dinfo, void *stream);

/* B := alpha * op(A)^{-1} B (side 'L') or alpha * B op(A)^{-1} (side 'R'),
 * A lower (uplo 'L') or upper ('U') triangular, non-unit ('N') or unit ('U')
 * diagonal, op = N or T ('C' == 'T').  Replaces elem::Trsm(Matrix)
 * (include/elemental/blas-like/level3/Trsm.hpp:49-86) -> blas::Trsm ->
 * dtrsm_ (src/core/imports/blas.cpp:1203-1210); every combination the
 * reference accepts.  side 'R', lower, trans 'T', non-unit with n <= 128 is
 * the Cholesky panel solve, bit-identical to the column-oriented reference
 * BLAS loop; side 'L' runs the persistent chain solves; the other side 'R'
 * forms solve B^T on the left through a transposed copy.  alpha != 1
 * pre-scales B.  Every argument is checked before anything is enqueued (an
 * ELEMAMD_ERR_ARG return leaves B untouched).  Synchronous when it runs a
 * persistent solve (a bounded wait that gave up returns ELEMAMD_ERR_HIP). */
int elemamd_trsm_f64(char side, char uplo, char trans, char diag, int m, int n, double alpha,
                     const double *A, int64_t lda, double *B, int64_t ldb, void *stream);

/* C := alpha*op(A)*op(A)^T + beta*C on the uplo triangle of C ('L' or 'U'),
 * op(A) = A (trans 'N', A n x k) or A^T (trans 'T'/'C', A k x n).  Replaces
 * elem::Herk / Syrk(Matrix) (include/elemental/blas-like/level3/Herk.hpp:17-27,
 * Syrk.hpp:20-57) -> blas::Herk -> dsyrk_ (src/core/imports/blas.cpp:1043-1050).
 * The other triangle of C is neither read nor written.  ('L', 'N') is the
 * Cholesky trailing update's kernel; the other forms run the TN product on
 * the requested triangle (trans 'N' through a transposed copy of A). */
int elemamd_herk_f64(char uplo, char trans, int n, int k, double alpha, const double *A, int64_t lda,
                     double beta, double *C, int64_t ldc, void *stream);

/* Float-storage trailing update of the mixed-precision factor (no reference
 * counterpart): mode ELEMAMD_TRAILING_F32 (FP32 MFMA),
 * ELEMAMD_TRAILING_BF16X3 (split-BF16x3, FP32 accumulation: A is split into
 * bf16 hi/lo planes in a per-stream scratch buffer, then the three products
 * hi*hi + hi*lo + lo*hi accumulate in FP32) or ELEMAMD_TRAILING_BF16X6
 * (three planes hi/mid/lo and the six products down to 2^-24 of a term;
 * the bf16 MFMA drops the bits of small products below the largest one's
 * 24-bit window, so the result is less accurate than FP32's). */
int elemamd_herk_f32(int mode, int n, int k, float alpha, const float *A, int64_t lda, float beta, float *C,
                     int64_t ldc, void *stream);

/* The uplo triangle ('L' or 'U') of C := alpha * op(A) * op(B) + beta*C,
 * op(A) n x k, op(B) k x n, orientations 'N' or 'T'/'C'.  Replaces
 * elem::Trrk(Matrix) / LocalTrrk (src/blas-like/Trrk.cpp:19-38,
 * src/blas-like/Trrk/Local.hpp:733-775 and :931-976); the TN form is the
 * direct product, the other orientations go through transposed copies. */
int elemamd_trrk_f64(char uplo, char orientA, char orientB, int n, int k, double alpha, const double *A,
                     int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc, void *stream);

/* x := op(A)^{-1} x for one vector (incx must be 1), A lower or upper,
 * non-unit or unit diagonal.  Replaces elem::Trsv
 * (include/elemental/blas-like/level2/Trsv.hpp:20-44) -> dtrsv_
 * (src/core/imports/blas.cpp:901-904).  Synchronous. */
int elemamd_trsv_f64(char uplo, char trans, char diag, int n, const double *A, int64_t lda, double *x,
                     int64_t incx, void *stream);

/* ------------------------------------------------------------------ */
/* The path.                                                          */

/* Blocked right-looking Cholesky (variant 3) with block size nb (1..128;
 * Elemental's default Blocksize() is 128, src/core/global.cpp:194).
 * Replaces elem::Cholesky(uplo, Matrix<F>&)
 * (include/elemental/lapack-like/Cholesky.hpp:49-62).  uplo 'L' ->
 * cholesky::LVar3 (include/elemental/lapack-like/Cholesky/LVar3.hpp:83-123):
 * the lower triangle of A becomes L, the strictly upper triangle is
 * untouched.  uplo 'U' -> cholesky::UVar3 (Cholesky/UVar3.hpp:85-125): the
 * upper triangle becomes U with A = U^T U, the strictly lower triangle is
 * untouched; it runs as the lower factorization of A^T (UVar3Unb on A is
 * LVar3Unb on A^T operation by operation), with A transposed in place
 * before and after.  Asynchronous: the outcome is written to *dinfo. */
int elemamd_cholesky_f64(char uplo, int n, double *A, int64_t lda, int nb, int *dinfo, void *stream);

/* B := A^{-1} B given the factor in A's uplo triangle.  Replaces
 * cholesky::SolveAfter(uplo, orientation, A, B)
 * (include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75): two Trsv
 * for one column, two Trsm otherwise; LOWER solves with L then L^T, UPPER
 * with U^T then U.  orientation 'N', 'T' or 'C' (identical for real data).
 * Synchronous: returns once the solves are done, ELEMAMD_ERR_HIP if a
 * persistent solve's bounded wait gave up (the result is then wrong). */
int elemamd_solve_after_f64(char uplo, char orientation, int n, int k, const double *A, int64_t lda, double *B,
                            int64_t ldb, void *stream);

/* Cholesky then SolveAfter.  Replaces elem::HPDSolve(uplo, orientation,
 * Matrix<F>& A, Matrix<F>& B) (include/elemental/lapack-like/HPDSolve.hpp:18-28).
 * The solve is skipped when *dinfo becomes nonzero.  Synchronous, as
 * SolveAfter. */
int elemamd_hpd_solve_f64(char uplo, char orientation, int n, int k, double *A, int64_t lda, double *B,
                          int64_t ldb, int nb, int *dinfo, void *stream);

/* C := alpha*op(A)*op(B) + beta*C, op = N or T ('C' == 'T'), op(A) m x k,
 * op(B) k x n.  Replaces elem::Gemm(Orientation, Orientation, T, const
 * Matrix<T>&, const Matrix<T>&, T, Matrix<T>&)
 * (include/elemental/blas-like/level3/Gemm.hpp) -> blas::Gemm -> dgemm_.
 * beta == 0: C is not read.  Asynchronous. */
int elemamd_gemm_f64(char transa, char transb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                     const double *B, int64_t ldb, double beta, double *C, int64_t ldc, void *stream);

/* A := A^{-1} for symmetric positive definite A, on the uplo triangle (the
 * other triangle is untouched).  Replaces elem::HPDInverse(uplo, Matrix<F>& A)
 * (include/elemental/lapack-like/Inverse.hpp:29-40) ->
 * hpd_inverse::CholeskyLVar2 (Inverse/CholeskyLVar2.hpp:29-75); UPPER is the
 * same algorithm on A^T.  Block size nb (1..128; the reference's default
 * 128).  A non-positive pivot sets *dinfo to its 1-based column and stops
 * (A is left as the reference leaves it when Cholesky(A11) throws).
 * Asynchronous: read *dinfo after the stream. */
int elemamd_hpd_inverse_f64(char uplo, int n, double *A, int64_t lda, int nb, int *dinfo, void *stream);

/* A := A^{-1} for triangular A ('L' or 'U'), unit ('U') or non-unit ('N')
 * diagonal; the other triangle (and a unit diagonal) is untouched.
 * Replaces elem::TriangularInverse(uplo, diag, Matrix<F>& A)
 * (include/elemental/lapack-like/TriangularInverse.hpp) ->
 * triangular_inverse::LVar3 (TriangularInverse/LVar3.hpp:19-95).
 * Asynchronous. */
int elemamd_triangular_inverse_f64(char uplo, char diag, int n, double *A, int64_t lda, int nb, void *stream);

/* Generalized-definite symmetric eigenproblem.  Replaces
 * elem::HermitianGenDefiniteEig(type, uplo, Matrix<F>& A, Matrix<F>& B,
 * Matrix<R>& w[, Matrix<F>& X][, a, b]) (include/elemental/lapack-like/
 * HermitianGenDefiniteEig.hpp:22-126, 135-250): Cholesky(uplo, B), then
 * A := inv(L) A inv(L)^T (type AXBX) or L^T A L (ABX, BAX), the standard
 * eigenproblem of A, and, when X is not NULL, X := inv(L)^T Z (AXBX, ABX) or
 * L Z (BAX).  Only the uplo triangles of A and B are read; on return B's
 * holds the factor and A's the reduced matrix.  range: ELEMAMD_EIG_ALL,
 * ELEMAMD_EIG_INDEX (eigenvalues il..iu, 0-based inclusive, ascending) or
 * ELEMAMD_EIG_VALUE (eigenvalues in (vl, vu]).  w: device array of n doubles
 * (the first *nev hold the eigenvalues, ascending); X: device n x *nev
 * (ldx >= n).  *info (host) = B's failing column (1-based) when B is not
 * numerically HPD (nothing else is done), else 0.  Synchronous; returns
 * ELEMAMD_ERR_NOT_CONVERGED if the eigensolver does not converge. */
#define ELEMAMD_AXBX 1
#define ELEMAMD_ABX 2
#define ELEMAMD_BAX 3
#define ELEMAMD_EIG_ALL 0
#define ELEMAMD_EIG_INDEX 1
#define ELEMAMD_EIG_VALUE 2
int elemamd_hermitian_gen_definite_eig_f64(int type, char uplo, int n, double *A, int64_t lda, double *B,
                                           int64_t ldb, double *w, double *X, int64_t ldx, int range, int il, int iu,
                                           double vl, double vu, int *nev, int *info, void *stream);

/* The same three entry points for Matrix<float> (the reference templates
 * Cholesky<F>, cholesky::SolveAfter<F> and HPDSolve<F> on the field,
 * Cholesky.hpp:47-62, SolveAfter.hpp:20-75, HPDSolve.hpp:18-28): the float
 * factorization keeps each 128 x 128 diagonal block's factor in FP64 before
 * storing it as float, solves the panels by products with the inverses of
 * its 32 x 32 diagonal pieces and updates the trailing matrix with FP32
 * MFMA; the float solves use the inverses of the 64 x 64 diagonal blocks. */
int elemamd_cholesky_f32(char uplo, int n, float *A, int64_t lda, int nb, int *dinfo, void *stream);
int elemamd_solve_after_f32(char uplo, char orientation, int n, int k, const float *A, int64_t lda, float *B,
                            int64_t ldb, void *stream);
int elemamd_hpd_solve_f32(char uplo, char orientation, int n, int k, float *A, int64_t lda, float *B,
                          int64_t ldb, int nb, int *dinfo, void *stream);

/* ------------------------------------------------------------------ */
/* Mixed precision with FP64 iterative refinement (no reference         */
/* counterpart; results are held to the FP64 backward-error target).    */

typedef struct {
    int nb;               /* factorization block size, 1..128 (0 -> 128) */
    int trailing;         /* ELEMAMD_TRAILING_F32, _BF16X3 or _BF16X6 (_BF16X6: nb a multiple of 32) */
    int max_iter;         /* refinement steps before giving up (0 -> 30) */
    int gmres_restart;    /* Krylov dimension of a GMRES-IR cycle: 0 -> 20, negative disables GMRES-IR */
    int force_gmres;      /* 1: use GMRES-IR from the first correction */
    int allow_shift;      /* 1: on low-precision breakdown, retry the factor of A + sigma*diag(A) */
    double tol;           /* target normwise backward error (inf-norms); 0 -> 1e-14 */
    int refine;           /* ELEMAMD_REFINE_IR (classic) or ELEMAMD_REFINE_CG (preconditioned CG) */
    int fallback;         /* 1: when the low-precision route fails (breakdown without a shift, a GMRES-IR
                           * cycle that does not halve the backward error, or max_iter reached), solve with
                           * the plain FP64 Cholesky + SolveAfter on a copy of A in the workspace */
} elemamd_ir_opts;

/* Refinement of the mixed solve.  IR: x += M^{-1} (b - A x) with the
 * low-precision factor M, GMRES-IR on stagnation.  CG: conjugate gradients
 * on A x = b preconditioned by M (flexible direction update, FP64 products,
 * iterates and residuals; the true residual confirms convergence), classic
 * IR taking over if it stagnates.  report->iterations counts FP64 products
 * with A (residuals for IR, CG steps for CG). */
#define ELEMAMD_REFINE_IR 0
#define ELEMAMD_REFINE_CG 1

typedef struct {
    int info;             /* failing column of the FP64 fallback factor (0 = none) */
    int lowprec_info;     /* failing column of the last low-precision factor attempt (0 = none) */
    int iterations;       /* refinement steps (FP64 residuals computed) */
    int gmres_iterations; /* total inner GMRES iterations */
    int used_gmres;       /* 1 if GMRES-IR was used */
    int shifted;          /* 1 if the low-precision factor needed a diagonal shift */
    int converged;        /* 1 if backward_error <= tol */
    double backward_error;/* max over columns of ||b-Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
    double anorm_inf;     /* ||A||_inf */
    int fallback;         /* 1 if X came from the plain FP64 fallback solve */
} elemamd_ir_report;

/* Workspace bytes needed by elemamd_hpd_solve_mixed_f64. */
size_t elemamd_hpd_solve_mixed_workspace(int n, int k, const elemamd_ir_opts *opts);

/* Solve A X = B for symmetric positive definite A (lower triangle read, A is
 * NOT modified) to FP64 backward error: low-precision factor (FP64 diagonal
 * potrf, FP32-stored panels, FP32 or split-BF16x3 trailing update) then FP64
 * residual iterative refinement, falling back to GMRES-IR preconditioned by
 * the low-precision factor when classic refinement stagnates.  B is
 * overwritten with X.  Synchronous (it reads convergence norms each step).
 * `work` must hold elemamd_hpd_solve_mixed_workspace() bytes of device memory.
 * With opts->fallback the solve is finished in plain FP64 when the
 * low-precision route fails (the workspace then includes an n x n FP64 copy
 * of A).  Returns ELEMAMD_OK, ELEMAMD_ERR_NOT_CONVERGED (X holds the best
 * iterate; with the fallback: A is not numerically HPD, report->info is the
 * failing column, or the input holds NaN/Inf), or another error; *report
 * (may be NULL) receives the details. */
int elemamd_hpd_solve_mixed_f64(int n, int k, const double *A, int64_t lda, double *B, int64_t ldb,
                                const elemamd_ir_opts *opts, void *work, size_t work_bytes,
                                elemamd_ir_report *report, void *stream);

/* ------------------------------------------------------------------ */
/* Batched factor-and-solve of many independent matrices of uneven     */
/* size (no reference counterpart: Elemental factors one matrix per    */
/* call; each matrix goes through the same LVar3 + SolveAfter steps).  */

#define ELEMAMD_BATCH_F64 0         /* plain FP64, A_i overwritten with L_i */
#define ELEMAMD_BATCH_F32_IR 1      /* FP32 factor + FP64-residual refinement, A_i read only */
#define ELEMAMD_BATCH_BF16X3_IR 2   /* as 1 with split-BF16x3 trailing updates */

typedef struct elemamd_batch_plan elemamd_batch_plan;

/* Plan for `count` matrices of sizes n[i] with k[i] right-hand sides.
 * Matrix i is assigned to the first size bucket b with n[i] <= bucket_max_n[b]
 * (increasing; the last bucket also takes every larger matrix) and is
 * factored with that bucket's block size bucket_nb[b] (1..128).  Inputs are
 * packed: A_i column-major with ld = n[i] at element offset a_off[i] of one
 * buffer, B_i (n[i] x k[i], ld = n[i]) at b_off[i] (see _layout). */
int elemamd_batch_plan_create(int count, const int *n, const int *k, int precision, int nbuckets,
                              const int *bucket_max_n, const int *bucket_nb, elemamd_batch_plan **plan);
int elemamd_batch_plan_destroy(elemamd_batch_plan *plan);
/* Launch form of each size bucket (the create call's bucket order):
 * bucket_grid[b] = 0 runs bucket b as one potrf, one panel and one trailing
 * launch per block step (the default); g > 0 runs its whole factorization,
 * and for ELEMAMD_BATCH_F64 its solves, as ONE persistent launch of g
 * workgroups that take the bucket's matrices largest first.  Results are
 * the same bit for bit.  No reference counterpart
 * (for callers factoring many small matrices). */
int elemamd_batch_set_launch(elemamd_batch_plan *plan, const int *bucket_grid);
/* Element offsets of every A_i / B_i and the total element counts. */
int elemamd_batch_plan_layout(const elemamd_batch_plan *plan, int64_t *a_off, int64_t *b_off, int64_t *a_total,
                              int64_t *b_total);
/* Refinement of the low-precision modes: a fixed number of sweeps (so the
 * launch sequence can be captured and replayed) and the per-matrix target;
 * a matrix whose backward error meets `tol` skips the remaining sweeps.
 * Defaults: 4 sweeps, 1e-14. */
int elemamd_batch_set_refinement(elemamd_batch_plan *plan, int sweeps, double tol);
size_t elemamd_batch_workspace(const elemamd_batch_plan *plan);
/* Factor and solve every matrix; B_i is overwritten with X_i.  With
 * use_graph = 1 the launch sequence is captured in a hipGraph on the first
 * call (on a private stream; again whenever A, B or work change) and
 * launched on `stream` as one submission.  Asynchronous on `stream`. */
int elemamd_batch_hpd_solve(elemamd_batch_plan *plan, double *A, double *B, void *work, size_t work_bytes,
                            int use_graph, void *stream);
/* Synchronize and fetch per-matrix results in the caller's order: factor
 * status (0, or the 1-based failing column), normwise backward error (the
 * refinement modes; NaN for plain FP64), refinement sweeps used.  Any
 * pointer may be NULL. */
int elemamd_batch_results(const elemamd_batch_plan *plan, int *info, double *backward_error, int *iterations,
                          void *stream);
/* eta[i] = max_c ||b_c - A_i x_c||_inf / (||A_i||_inf ||x_c||_inf + ||b_c||_inf)
 * for packed A (symmetric, lower triangle read), X and B; R receives the
 * residuals (b_total elements).  Synchronous. */
int elemamd_batch_backward_error(const elemamd_batch_plan *plan, const double *A, const double *X, const double *B,
                                 double *R, double *eta, void *stream);

/* ------------------------------------------------------------------ */
/* Serving loop (no reference counterpart): continuous batching of      */
/* factor-and-solve requests of mixed size on one GPU.                  */

#define ELEMAMD_ERR_NOT_HPD (-5)   /* a served FP64 factorization met a non-positive pivot */

typedef struct elemamd_server elemamd_server;

typedef struct {
    int slots;              /* requests running at once, one HIP stream each (1..64) */
    int precision;          /* ELEMAMD_BATCH_F64, _F32_IR or _BF16X3_IR */
    int nb;                 /* block size (0 -> 128) */
    double tol;             /* refinement target (0 -> 1e-14) */
    size_t memory_budget;   /* bytes of HBM the server may reserve (0 -> free HBM at creation minus a reserve) */
} elemamd_server_config;

typedef struct {
    int64_t ticket;
    int status;             /* ELEMAMD_OK or a negative error (ELEMAMD_ERR_NOT_HPD: info = failing column) */
    int info;
    int factored;           /* 1: this request computed its factor; 0: it reused a resident one */
    int iterations;         /* refinement steps (low-precision modes) */
    int gmres_iterations;
    int slot;
    double backward_error;  /* low-precision modes: final normwise backward error */
    double t_submit, t_start, t_done;  /* seconds since the server was created */
} elemamd_serve_result;

typedef struct {
    int64_t completed, reuse_hits, evictions;
    size_t peak_reserved, reserved, budget;
    int resident_factors;
} elemamd_server_statistics;

/* Start `slots` worker threads on the current device. */
int elemamd_server_create(const elemamd_server_config *cfg, elemamd_server **server);
/* Queue a request: solve A X = B (A n x n device, lower triangle read and
 * never modified; B n x k device, overwritten with X).  The factor is kept
 * resident under `key`; with reuse = 1 a resident factor of `key` is used
 * instead of factoring again.  The request starts after the work enqueued so
 * far on `ready_stream` (the producer of A and B; NULL = the default
 * stream); the server runs on streams of its own.  A and B must stay valid until the
 * request is waited for.  Thread-safe; returns at once with a ticket. */
int elemamd_server_submit(elemamd_server *server, int key, int reuse, int n, int k, const double *A, int64_t lda,
                          double *B, int64_t ldb, void *ready_stream, int64_t *ticket);
/* Block until the ticket's request completes; fills *result (may be NULL). */
int elemamd_server_wait(elemamd_server *server, int64_t ticket, elemamd_serve_result *result);
/* Release the resident factor of `key` (an error if a request is using it). */
int elemamd_server_evict(elemamd_server *server, int key);
int elemamd_server_stats(elemamd_server *server, elemamd_server_statistics *stats);
/* Drain the queue, stop the workers, free every resident factor. */
int elemamd_server_destroy(elemamd_server *server);

/* ------------------------------------------------------------------ */
/* Utilities.                                                          */

/* Fill a rows x cols column-major block with N(0,1) numbers: entry t in
 * column-major order is number offset+t of the splitmix64 / Box-Muller
 * stream `seed`. */
int elemamd_fill_normal_f64(double *out, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, int64_t offset,
                            void *stream);

/* R := B - A*X with A symmetric, lower triangle read (the FP64 refinement
 * residual), n x n; X, B, R n x k. */
int elemamd_sym_residual_f64(int n, int k, const double *A, int64_t lda, const double *X, int64_t ldx,
                             const double *B, int64_t ldb, double *R, int64_t ldr, void *stream);

/* ------------------------------------------------------------------ */
/* Allocation-free SolveAfter.                                         */

/* elemamd_solve_after_f64 without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (tight serving loops
 * and harnesses that forbid hipMalloc; like elemamd_solve_after_f64 it ends
 * by synchronizing the stream to read the solve's error word).  The chain
 * solves' flags and partial sums live in a fixed per-GPU device array of
 * elemamd_solve_after_f64_noalloc_limit() bytes (9 MiB: 8*n*k bytes of partial
 * sums plus the flags, e.g. n = 32768 with k = 32), held until the process
 * ends; larger n*k returns ELEMAMD_ERR_WORKSPACE before anything is enqueued.
 * Arguments, results and errors are otherwise those of
 * elemamd_solve_after_f64.  Calls that overlap on one GPU share its array and
 * race: serialise them. */
size_t elemamd_solve_after_f64_noalloc_limit(void);
int elemamd_solve_after_f64_noalloc(char uplo, char orientation, int n, int k, const double *A, int64_t lda,
                                    double *B, int64_t ldb, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ELEMENTAL_AMD_H */
