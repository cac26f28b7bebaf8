// Launchers of the device kernels (each in its own .hip translation unit).
#pragma once

#include "common.hpp"

namespace elemamd {

// Factorization kernels skip their work when *info != 0 && *info <= limit
// (info is the 1-based failing column): a kernel belonging to the step whose
// diagonal block ends at column c passes limit = c, so a failure in a later
// step never cancels it and A ends in the state the reference leaves.
constexpr int kNoLimit = 0x7fffffff;

// potrf.hip -- LVar3Unb on an n x n (n <= 128) diagonal block starting at
// global column col_offset (skips if an earlier column failed).
template <typename T>
hipError_t launch_potrf_unb(int n, T* A, i64 lda, int* info, int col_offset, hipStream_t s);

// trsm_panel.hip -- B := B * inv(L)^T, L nb x nb lower (nb <= 128), B m x nb.
template <typename T>
hipError_t launch_trsm_rlt(int m, int nb, const T* L, i64 ldl, T* B, i64 ldb, const int* info, int limit,
                           hipStream_t s);


// panel_f64.hip -- the FP64 factorization's panel solve B := B L^{-T}
// (b <= 128): substitution inside 32-column pieces, FP64 MFMA products
// between them, rows in registers (not bit-identical to launch_trsm_rlt).
hipError_t launch_panel_f64(int m, int b, const double* L, i64 ldl, double* B, i64 ldb, const int* info, int limit,
                            hipStream_t s);

// potrf_f32.hip -- panel step of the float-storage factor: the b x b
// (b <= 128) diagonal block factored in place (FP64 on 32 x 32 pieces, whose
// inverses go to Xd, 4 x 32 x 32 floats), and B (m x b) := B L^{-T} with the
// factored block L and those inverses; planes (may be null): the result's np
// bf16 planes (launch_split_bf16's form) are also written, row r of B to
// planes + p*ps + r*ldk (columns 0..b-1).  upd (may be null): B first takes
// B -= P_B P_top^T over uk columns of the same np planes, P_top = upd rows
// 0..b-1 and P_B = upd rows b..b+m-1 (the block column's pending in-group
// update; the planes at upd must not overlap B's output planes).
hipError_t launch_potrf_f32(int b, float* A, i64 lda, float* Xd, int* info, int col_offset, hipStream_t s);
hipError_t launch_panel_f32(int m, int b, const float* L, i64 ldl, const float* Xd, float* B, i64 ldb,
                            const int* info, int limit, unsigned short* planes, int np, i64 ps, i64 ldk,
                            hipStream_t s, const unsigned short* upd = nullptr, int uk = 0);

// syrk.hip -- C := beta*C + alpha*A*B^T on the entries j <= i of C (m x n, n <= m).
// With info != null the k columns are factor columns [kcol0, kcol0 + k) and a
// recorded failure drops the panels of its nbp-block and later ones.
// beside_chain: the launch runs concurrently with the factorization's chain
// kernels; its workgroups reserve extra LDS so that room for them stays.
hipError_t launch_syrk_ln_f64(int m, int n, int k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                              double beta, double* C, i64 ldc, const int* info, int kcol0, int nbp, hipStream_t s,
                              bool beside_chain = false);
// float storage; bf16x3 = false: FP32 MFMA, true: split-BF16x3 MFMA with FP32 accumulation.
hipError_t launch_syrk_ln_f32(bool bf16x3, int m, int n, int k, float alpha, const float* A, i64 lda, const float* B,
                              i64 ldb, float beta, float* C, i64 ldc, const int* info, int kcol0, int nbp,
                              hipStream_t s);

// Split-BF16 from pre-split panels: np bf16 planes of the panel (plane p at
// planes + p*ps, row-major, row stride ldk a multiple of 32), made by
// launch_split_bf16 from a float m x k panel; rows of the planes = rows (and
// columns) of C.  np = 2: split-BF16x3 (three MFMA terms), np = 3:
// split-BF16x6 (six terms, down to 2^-24 of a product).
hipError_t launch_split_bf16(int np, int m, int k, const float* P, i64 ldp, unsigned short* planes, i64 ps, i64 ldk,
                             hipStream_t s);
// chain = 1: the launch is on the factorization's critical chain (its waves
// take issue priority over other work sharing a SIMD).
hipError_t launch_syrk_ln_bf16_planes(int np, int m, int n, int k, float alpha, const unsigned short* planes, i64 ps,
                                      i64 ldk, float beta, float* C, i64 ldc, const int* info, int kcol0, int nbp,
                                      hipStream_t s, int chain = 0);

// The in-group update of one block column's diagonal block (ncol <= 128),
// with many workgroups: C (ncol x ncol, lower) -= P P^T, P = the first ncol
// rows of the planes, k columns (the group's earlier panels); the rows below
// it take theirs inside launch_panel_f32.
hipError_t launch_syrk_diag_bf16_planes(int np, int ncol, int k, const unsigned short* planes, i64 ps, i64 ldk,
                                        float* C, i64 ldc, const int* info, int kcol0, int nbp, hipStream_t s);

// gemm.hip -- C := beta*C + alpha*op(A)*op(B), op(A) m x k, op(B) k x n;
// ta: A stored k x m (op = T), tb: B stored n x k (op = T).
// tri: 0 all of C; 1 its lower triangle; 2 its upper triangle (1, 2: m == n).
// Operand shapes mask the stored matrix: outside the triangle reads as zero,
// a unit diagonal as one.  info (may be null): the launch does nothing when
// *info != 0.  beta == 0: C is not read.
enum : int {
    kGemmShapeFull = 0,
    kGemmShapeLower = 1,
    kGemmShapeLowerUnit = 2,
    kGemmShapeUpper = 3,
    kGemmShapeUpperUnit = 4,
};
hipError_t launch_gemm_f64(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, i64 lda, int ashape,
                           const double* B, i64 ldb, int bshape, double beta, double* C, i64 ldc, int tri,
                           const int* info, hipStream_t s);
// the TN form without masks (Trrk)
hipError_t launch_gemm_tn_f64(int m, int n, int k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                              double beta, double* C, i64 ldc, int tri, hipStream_t s);

// inverse.hip -- n x n (n <= 128) blocks by one workgroup, bit-identical to
// the reference's unblocked loops in IEEE double (no contraction).
// triangular_inverse::LVar3Unb on the lower triangle of L: X (ldx, full n x n:
// zeros above the diagonal, ones on it for a unit diagonal) and, when Lout
// is not null, the lower triangle of Lout (strictly lower for unit) receive
// inv(L).  L and Lout may be the same block.
hipError_t launch_trtri_unb_f64(int n, const double* L, i64 ldl, bool unit, double* X, i64 ldx, double* Lout,
                                i64 ldo, const int* info, hipStream_t s);
// The same inverse by recursive doubling (7 levels of short products, not
// the reference's operation order): for the diagonal blocks of the blocked
// algorithms.  X (ldx >= 128) receives the full 128 x 128 inverse (the
// identity past n), Lout (may be null) the lower triangle.
hipError_t launch_trtri_fast_f64(int n, const double* L, i64 ldl, bool unit, double* X, i64 ldx, double* Lout,
                                 i64 ldo, const int* info, hipStream_t s);
// internal::TrtrmmLUnblocked(TRANSPOSE) on the lower triangle of X: the lower
// triangle of C := X^T X.
hipError_t launch_trtrmm_unb_f64(int n, const double* X, i64 ldx, double* C, i64 ldc, const int* info, hipStream_t s);

// solve.hip -- persistent triangular solves with precomputed inverses of the
// 64x64 diagonal blocks.  Dinv holds solve_dinv_count(n) elements, ctrl
// solve_ctrl_count(n, kc) ints (ctrl[1] != 0 after a solve means a bounded wait
// expired).
size_t solve_ctrl_count(int n, int kc);
// solve_chain.hip -- chain + helpers triangular solve (forward, or backward with L^T)
size_t chain_ctrl_count(int n, int kc);
// trans = false: the forward solve B := inv(L) B; trans = true: the backward
// solve B := inv(L^T) B (blocks in reverse order).  upper: the factor is
// stored in the upper triangle as U = L^T, so the forward solve is with U^T
// and the backward one with U.  unit: unit diagonal (never read).
template <typename T>
hipError_t launch_trsm_chain(bool trans, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv, int* ctrl,
                             T* part, const int* info, hipStream_t s, bool upper = false, bool unit = false);
size_t solve_dinv_count(int n);
template <typename T>
hipError_t launch_trtri_diag64(int n, const T* L, i64 ldl, T* Dinv, hipStream_t s);
// the 64-blocks I0 <= I < I1 only
template <typename T>
hipError_t launch_trtri_diag64_range(int n, int I0, int I1, const T* L, i64 ldl, T* Dinv, hipStream_t s);
template <typename T>
hipError_t launch_trsm_persist(bool trans, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv,
                               int* ctrl, const int* info, hipStream_t s);

// solve_sweep.hip -- block-inverse sweeps for the float correction solves of
// refinement.  Preparation (once per factor S, lower triangle = L): writes the
// mirror of L into the strict upper triangle of S, the diagonal-block
// inverses into Xp (sweep_xinv_count floats, scratch) and the sweep operands
// into tiles (sweep_tiles_count floats).  Dinv: launch_trtri_diag64 output.
size_t sweep_tiles_count(int n);
size_t sweep_xinv_count(int n);
size_t sweep_work_bytes(int n);
// The strict upper triangle of S := the mirror of its lower one (the
// backward sweep's operand); it touches nothing launch_sweep_prepare or
// the diagonal-block inverses read, so it may run beside them.
hipError_t launch_sweep_mirror(int n, float* S, i64 lds, hipStream_t s);
hipError_t launch_sweep_prepare(int n, const float* S, i64 lds, const float* Dinv, float* Xp, float* tiles,
                                hipStream_t s);
// Column blocks of the sweeps (512): block b's inverse and sweep operands
// (both directions) read only S's columns of block b (lower triangle) and the
// diagonal-block inverses of its rows, so they may be made as soon as those
// columns of the factor are final.
int sweep_block_cols();
hipError_t launch_sweep_prepare_blocks(int n, int b0, int b1, const float* S, i64 lds, const float* Dinv, float* Xp,
                                       float* tiles, hipStream_t s);
// B := inv(L^T) inv(L) B for B n x k (column-major); work: sweep_work_bytes(n).
hipError_t launch_sweep_solve(int n, int k, const float* S, i64 lds, const float* tiles, float* B, i64 ldb, void* work,
                              hipStream_t s);
// The work's error word: set by a sweep whose bounded wait expired (its
// result is then wrong), kept until cleared.
hipError_t launch_sweep_clear_error(int n, void* work, hipStream_t s);
hipError_t sweep_read_error(int n, void* work, int* out, hipStream_t s);

// blas1.hip -- elementwise helpers, conversions, residual and norms.
hipError_t launch_fill_int(int* p, int value, int count, hipStream_t s);
hipError_t launch_fill_normal_f64(double* out, i64 rows, i64 cols, i64 ld, uint64_t seed, i64 offset, hipStream_t s);
hipError_t launch_scale_f64(int m, int n, double alpha, double* A, i64 lda, hipStream_t s);
// A := A^T in place (n x n); B (n x m) := A^T (A m x n).
hipError_t launch_transpose_inplace_f64(int n, double* A, i64 lda, hipStream_t s);
template <typename T>
hipError_t launch_transpose_inplace(int n, T* A, i64 lda, hipStream_t s);
hipError_t launch_transpose_f64(int m, int n, const double* A, i64 lda, double* B, i64 ldb, hipStream_t s);
// F (n x n) := the symmetric matrix stored in A's lower triangle; or A's
// lower triangle with zeros above.  A's strictly upper triangle is not read.
hipError_t launch_symmetrize_lower_f64(int n, const double* A, i64 lda, double* F, i64 ldf, hipStream_t s);
hipError_t launch_copy_lower_zero_upper_f64(int n, const double* A, i64 lda, double* F, i64 ldf, hipStream_t s);
hipError_t launch_copy_lower_f64_to_f32(int n, const double* A, i64 lda, float* S, i64 lds, hipStream_t s);
// columns j0 <= j < j1 only
hipError_t launch_copy_lower_f64_to_f32_cols(int n, int j0, int j1, const double* A, i64 lda, float* S, i64 lds,
                                            hipStream_t s);
hipError_t launch_copy_f64_to_f32_scaled(int m, int n, const double* X, i64 ldx, const double* colscale, float* Y,
                                         i64 ldy, hipStream_t s);
hipError_t launch_axpy_f32_to_f64_scaled(int m, int n, const float* D, i64 ldd, const double* colscale, double* X,
                                         i64 ldx, hipStream_t s);
hipError_t launch_copy_f64(int m, int n, const double* X, i64 ldx, double* Y, i64 ldy, hipStream_t s);
// R := B - A*X with A symmetric (lower triangle read), n x n; X, B, R n x k.
// `part` (symm_part_count(n, k) doubles, or null for one unsplit sweep)
// holds the partial sums of the split tile sweep; they are added in a fixed
// order, so the result does not depend on scheduling.
int symm_splits(int n, int k);
size_t symm_part_count(int n, int k);
hipError_t launch_symm_lower_residual(int n, int k, const double* A, i64 lda, const double* X, i64 ldx,
                                      const double* B, i64 ldb, double* R, i64 ldr, double* part, hipStream_t s);
// Y := A*X with A symmetric (lower triangle read).
hipError_t launch_symm_lower(int n, int k, const double* A, i64 lda, const double* X, i64 ldx, double* Y, i64 ldy,
                             double* part, hipStream_t s);
// colmax[c] = max_i |X(i,c)| for c < k.
hipError_t launch_col_absmax(int m, int k, const double* X, i64 ldx, double* colmax, hipStream_t s);
// rowsum[i] = sum_j |A(i,j)| of the symmetric matrix stored in the lower
// triangle; part: symm_part_count(n, 1) doubles or null.
hipError_t launch_sym_lower_abs_rowsum(int n, const double* A, i64 lda, double* rowsum, double* part, hipStream_t s);
// out[0] = max_i v[i]
hipError_t launch_vec_max(int n, const double* v, double* out, hipStream_t s);

hipError_t launch_batched_dots(int n, int k, int nv, const double* V, i64 ldv, i64 vstride, const double* W, i64 ldw,
                               double* out, hipStream_t s);
hipError_t launch_batched_gemv(int n, int k, int nv, double alpha, const double* V, i64 ldv, i64 vstride,
                               const double* h, double beta, double* W, i64 ldw, hipStream_t s);
hipError_t launch_col_scale(int n, int k, const double* Y, i64 ldy, const double* sc, int invert, double* X, i64 ldx,
                            hipStream_t s);
hipError_t launch_copy_f32_to_f64_scaled(int m, int n, const float* D, i64 ldd, const double* colscale, double* Y,
                                         i64 ldy, hipStream_t s);
hipError_t launch_axpy_f64(int m, int n, double alpha, const double* X, i64 ldx, double* Y, i64 ldy, hipStream_t s);
// Preconditioned CG refinement steps (blocks n-row, ld m): X += alpha P and
// R -= alpha Q with alpha = rz/pq per active column; P := Z - (zq/pq) P.
hipError_t launch_cg_xr(int m, int n, const double* rz, const double* pq, const int* active, const double* P,
                        const double* Q, double* X, double* R, hipStream_t s);
hipError_t launch_cg_p(int m, int n, const double* zq, const double* pq, const double* Z, double* P, hipStream_t s);
hipError_t launch_shift_diag(int n, const double* A, i64 lda, float* S, i64 lds, double sigma, hipStream_t s);

}  // namespace elemamd
