// Host drivers of the path (internal C++ interface under the C ABI).
#pragma once

#include <functional>

#include "common.hpp"

namespace elemamd {

enum : int { kTrailingF64 = 0, kTrailingF32 = 1, kTrailingBf16x3 = 2, kTrailingBf16x6 = 3 };
// Refinement of the mixed solve: classic iterative refinement (GMRES-IR on
// stagnation) or preconditioned CG on A with the low-precision factor as the
// preconditioner and FP64 products, residuals and iterates (classic IR takes
// over if it stagnates).
enum : int { kRefineIR = 0, kRefineCG = 1 };

// Columns of panels applied per pass over the trailing matrix (a multiple of
// the block size nb is used: max(1, depth / nb) panels).
constexpr int kUpdateDepthF64 = 256;
constexpr int kUpdateDepthF32 = 512;
constexpr int kUpdateDepthBf16x3 = 512;

// cholesky::LVar3 on device memory; T = double (FP64 path) or float (the
// low-precision factor, diagonal blocks still factored in double).
// trailing_ready (may be null): an event after which the columns right of
// the first group of block columns hold their input (they may still be
// being written when the call starts; the first update that reads them
// waits for it).
// after_group (may be null): called on the host once the launches of each
// group of block columns are enqueued, with the group's end column c1; work
// ordered after that point of s sees columns [0, c1) of the factor final.
// scratch (may be null): lowprec_scratch_bytes(n, nb, trailing) bytes of
// device memory for the float factor's bf16 planes and diagonal-piece
// inverses; null takes them from the per-stream scratch buffer.
template <typename T>
int cholesky_lower(int n, T* A, i64 lda, int nb, int trailing, int* dinfo, hipStream_t s,
                   hipEvent_t trailing_ready = nullptr, const std::function<int(int)>* after_group = nullptr,
                   void* scratch = nullptr);
size_t lowprec_scratch_bytes(int n, int nb, int trailing);

// hpd_inverse::CholeskyLVar2 / triangular_inverse::LVar3 on the lower
// triangle (hpd_inverse.cpp); scratch: inverse_scratch_bytes(n, nb) bytes.
size_t inverse_scratch_bytes(int n, int nb);
int hpd_inverse_lower(int n, double* A, i64 lda, int nb, int* dinfo, void* scratch, hipStream_t s);
int triangular_inverse_lower(int n, double* L, i64 ldl, bool unit, int nb, void* scratch, hipStream_t s);

// elem::HermitianGenDefiniteEig on the lower triangles of A and B
// (gen_eig.cpp).  type 1 AXBX, 2 ABX, 3 BAX; range 0 all, 1 indices
// [il, iu] (0-based), 2 values in (vl, vu].  w: n doubles (device); X
// (device, may be null): n x nev eigenvectors.  *nev: eigenvalues found;
// *info: B's failing column (1-based, 0 = none; nothing else is done).
// scratch: gen_eig_scratch_bytes(n).  Synchronous.
size_t gen_eig_scratch_bytes(int n);
int hermitian_gen_definite_eig_lower(int type, int n, double* A, i64 lda, double* B, i64 ldb, double* w, double* X,
                                     i64 ldx, int range, int il, int iu, double vl, double vu, int* nev, int* info,
                                     void* scratch, hipStream_t s);

// cholesky::SolveAfter(LOWER, NORMAL): B := inv(L^T) inv(L) B.
// `work` holds solve_work_bytes<T>(n, k) bytes of device memory.
template <typename T>
size_t solve_work_bytes(int n, int k);
// upper: the factor is stored in the upper triangle as U = L^T
// (cholesky::SolveAfter(UPPER, ...): the forward solve with U^T, the backward
// one with U).  inverse_blocks: the 64 x 64 diagonal blocks are applied as
// products with their inverses (computed first) instead of by substitution
// (always for float; for double the faster form that HPDSolve uses, while
// the SolveAfter entry point keeps the substitution, exact on integer
// problems); lower-stored factors only.
template <typename T>
int solve_after_lower(int n, int k, const T* L, i64 ldl, T* B, i64 ldb, const int* dinfo, void* work, hipStream_t s,
                      bool upper = false, bool inverse_blocks = sizeof(T) == 4);
// The control words of solve_after_lower inside its `work` buffer.
template <typename T>
int* solve_ctrl(int n, void* work);
// Synchronize `s` and report a persistent solve whose bounded wait gave up
// (kErrHip; the solution is then wrong) since solve_apply cleared ctrl[1].
int solve_read_error(const int* ctrl, hipStream_t s);
// The two halves of it for repeated solves with one factor: the diagonal
// block inverses are computed once (solve_prepare) and reused (solve_apply).
template <typename T>
int solve_prepare(int n, const T* L, i64 ldl, T* Dinv, hipStream_t s);
template <typename T>
int solve_apply(bool forward, bool backward, int n, int k, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv,
                int* ctrl, const int* dinfo, hipStream_t s, bool upper = false, bool unit = false);

struct MixedOpts {
    int nb = 128;
    int trailing = kTrailingF32;
    int max_iter = 30;
    int gmres_restart = 20;
    int force_gmres = 0;
    int allow_shift = 1;
    double tol = 1e-14;
    int refine = kRefineIR;
    // finish in plain FP64 (Cholesky + SolveAfter on a copy of A) when the
    // low-precision route fails
    int fallback = 1;
};

struct MixedReport {
    int info;
    int lowprec_info;
    int iterations;
    int gmres_iterations;
    int used_gmres;
    int shifted;
    int converged;
    double backward_error;
    double anorm_inf;
    int fallback;
};

// Low-precision factor kept for repeated solves (device memory of
// mixed_factor_bytes(n), laid out by carve_factor).
struct MixedFactor {
    float* S;         // n x n float factor (ld n): L below the diagonal, its mirror above
    float* Dinv;      // inverses of its 64x64 diagonal blocks
    float* Xp;        // inverses of its 256x256 diagonal blocks (sweep preparation scratch)
    float* tiles;     // operands of the block-inverse sweeps
    int* dinfo;       // device status word
    double* scratch;  // n + 8 doubles, then the split row-sum partials
    double anorm;     // ||A||_inf
    int lowprec_info;
    int shifted;
};
size_t mixed_factor_bytes(int n);
MixedFactor carve_factor(void* mem, int n);
// scratch: lowprec_scratch_bytes(n, o.nb, o.trailing) bytes (null: the
// per-stream scratch buffer).
int mixed_factor(int n, const double* A, i64 lda, const MixedOpts& o, MixedFactor& f, hipStream_t s,
                 void* scratch = nullptr);
size_t mixed_refine_bytes(int n, int k, const MixedOpts& o);
int mixed_refine(int n, int k, const double* A, i64 lda, double* B, i64 ldb, const MixedOpts& o,
                 const MixedFactor& f, void* work, size_t work_bytes, MixedReport* rep, hipStream_t s);

size_t mixed_workspace_bytes(int n, int k, const MixedOpts& o);
int hpd_solve_mixed(int n, int k, const double* A, i64 lda, double* B, i64 ldb, const MixedOpts& o, void* work,
                    size_t work_bytes, MixedReport* rep, hipStream_t s);

}  // namespace elemamd
