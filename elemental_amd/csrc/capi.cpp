// extern "C" entry points declared in include/elemental_amd.h.
#include "../../include/elemental_amd.h"

#include <algorithm>

#include "batch.hpp"
#include "kernels.hpp"
#include "lapack.hpp"
#include "runtime.hpp"
#include "server.hpp"

using namespace elemamd;

struct elemamd_batch_plan {
    BatchPlan* p;
};
struct elemamd_server {
    Server* s;
};

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }
inline bool is_trans(char c) { c = up(c); return c == 'T' || c == 'C'; }
inline bool valid_orient(char c) { c = up(c); return c == 'N' || c == 'T' || c == 'C'; }

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// B := inv(op(T)) B for an n x n triangular T stored in the lower (upper =
// false) or upper triangle, op = N or T, on the calling stream; `work` holds
// solve_work_bytes<double>(n, k) bytes (null: the per-stream scratch).
// Lower N and upper T run the forward chain, lower T and upper N the
// backward one.  Synchronous: a persistent solve whose bounded wait gave up
// is reported as ELEMAMD_ERR_HIP.
int solve_left(bool upper, bool trans, bool unit, int n, int k, const double* L, i64 ldl, double* B, i64 ldb,
               void* work, hipStream_t s) {
    if (work == nullptr) work = stream_scratch(solve_work_bytes<double>(n, k), s);
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    int* ctrl = solve_ctrl<double>(n, work);
    const bool backward = upper != trans;
    const int rc = solve_apply<double>(!backward, backward, n, k, L, ldl, B, ldb, nullptr, ctrl, nullptr, s, upper,
                                       unit);
    return rc ? rc : solve_read_error(ctrl, s);
}

}  // namespace

extern "C" {

const char* elemamd_version(void) { return "elemental_amd 0.1 (gfx950)"; }

int elemamd_info_read(const int* dinfo, void* stream) {
    if (dinfo == nullptr) return ELEMAMD_ERR_ARG;
    int v = 0;
    if (hipMemcpyAsync(&v, dinfo, sizeof(int), hipMemcpyDeviceToHost, S(stream)) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (hipStreamSynchronize(S(stream)) != hipSuccess) return ELEMAMD_ERR_HIP;
    return v;
}

int elemamd_potrf_unb_f64(int n, double* A, int64_t lda, int* dinfo, void* stream) {
    if (n < 0 || n > 128 || lda < std::max(1, n) || (n > 0 && A == nullptr)) return ELEMAMD_ERR_ARG;
    return launch_potrf_unb<double>(n, A, lda, dinfo, 0, S(stream)) == hipSuccess ? ELEMAMD_OK : ELEMAMD_ERR_HIP;
}

int elemamd_trsm_f64(char side, char uplo, char trans, char diag, int m, int n, double alpha, const double* A,
                     int64_t lda, double* B, int64_t ldb, void* stream) {
    side = up(side);
    uplo = up(uplo);
    diag = up(diag);
    // every argument is validated before anything is enqueued, so an
    // ELEMAMD_ERR_ARG return leaves B untouched
    if ((uplo != 'L' && uplo != 'U') || (diag != 'N' && diag != 'U') || !valid_orient(trans)) return ELEMAMD_ERR_ARG;
    if (side != 'L' && side != 'R') return ELEMAMD_ERR_ARG;
    if (m < 0 || n < 0 || ldb < std::max(1, m)) return ELEMAMD_ERR_ARG;
    const int na = (side == 'L') ? m : n;
    if (lda < std::max(1, na)) return ELEMAMD_ERR_ARG;
    if (m == 0 || n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    const bool upper = uplo == 'U', tr = is_trans(trans), unit = diag == 'U';
    // the Cholesky panel form keeps the reference BLAS column order (bit-exact)
    const bool panel = side == 'R' && !upper && tr && !unit && n <= 128;
    if (alpha != 1.0 && launch_scale_f64(m, n, alpha, B, ldb, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (panel)
        return launch_trsm_rlt<double>(m, n, A, lda, B, ldb, nullptr, kNoLimit, s) == hipSuccess ? ELEMAMD_OK
                                                                                             : ELEMAMD_ERR_HIP;
    if (side == 'L') return solve_left(upper, tr, unit, m, n, A, lda, B, ldb, nullptr, s);
    // X op(T) = B  <=>  op(T)^T X^T = B^T: a left solve of B^T (n x m) with
    // the other orientation, through a transposed copy
    const size_t bt = al256(sizeof(double) * (size_t)n * m);
    char* w = (char*)stream_scratch(bt + solve_work_bytes<double>(n, m), s);
    if (w == nullptr) return ELEMAMD_ERR_HIP;
    double* Bt = (double*)w;
    if (launch_transpose_f64(m, n, B, ldb, Bt, n, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    const int rc = solve_left(upper, !tr, unit, n, m, A, lda, Bt, n, w + bt, s);
    if (rc) return rc;
    return launch_transpose_f64(n, m, Bt, n, B, ldb, s) == hipSuccess ? ELEMAMD_OK : ELEMAMD_ERR_HIP;
}

int elemamd_herk_f64(char uplo, char trans, int n, int k, double alpha, const double* A, int64_t lda, double beta,
                     double* C, int64_t ldc, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(trans)) return ELEMAMD_ERR_ARG;
    const bool tr = is_trans(trans);
    if (n < 0 || k < 0 || lda < std::max(1, tr ? k : n) || ldc < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    if (uplo == 'L' && !tr)  // the Cholesky trailing update's form
        return launch_syrk_ln_f64(n, n, k, alpha, A, lda, A, lda, beta, C, ldc, nullptr, 0, 1, s) == hipSuccess
                   ? ELEMAMD_OK
                   : ELEMAMD_ERR_HIP;
    // op(A) op(A)^T = At^T At with At = op(A)^T (k x n): the TN product on
    // the requested triangle; trans 'N' goes through a transposed copy
    const double* At = A;
    i64 ldat = lda;
    if (!tr) {
        double* w = (double*)stream_scratch(sizeof(double) * (size_t)std::max(1, k) * n, s);
        if (w == nullptr) return ELEMAMD_ERR_HIP;
        if (k > 0 && launch_transpose_f64(n, k, A, lda, w, std::max(1, k), s) != hipSuccess) return ELEMAMD_ERR_HIP;
        At = w;
        ldat = std::max(1, k);
    }
    return launch_gemm_tn_f64(n, n, k, alpha, At, ldat, At, ldat, beta, C, ldc, uplo == 'L' ? 1 : 2, s) ==
                   hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_herk_f32(int mode, int n, int k, float alpha, const float* A, int64_t lda, float beta, float* C,
                     int64_t ldc, void* stream) {
    if (mode != ELEMAMD_TRAILING_F32 && mode != ELEMAMD_TRAILING_BF16X3 && mode != ELEMAMD_TRAILING_BF16X6)
        return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldc < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    if (mode != ELEMAMD_TRAILING_F32 && k > 0) {
        // the factorization's form: split A into bf16 planes once, then the planes update
        const int np = mode == ELEMAMD_TRAILING_BF16X3 ? 2 : 3;
        const i64 ldk = ((i64)k + 31) / 32 * 32;
        const i64 ps = (i64)n * ldk;
        auto* planes = (unsigned short*)stream_scratch((size_t)np * ps * sizeof(unsigned short), s);
        if (planes == nullptr) return ELEMAMD_ERR_HIP;
        if (launch_split_bf16(np, n, k, A, lda, planes, ps, ldk, s) != hipSuccess) return ELEMAMD_ERR_HIP;
        return launch_syrk_ln_bf16_planes(np, n, n, k, alpha, planes, ps, ldk, beta, C, ldc, nullptr, 0, 1, s) ==
                       hipSuccess
                   ? ELEMAMD_OK
                   : ELEMAMD_ERR_HIP;
    }
    return launch_syrk_ln_f32(false, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, nullptr, 0, 1, s) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_trrk_f64(char uplo, char orientA, char orientB, int n, int k, double alpha, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(orientA) || !valid_orient(orientB)) return ELEMAMD_ERR_ARG;
    const bool ta = is_trans(orientA), tb = is_trans(orientB);
    // op(A) is n x k (A: k x n when transposed), op(B) is k x n (B: n x k when transposed)
    if (n < 0 || k < 0 || lda < std::max(1, ta ? k : n) || ldb < std::max(1, tb ? n : k) || ldc < std::max(1, n))
        return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    // the TN product takes op(A)^T and op(B) as k x n column-major operands;
    // the other orientations are transposed into scratch first
    const size_t one = al256(sizeof(double) * (size_t)std::max(1, k) * n);
    char* w = (!ta || tb) ? (char*)stream_scratch(2 * one, s) : nullptr;
    if ((!ta || tb) && w == nullptr) return ELEMAMD_ERR_HIP;
    const int ldk = std::max(1, k);
    const double* Ao = A;
    i64 ldao = lda;
    if (!ta) {
        if (k > 0 && launch_transpose_f64(n, k, A, lda, (double*)w, ldk, s) != hipSuccess) return ELEMAMD_ERR_HIP;
        Ao = (const double*)w;
        ldao = ldk;
    }
    const double* Bo = B;
    i64 ldbo = ldb;
    if (tb) {
        if (k > 0 && launch_transpose_f64(n, k, B, ldb, (double*)(w + one), ldk, s) != hipSuccess)
            return ELEMAMD_ERR_HIP;
        Bo = (const double*)(w + one);
        ldbo = ldk;
    }
    return launch_gemm_tn_f64(n, n, k, alpha, Ao, ldao, Bo, ldbo, beta, C, ldc, uplo == 'L' ? 1 : 2, s) ==
                   hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_trsv_f64(char uplo, char trans, char diag, int n, const double* A, int64_t lda, double* x, int64_t incx,
                     void* stream) {
    uplo = up(uplo);
    diag = up(diag);
    if ((uplo != 'L' && uplo != 'U') || (diag != 'N' && diag != 'U') || !valid_orient(trans) || incx != 1)
        return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    return solve_left(uplo == 'U', is_trans(trans), diag == 'U', n, 1, A, lda, x, std::max(1, n), nullptr, S(stream));
}

int elemamd_cholesky_f64(char uplo, int n, double* A, int64_t lda, int nb, int* dinfo, void* stream) {
    uplo = up(uplo);
    if (uplo != 'L' && uplo != 'U') return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr) return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    if (uplo == 'L') return cholesky_lower<double>(n, A, lda, nb, kTrailingF64, dinfo, s);
    // cholesky::UVar3 on A is cholesky::LVar3 on A^T, operation by operation
    // (real data): transpose in place, factor the lower triangle, transpose
    // back.  The strictly lower triangle of A ends where it started.
    if (launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    const int rc = cholesky_lower<double>(n, A, lda, nb, kTrailingF64, dinfo, s);
    if (launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    return rc;
}

int elemamd_solve_after_f64(char uplo, char orientation, int n, int k, const double* A, int64_t lda, double* B,
                            int64_t ldb, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0) return ELEMAMD_ERR_ARG;
    void* work = stream_scratch(solve_work_bytes<double>(std::max(n, 1), std::max(k, 1)), S(stream));
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    const int rc = solve_after_lower<double>(n, k, A, lda, B, ldb, nullptr, work, S(stream), uplo == 'U');
    if (rc || n == 0 || k == 0) return rc;
    return solve_read_error(solve_ctrl<double>(n, work), S(stream));
}

int elemamd_hpd_solve_f64(char uplo, char orientation, int n, int k, double* A, int64_t lda, double* B, int64_t ldb,
                          int nb, int* dinfo, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldb < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr)
        return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    void* work = stream_scratch(solve_work_bytes<double>(std::max(n, 1), std::max(k, 1)), s);
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    const bool upper = uplo == 'U';
    // UPPER: factor and solve on A^T (its lower triangle holds L = U^T), then
    // transpose back
    if (upper && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    int rc = cholesky_lower<double>(n, A, lda, nb, kTrailingF64, dinfo, s);
    if (rc == kOk) rc = solve_after_lower<double>(n, k, A, lda, B, ldb, dinfo, work, s, false, true);
    if (upper && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (rc || n == 0 || k == 0) return rc;
    return solve_read_error(solve_ctrl<double>(n, work), s);
}

// ------------------------------------------------- inverses and Gemm
int elemamd_gemm_f64(char ta, char tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream) {
    if (!valid_orient(ta) || !valid_orient(tb) || m < 0 || n < 0 || k < 0) return ELEMAMD_ERR_ARG;
    const bool TA = is_trans(ta), TB = is_trans(tb);
    if (lda < std::max(1, TA ? k : m) || ldb < std::max(1, TB ? n : k) || ldc < std::max(1, m))
        return ELEMAMD_ERR_ARG;
    return launch_gemm_f64(TA, TB, m, n, k, alpha, A, lda, kGemmShapeFull, B, ldb, kGemmShapeFull, beta, C, ldc, 0,
                           nullptr, S(stream)) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_hpd_inverse_f64(char uplo, int n, double* A, int64_t lda, int nb, int* dinfo, void* stream) {
    uplo = up(uplo);
    if (uplo != 'L' && uplo != 'U') return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr) return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    void* w = stream_scratch(inverse_scratch_bytes(n, nb), s);
    if (w == nullptr) return ELEMAMD_ERR_HIP;
    // UPPER (hpd_inverse::CholeskyUVar2) on real data is the lower algorithm
    // on A^T: transpose in place, invert, transpose back
    if (uplo == 'U' && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    const int rc = hpd_inverse_lower(n, A, lda, nb, dinfo, w, s);
    if (uplo == 'U' && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    return rc;
}

int elemamd_triangular_inverse_f64(char uplo, char diag, int n, double* A, int64_t lda, int nb, void* stream) {
    uplo = up(uplo);
    diag = up(diag);
    if ((uplo != 'L' && uplo != 'U') || (diag != 'N' && diag != 'U')) return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128) return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    void* w = stream_scratch(inverse_scratch_bytes(n, nb), s);
    if (w == nullptr) return ELEMAMD_ERR_HIP;
    // UPPER: inv(U) = inv(U^T)^T, the lower algorithm on the transpose
    if (uplo == 'U' && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    const int rc = triangular_inverse_lower(n, A, lda, diag == 'U', nb, w, s);
    if (uplo == 'U' && launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    return rc;
}

int elemamd_hermitian_gen_definite_eig_f64(int type, char uplo, int n, double* A, int64_t lda, double* B,
                                           int64_t ldb, double* w, double* X, int64_t ldx, int range, int il, int iu,
                                           double vl, double vu, int* nev, int* info, void* stream) {
    uplo = up(uplo);
    if (type < ELEMAMD_AXBX || type > ELEMAMD_BAX || (uplo != 'L' && uplo != 'U')) return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n) || ldb < std::max(1, n) || nev == nullptr || info == nullptr)
        return ELEMAMD_ERR_ARG;
    if ((n > 0 && w == nullptr) || (X != nullptr && ldx < std::max(1, n))) return ELEMAMD_ERR_ARG;
    if (range == ELEMAMD_EIG_INDEX) {
        if (n > 0 && (il < 0 || il > iu || iu >= n)) return ELEMAMD_ERR_ARG;
    } else if (range == ELEMAMD_EIG_VALUE) {
        if (!(vl < vu)) return ELEMAMD_ERR_ARG;
    } else if (range != ELEMAMD_EIG_ALL) {
        return ELEMAMD_ERR_ARG;
    }
    *nev = 0;
    *info = 0;
    if (n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    void* work = stream_scratch(gen_eig_scratch_bytes(n), s);
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    // UPPER on real data: the same reduction with U = L^T, run on the
    // transposes of A and B and transposed back (the factor and the reduced
    // matrix end in the upper triangles)
    const bool upper = uplo == 'U';
    if (upper && (launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess ||
                  launch_transpose_inplace_f64(n, B, ldb, s) != hipSuccess))
        return ELEMAMD_ERR_HIP;
    const int rc =
        hermitian_gen_definite_eig_lower(type, n, A, lda, B, ldb, w, X, ldx, range, il, iu, vl, vu, nev, info, work, s);
    if (upper && (launch_transpose_inplace_f64(n, A, lda, s) != hipSuccess ||
                  launch_transpose_inplace_f64(n, B, ldb, s) != hipSuccess))
        return ELEMAMD_ERR_HIP;
    if (upper && hipStreamSynchronize(s) != hipSuccess) return ELEMAMD_ERR_HIP;
    return rc;
}

// ------------------------------------------------------- Matrix<float>
// Cholesky / SolveAfter / HPDSolve with F = float: the float factorization
// (FP64 diagonal blocks, FP32 panels by products with the diagonal pieces'
// inverses, FP32 MFMA trailing updates) and float solves with the 64 x 64
// diagonal-block inverses.
int elemamd_cholesky_f32(char uplo, int n, float* A, int64_t lda, int nb, int* dinfo, void* stream) {
    uplo = up(uplo);
    if (uplo != 'L' && uplo != 'U') return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr) return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    if (uplo == 'U' && launch_transpose_inplace<float>(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    const int rc = cholesky_lower<float>(n, A, lda, nb, kTrailingF32, dinfo, s);
    if (uplo == 'U' && launch_transpose_inplace<float>(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    return rc;
}

namespace {
// float SolveAfter with the factor in the lower triangle of L; `extra` bytes
// of the per-stream scratch precede the solve's own work (for a transposed copy)
int solve_after_f32_lower(int n, int k, const float* L, i64 ldl, float* B, i64 ldb, const int* dinfo, char* work,
                          hipStream_t s) {
    const int rc = solve_after_lower<float>(n, k, L, ldl, B, ldb, dinfo, work, s);
    if (rc || n == 0 || k == 0) return rc;
    return solve_read_error(solve_ctrl<float>(n, work), s);
}
}  // namespace

int elemamd_solve_after_f32(char uplo, char orientation, int n, int k, const float* A, int64_t lda, float* B,
                            int64_t ldb, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0 || k == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    // UPPER: the solves run on a transposed copy (L = U^T in its lower triangle)
    const size_t tb = uplo == 'U' ? al256(sizeof(float) * (size_t)n * n) : 0;
    char* w = (char*)stream_scratch(tb + solve_work_bytes<float>(n, k), s);
    if (w == nullptr) return ELEMAMD_ERR_HIP;
    const float* L = A;
    i64 ldl = lda;
    if (uplo == 'U') {
        float* Lt = (float*)w;
        if (hipMemcpy2DAsync(Lt, sizeof(float) * n, A, sizeof(float) * lda, sizeof(float) * n, n,
                             hipMemcpyDeviceToDevice, s) != hipSuccess ||
            launch_transpose_inplace<float>(n, Lt, n, s) != hipSuccess)
            return ELEMAMD_ERR_HIP;
        L = Lt;
        ldl = n;
    }
    return solve_after_f32_lower(n, k, L, ldl, B, ldb, nullptr, w + tb, s);
}

int elemamd_hpd_solve_f32(char uplo, char orientation, int n, int k, float* A, int64_t lda, float* B, int64_t ldb,
                          int nb, int* dinfo, void* stream) {
    uplo = up(uplo);
    if ((uplo != 'L' && uplo != 'U') || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldb < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr)
        return ELEMAMD_ERR_ARG;
    hipStream_t s = S(stream);
    char* w = (char*)stream_scratch(solve_work_bytes<float>(std::max(n, 1), std::max(k, 1)), s);
    if (w == nullptr) return ELEMAMD_ERR_HIP;
    const bool upper = uplo == 'U';
    if (upper && launch_transpose_inplace<float>(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    int rc = cholesky_lower<float>(n, A, lda, nb, kTrailingF32, dinfo, s);
    if (rc == kOk && n > 0 && k > 0) rc = solve_after_lower<float>(n, k, A, lda, B, ldb, dinfo, w, s);
    if (upper && launch_transpose_inplace<float>(n, A, lda, s) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (rc || n == 0 || k == 0) return rc;
    return solve_read_error(solve_ctrl<float>(n, w), s);
}

static MixedOpts to_opts(const elemamd_ir_opts* o) {
    MixedOpts m;
    if (o) {
        m.nb = o->nb > 0 ? o->nb : 128;
        m.trailing = o->trailing;
        m.max_iter = o->max_iter > 0 ? o->max_iter : 30;
        m.gmres_restart = o->gmres_restart >= 0 ? (o->gmres_restart == 0 ? 20 : o->gmres_restart) : 0;
        m.force_gmres = o->force_gmres;
        m.allow_shift = o->allow_shift;
        m.tol = o->tol > 0.0 ? o->tol : 1e-14;
        m.refine = o->refine;
        m.fallback = o->fallback;
    }
    return m;
}

size_t elemamd_hpd_solve_mixed_workspace(int n, int k, const elemamd_ir_opts* opts) {
    if (n < 0 || k < 0) return 0;
    return mixed_workspace_bytes(n, k, to_opts(opts));
}

int elemamd_hpd_solve_mixed_f64(int n, int k, const double* A, int64_t lda, double* B, int64_t ldb,
                                const elemamd_ir_opts* opts, void* work, size_t work_bytes, elemamd_ir_report* report,
                                void* stream) {
    MixedOpts o = to_opts(opts);
    if (o.trailing != kTrailingF32 && o.trailing != kTrailingBf16x3 && o.trailing != kTrailingBf16x6)
        return ELEMAMD_ERR_ARG;
    if (o.refine != ELEMAMD_REFINE_IR && o.refine != ELEMAMD_REFINE_CG) return ELEMAMD_ERR_ARG;
    MixedReport r{};
    int rc = hpd_solve_mixed(n, k, A, lda, B, ldb, o, work, work_bytes, &r, S(stream));
    if (report) {
        report->info = r.info;
        report->lowprec_info = r.lowprec_info;
        report->iterations = r.iterations;
        report->gmres_iterations = r.gmres_iterations;
        report->used_gmres = r.used_gmres;
        report->shifted = r.shifted;
        report->converged = r.converged;
        report->backward_error = r.backward_error;
        report->anorm_inf = r.anorm_inf;
        report->fallback = r.fallback;
    }
    return rc;
}

int elemamd_fill_normal_f64(double* out, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, int64_t offset,
                            void* stream) {
    if (rows < 0 || cols < 0 || ld < std::max<int64_t>(1, rows)) return ELEMAMD_ERR_ARG;
    return launch_fill_normal_f64(out, rows, cols, ld, seed, offset, S(stream)) == hipSuccess ? ELEMAMD_OK
                                                                                             : ELEMAMD_ERR_HIP;
}

int elemamd_sym_residual_f64(int n, int k, const double* A, int64_t lda, const double* X, int64_t ldx,
                             const double* B, int64_t ldb, double* R, int64_t ldr, void* stream) {
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldx < std::max(1, n) || ldb < std::max(1, n) ||
        ldr < std::max(1, n))
        return ELEMAMD_ERR_ARG;
    double* part = nullptr;
    const size_t pc = symm_part_count(n, k);
    if (pc > 0 && (part = (double*)stream_scratch(pc * sizeof(double), S(stream))) == nullptr) return ELEMAMD_ERR_HIP;
    return launch_symm_lower_residual(n, k, A, lda, X, ldx, B, ldb, R, ldr, part, S(stream)) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_batch_plan_create(int count, const int* n, const int* k, int precision, int nbuckets,
                              const int* bucket_max_n, const int* bucket_nb, elemamd_batch_plan** plan) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    *plan = nullptr;
    BatchPlan* P = nullptr;
    const int rc = batch_plan_create(count, n, k, precision, nbuckets, bucket_max_n, bucket_nb, &P);
    if (rc != kOk) return rc;
    *plan = new elemamd_batch_plan{P};
    return ELEMAMD_OK;
}

int elemamd_batch_set_launch(elemamd_batch_plan* plan, const int* bucket_grid) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    return batch_set_launch(plan->p, bucket_grid);
}

int elemamd_batch_plan_destroy(elemamd_batch_plan* plan) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    delete plan->p;
    delete plan;
    return ELEMAMD_OK;
}

int elemamd_batch_plan_layout(const elemamd_batch_plan* plan, int64_t* a_off, int64_t* b_off, int64_t* a_total,
                              int64_t* b_total) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    const BatchPlan* P = plan->p;
    for (int i = 0; i < P->count; ++i) {
        if (a_off) a_off[i] = P->a_off[i];
        if (b_off) b_off[i] = P->b_off[i];
    }
    if (a_total) *a_total = P->a_total;
    if (b_total) *b_total = P->b_total;
    return ELEMAMD_OK;
}

int elemamd_batch_set_refinement(elemamd_batch_plan* plan, int sweeps, double tol) {
    if (plan == nullptr || sweeps < 0 || !(tol > 0.0)) return ELEMAMD_ERR_ARG;
    BatchPlan* P = plan->p;
    if (P->ir_sweeps != sweeps || P->tol != tol) {
        P->ir_sweeps = sweeps;
        P->tol = tol;
        P->g_A = nullptr;  // the captured sequence no longer matches: capture again
    }
    return ELEMAMD_OK;
}

size_t elemamd_batch_workspace(const elemamd_batch_plan* plan) {
    return plan == nullptr ? 0 : batch_workspace_bytes(plan->p);
}

int elemamd_batch_hpd_solve(elemamd_batch_plan* plan, double* A, double* B, void* work, size_t work_bytes,
                            int use_graph, void* stream) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    return batch_hpd_solve(plan->p, A, B, work, work_bytes, use_graph != 0, S(stream));
}

int elemamd_batch_results(const elemamd_batch_plan* plan, int* info, double* backward_error, int* iterations,
                          void* stream) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    return batch_results(plan->p, info, backward_error, iterations, S(stream));
}

int elemamd_batch_backward_error(const elemamd_batch_plan* plan, const double* A, const double* X, const double* B,
                                 double* R, double* eta, void* stream) {
    if (plan == nullptr || A == nullptr || X == nullptr || B == nullptr || R == nullptr) return ELEMAMD_ERR_ARG;
    return batch_backward_error(plan->p, A, X, B, R, eta, S(stream));
}

int elemamd_server_create(const elemamd_server_config* cfg, elemamd_server** server) {
    if (cfg == nullptr || server == nullptr) return ELEMAMD_ERR_ARG;
    *server = nullptr;
    ServerConfig c;
    c.slots = cfg->slots;
    c.precision = cfg->precision;
    c.nb = cfg->nb > 0 ? cfg->nb : 128;
    c.tol = cfg->tol > 0.0 ? cfg->tol : 1e-14;
    c.memory_budget = cfg->memory_budget;
    auto* S = new Server(c);
    const int rc = S->start();
    if (rc != kOk) {
        delete S;
        return rc;
    }
    *server = new elemamd_server{S};
    return ELEMAMD_OK;
}

int elemamd_server_submit(elemamd_server* server, int key, int reuse, int n, int k, const double* A, int64_t lda,
                          double* B, int64_t ldb, void* ready_stream, int64_t* ticket) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    ServeRequest q{key, reuse, n, k, A, lda, B, ldb, S(ready_stream)};
    return server->s->submit(q, ticket);
}

int elemamd_server_wait(elemamd_server* server, int64_t ticket, elemamd_serve_result* result) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    ServeResult r{};
    const int rc = server->s->wait(ticket, &r);
    if (rc != kOk) return rc;
    if (result) {
        result->ticket = r.ticket;
        result->status = r.status;
        result->info = r.info;
        result->factored = r.factored;
        result->iterations = r.iterations;
        result->gmres_iterations = r.gmres_iterations;
        result->slot = r.slot;
        result->backward_error = r.backward_error;
        result->t_submit = r.t_submit;
        result->t_start = r.t_start;
        result->t_done = r.t_done;
    }
    return ELEMAMD_OK;
}

int elemamd_server_evict(elemamd_server* server, int key) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    return server->s->evict(key);
}

int elemamd_server_stats(elemamd_server* server, elemamd_server_statistics* stats) {
    if (server == nullptr || stats == nullptr) return ELEMAMD_ERR_ARG;
    const ServerStats t = server->s->stats();
    stats->completed = t.completed;
    stats->reuse_hits = t.reuse_hits;
    stats->evictions = t.evictions;
    stats->peak_reserved = t.peak_used;
    stats->reserved = t.used;
    stats->budget = t.budget;
    stats->resident_factors = t.resident;
    return ELEMAMD_OK;
}

int elemamd_server_destroy(elemamd_server* server) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    delete server->s;
    delete server;
    return ELEMAMD_OK;
}

}  // extern "C"
