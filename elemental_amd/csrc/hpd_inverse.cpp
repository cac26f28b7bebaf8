// Host drivers of the direct callers of the Cholesky factor that invert:
// the HPD inverse (hpd_inverse::CholeskyLVar2) and the triangular inverse
// (triangular_inverse::LVar3).  Everything is enqueued on the caller's
// stream.  The diagonal blocks' Cholesky factor is the bit-exact LVar3Unb
// kernel (potrf.hip); their inverse and Trtrmm are the reference's unblocked
// loops bit for bit when the whole matrix is one block, and the
// recursive-doubling inverse and an MFMA product otherwise (inverse.hip,
// gemm.hip); every other product runs on the FP64 MFMA GEMM.
#include <algorithm>

#include "kernels.hpp"
#include "lapack.hpp"

namespace elemamd {

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// scratch: the diagonal block's inverse X11 (128 x 128, ld 128), then W
// (nb x n) and P (n x nb)
constexpr i64 kLdX = 128;
size_t inverse_scratch_bytes(int n, int nb) {
    const size_t b = (size_t)std::max(1, std::min(nb, 128));
    return al256(kLdX * kLdX * sizeof(double)) + 2 * al256(b * (size_t)std::max(1, n) * sizeof(double));
}

// X11 = inv(L11) of a b x b diagonal block: the reference's unblocked loop
// bit for bit when the whole matrix is this one block (n <= nb), otherwise
// the recursive-doubling kernel (the blocked steps only need an accurate
// inverse; the unblocked loop is a 128-step chain on one workgroup).
static hipError_t block_inverse(bool single, int b, const double* L11, i64 ld, bool unit, double* Xs, double* Lout,
                                const int* info, hipStream_t s) {
    if (single) return launch_trtri_unb_f64(b, L11, ld, unit, Xs, kLdX, Lout, ld, info, s);
    return launch_trtri_fast_f64(b, L11, ld, unit, Xs, kLdX, Lout, ld, info, s);
}

// hpd_inverse::CholeskyLVar2 (include/elemental/lapack-like/Inverse/
// CholeskyLVar2.hpp:29-75) on the lower triangle: per diagonal block (A00
// the inverted part above-left of it, A10 / A21 its row / column panels,
// A20 and A22 below):
//     Cholesky(A11);  A10 := inv(L11) A10;  A21 := A21 inv(L11)^T;
//     A00 += A10^T A10;  A20 -= A21 A10;  A22 -= A21 A21^T;
//     A10 := inv(L11)^T A10;  A21 := -A21 inv(L11);
//     A11 := inv(L11)^T inv(L11)        (TriangularInverse, then Trtrmm)
// The four triangular solves with L11 become products with X11 = inv(L11)
// (computed once per block), so the
// whole step is one potrf, one small inverse and GEMMs:
//     W = X11 A10,  P = A21 X11^T,  A00 += W^T W,  A20 -= P W,
//     A22 -= P P^T,  A10 := X11^T W,  A21 := -P X11,  A11 := X11^T X11.
// A failed pivot stops everything after it (every launch checks *dinfo), so
// A is left as the reference leaves it when Cholesky(A11) throws.
int hpd_inverse_lower(int n, double* A, i64 lda, int nb, int* dinfo, void* scratch, hipStream_t s) {
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr) return kErrArg;
    ELEMAMD_HIP_TRY(launch_fill_int(dinfo, 0, 1, s));
    if (n == 0) return kOk;
    if (scratch == nullptr) return kErrWorkspace;
    double* Xs = (double*)scratch;
    double* Ws = (double*)((char*)scratch + al256(kLdX * kLdX * sizeof(double)));
    double* Ps = (double*)((char*)Ws + al256((size_t)nb * n * sizeof(double)));
    const i64 ldw = nb, ldp = n;
    constexpr int F = kGemmShapeFull, Lo = kGemmShapeLower;
    for (int k0 = 0; k0 < n; k0 += nb) {
        const int b = std::min(nb, n - k0), j = k0, m = n - k0 - b;
        double* A11 = A + k0 + (i64)k0 * lda;
        double* A10 = A + k0;
        double* A21 = A11 + b;
        double* A20 = A + k0 + b;
        double* A22 = A11 + b + (i64)b * lda;
        ELEMAMD_HIP_TRY(launch_potrf_unb<double>(b, A11, lda, dinfo, k0, s));
        ELEMAMD_HIP_TRY(block_inverse(n <= nb, b, A11, lda, false, Xs, nullptr, dinfo, s));
        if (j > 0)  // W = X11 A10
            ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, b, j, b, 1.0, Xs, kLdX, Lo, A10, lda, F, 0.0, Ws, ldw, 0,
                                            dinfo, s));
        if (m > 0)  // P = A21 X11^T
            ELEMAMD_HIP_TRY(launch_gemm_f64(false, true, m, b, b, 1.0, A21, lda, F, Xs, kLdX, Lo, 0.0, Ps, ldp, 0,
                                            dinfo, s));
        if (j > 0)  // A00 += W^T W (lower)
            ELEMAMD_HIP_TRY(
                launch_gemm_f64(true, false, j, j, b, 1.0, Ws, ldw, F, Ws, ldw, F, 1.0, A, lda, 1, dinfo, s));
        if (j > 0 && m > 0)  // A20 -= P W
            ELEMAMD_HIP_TRY(
                launch_gemm_f64(false, false, m, j, b, -1.0, Ps, ldp, F, Ws, ldw, F, 1.0, A20, lda, 0, dinfo, s));
        if (m > 0)  // A22 -= P P^T (lower)
            ELEMAMD_HIP_TRY(
                launch_gemm_f64(false, true, m, m, b, -1.0, Ps, ldp, F, Ps, ldp, F, 1.0, A22, lda, 1, dinfo, s));
        if (j > 0)  // A10 := X11^T W
            ELEMAMD_HIP_TRY(launch_gemm_f64(true, false, b, j, b, 1.0, Xs, kLdX, Lo, Ws, ldw, F, 0.0, A10, lda, 0,
                                            dinfo, s));
        if (m > 0)  // A21 := -P X11
            ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, m, b, b, -1.0, Ps, ldp, F, Xs, kLdX, Lo, 0.0, A21, lda, 0,
                                            dinfo, s));
        // A11 := X11^T X11 (lower): the reference's TrtrmmLUnblocked bit for
        // bit for a single block, else the MFMA product
        if (n <= nb)
            ELEMAMD_HIP_TRY(launch_trtrmm_unb_f64(b, Xs, kLdX, A11, lda, dinfo, s));
        else
            ELEMAMD_HIP_TRY(
                launch_gemm_f64(true, false, b, b, b, 1.0, Xs, kLdX, Lo, Xs, kLdX, Lo, 0.0, A11, lda, 1, dinfo, s));
    }
    return kOk;
}

// triangular_inverse::LVar3 (include/elemental/lapack-like/
// TriangularInverse/LVar3.hpp:52-95) on the lower triangle: per diagonal
// block
//     L10 := -inv(L11) L10;  L20 += L21 L10;  L21 := L21 inv(L11);  L11 := inv(L11)
// with X11 = inv(L11) (written into L11 at once: later steps of this block
// use the copy in scratch) and the two solves as products with it.  unit: the diagonal is taken as one and
// left untouched.
int triangular_inverse_lower(int n, double* L, i64 ldl, bool unit, int nb, void* scratch, hipStream_t s) {
    if (n < 0 || ldl < std::max(1, n) || nb < 1 || nb > 128) return kErrArg;
    if (n == 0) return kOk;
    if (scratch == nullptr) return kErrWorkspace;
    double* Xs = (double*)scratch;
    double* Ws = (double*)((char*)scratch + al256(kLdX * kLdX * sizeof(double)));
    double* Ps = (double*)((char*)Ws + al256((size_t)nb * n * sizeof(double)));
    const i64 ldw = nb, ldp = n;
    const int F = kGemmShapeFull, T = unit ? kGemmShapeLowerUnit : kGemmShapeLower;
    for (int k0 = 0; k0 < n; k0 += nb) {
        const int b = std::min(nb, n - k0), j = k0, m = n - k0 - b;
        double* L11 = L + k0 + (i64)k0 * ldl;
        double* L10 = L + k0;
        double* L21 = L11 + b;
        double* L20 = L + k0 + b;
        ELEMAMD_HIP_TRY(block_inverse(n <= nb, b, L11, ldl, unit, Xs, L11, nullptr, s));
        if (j > 0) {  // W = -X11 L10;  L20 += L21 W;  L10 := W
            ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, b, j, b, -1.0, Xs, kLdX, T, L10, ldl, F, 0.0, Ws, ldw, 0,
                                            nullptr, s));
            if (m > 0)
                ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, m, j, b, 1.0, L21, ldl, F, Ws, ldw, F, 1.0, L20, ldl, 0,
                                                nullptr, s));
            ELEMAMD_HIP_TRY(launch_copy_f64(b, j, Ws, ldw, L10, ldl, s));
        }
        if (m > 0) {  // L21 := L21 X11
            ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, m, b, b, 1.0, L21, ldl, F, Xs, kLdX, T, 0.0, Ps, ldp, 0,
                                            nullptr, s));
            ELEMAMD_HIP_TRY(launch_copy_f64(m, b, Ps, ldp, L21, ldl, s));
        }
    }
    return kOk;
}

}  // namespace elemamd
