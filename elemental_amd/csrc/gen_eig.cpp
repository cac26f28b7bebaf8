// Host driver of the generalized-definite Hermitian eigensolver
// (elem::HermitianGenDefiniteEig, include/elemental/lapack-like/
// HermitianGenDefiniteEig.hpp:22-126): Cholesky of B, reduction of A to a
// standard symmetric problem with the factor, the standard eigensolver, and
// the back-transformation of the eigenvectors.
//
//   AXBX  A x = lambda B x:  A := inv(L) A inv(L)^T  (TwoSidedTrsm), X := inv(L)^T Z
//   ABX   A B x = lambda x:  A := L^T A L            (TwoSidedTrmm), X := inv(L)^T Z
//   BAX   B A x = lambda x:  A := L^T A L            (TwoSidedTrmm), X := L Z
//
// The Cholesky factor is this library's blocked LVar3; the two-sided
// reductions are done GPU-first as products with the factor or its inverse
// (the blocked triangular inverse, then two FP64 MFMA GEMMs on the full
// symmetric A) instead of the reference's variant-4 sweep of Trsm / Hemm /
// Her2k steps: same matrix, other rounding.  The standard eigenproblem of the
// reduced matrix (the reference calls LAPACK's dsyevr for a local Matrix,
// src/lapack-like/HermitianEig.cpp:224-240, 778-797, 1349-1365) runs on
// rocSOLVER's divide-and-conquer dsyevdx, with the reference's three
// selections: all eigenvalues, indices il..iu (0-based, inclusive), or values
// in (vl, vu].
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <map>

#include "kernels.hpp"
#include "lapack.hpp"

namespace elemamd {

namespace {

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// one rocBLAS handle per (thread, device), bound to the caller's stream
rocblas_handle solver_handle(hipStream_t s) {
    static thread_local std::map<int, rocblas_handle> handles;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    rocblas_handle& h = handles[dev];
    if (h == nullptr && rocblas_create_handle(&h) != rocblas_status_success) {
        h = nullptr;
        return nullptr;
    }
    if (rocblas_set_stream(h, s) != rocblas_status_success) return nullptr;
    return h;
}

}  // namespace

size_t gen_eig_scratch_bytes(int n) {
    const size_t nn = (size_t)std::max(1, n) * (size_t)std::max(1, n) * sizeof(double);
    return 3 * al256(nn) + al256(inverse_scratch_bytes(n, 128)) + 256;
}

int hermitian_gen_definite_eig_lower(int type, int n, double* A, i64 lda, double* B, i64 ldb, double* w, double* X,
                                     i64 ldx, int range, int il, int iu, double vl, double vu, int* nev, int* info,
                                     void* scratch, hipStream_t s) {
    *nev = 0;
    *info = 0;
    if (n == 0) return kOk;
    char* p = (char*)scratch;
    const size_t nn = al256((size_t)n * n * sizeof(double));
    double* F = (double*)p;            // the full symmetric A, then the eigensolver's input
    double* T = (double*)(p + nn);     // inv(L) (lower, zeros above)
    double* G = (double*)(p + 2 * nn); // products, then the eigenvectors Z
    void* inv_work = p + 3 * nn;
    int* dints = (int*)(p + 3 * nn + al256(inverse_scratch_bytes(n, 128)));  // factor info, nev, eig info
    constexpr int Fu = kGemmShapeFull, Lo = kGemmShapeLower;

    // Cholesky(LOWER, B)
    int rc = cholesky_lower<double>(n, B, ldb, 128, kTrailingF64, dints, s);
    if (rc) return rc;
    int hinfo = 0;
    ELEMAMD_HIP_TRY(hipMemcpyAsync(&hinfo, dints, sizeof(int), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    if (hinfo != 0) {
        *info = hinfo;
        return kOk;
    }
    // the two-sided reduction on the full symmetric A
    ELEMAMD_HIP_TRY(launch_symmetrize_lower_f64(n, A, lda, F, n, s));
    bool have_inv = false;
    auto make_inverse = [&]() -> int {
        ELEMAMD_HIP_TRY(launch_copy_lower_zero_upper_f64(n, B, ldb, T, n, s));
        const int r = triangular_inverse_lower(n, T, n, false, 128, inv_work, s);
        have_inv = r == kOk;
        return r;
    };
    if (type == 1) {  // A := inv(L) A inv(L)^T
        if ((rc = make_inverse())) return rc;
        ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, n, n, n, 1.0, T, n, Lo, F, n, Fu, 0.0, G, n, 0, nullptr, s));
        ELEMAMD_HIP_TRY(launch_gemm_f64(false, true, n, n, n, 1.0, G, n, Fu, T, n, Lo, 0.0, A, lda, 1, nullptr, s));
    } else {  // A := L^T A L
        ELEMAMD_HIP_TRY(launch_gemm_f64(true, false, n, n, n, 1.0, B, ldb, Lo, F, n, Fu, 0.0, G, n, 0, nullptr, s));
        ELEMAMD_HIP_TRY(launch_gemm_f64(false, false, n, n, n, 1.0, G, n, Fu, B, ldb, Lo, 0.0, A, lda, 1, nullptr, s));
    }
    // the standard eigenproblem on a copy (the solver overwrites its input)
    ELEMAMD_HIP_TRY(launch_copy_f64(n, n, A, lda, F, n, s));
    rocblas_handle h = solver_handle(s);
    if (h == nullptr) return kErrHip;
    const rocblas_erange er =
        range == 1 ? rocblas_erange_index : (range == 2 ? rocblas_erange_value : rocblas_erange_all);
    const rocblas_status st =
        rocsolver_dsyevdx(h, X != nullptr ? rocblas_evect_original : rocblas_evect_none, er, rocblas_fill_lower, n, F,
                          n, vl, vu, range == 1 ? il + 1 : 1, range == 1 ? iu + 1 : n, dints + 1, w, G, n, dints + 2);
    if (st != rocblas_status_success) return kErrHip;
    int hv[2] = {0, 0};
    ELEMAMD_HIP_TRY(hipMemcpyAsync(hv, dints + 1, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    *nev = hv[0];
    if (hv[1] != 0) return kErrNotConverged;
    if (X == nullptr || hv[0] == 0) return kOk;
    // back-transformation of the selected eigenvectors
    if (type == 3) {  // X := L Z
        ELEMAMD_HIP_TRY(
            launch_gemm_f64(false, false, n, hv[0], n, 1.0, B, ldb, Lo, G, n, Fu, 0.0, X, ldx, 0, nullptr, s));
    } else {  // X := inv(L)^T Z
        if (!have_inv && (rc = make_inverse())) return rc;
        ELEMAMD_HIP_TRY(
            launch_gemm_f64(true, false, n, hv[0], n, 1.0, T, n, Lo, G, n, Fu, 0.0, X, ldx, 0, nullptr, s));
    }
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    return kOk;
}

}  // namespace elemamd
