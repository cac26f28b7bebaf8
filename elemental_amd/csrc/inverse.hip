// Unblocked kernels of the HPD inverse and of the triangular inverse on an
// n x n (n <= 128) diagonal block, by one workgroup, bit-identical to the
// reference loops in IEEE double (contraction into FMA is off):
//
//  * triangular_inverse::LVar3Unb (include/elemental/lapack-like/
//    TriangularInverse/LVar3.hpp:19-49), lower, unit or non-unit diagonal.
//    Step j of the loop divides row j left of the diagonal by -lambda_j,
//    applies the rank-1 update L[j+1:, :j] += L[j+1:, j] L[j, :j] (Geru with
//    alpha = 1), divides column j below the diagonal by lambda_j and inverts
//    the pivot.  Followed element by element, X = inv(L) is
//        X[k][k] = 1 / L[k][k],
//        X[i][k] = (L[i][k]/L[k][k] + sum_{j=k+1}^{i-1} L[i][j] X[j][k]) / (-L[i][i])
//    with the sum taken in increasing j, each product and sum rounded: the
//    original L[i][j] enters the update of step j (its column is divided
//    only after the update) and X[j][k] is row j after its own division.
//    Columns of X are independent, so each thread owns one column k.
//  * internal::TrtrmmLUnblocked(TRANSPOSE) (include/elemental/blas-like/
//    level3/Trtrmm/Unblocked.hpp:15-57): step j adds l10^T l10 (row j left of
//    the diagonal) to the leading j x j triangle, then scales row j by
//    lambda_j and squares the pivot.  Element by element, C = X^T X is
//        C[i][k] = X[i][k] X[i][i] (X[i][i]^2 on the diagonal)
//                  + sum_{j=i+1}^{n-1} X[j][i] X[j][k]
//    in increasing j: every element is an independent sequence.
//
// Both keep the block packed in LDS (lower triangle, column-major, 66 KB per
// block of 128).
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

constexpr int kInvN = 128;
constexpr int kInvPacked = kInvN * (kInvN + 1) / 2;

__device__ __forceinline__ int pk(int i, int j) { return j * kInvN - (j * (j - 1)) / 2 + (i - j); }

__global__ __launch_bounds__(128) void trtri_unb_kernel(int n, const double* __restrict__ L, i64 ldl, int unit,
                                                        double* __restrict__ X, i64 ldx, double* Lout, i64 ldo,
                                                        const int* __restrict__ info) {
#pragma clang fp contract(off)
    __shared__ double sL[kInvPacked];
    __shared__ double sX[kInvPacked];
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x;
    for (int j = 0; j < n; ++j) {
        const int i = j + t;
        if (i < n) sL[pk(i, j)] = L[i + (i64)j * ldl];
    }
    __syncthreads();
    const int k = t;  // this thread's column of X
    const bool own = k < n;
    const double lam_k = own ? sL[pk(k, k)] : 1.0;
    if (own) sX[pk(k, k)] = unit ? 1.0 : 1.0 / lam_k;
    // uniform loops (every lane reads the same L[i][j], a broadcast); lane k
    // takes the terms j > k of rows i > k
    for (int i = 1; i < n; ++i) {
        const bool act = own && k < i;
        const double lik = act ? sL[pk(i, k)] : 0.0;
        double v = unit ? lik : lik / lam_k;
        for (int j = 1; j < i; ++j) {
            const double lij = sL[pk(i, j)];
            const double xjk = (j > k && act) ? sX[pk(j, k)] : 0.0;
            if (j > k) v = v + lij * xjk;
        }
        if (act) sX[pk(i, k)] = unit ? -v : v / (-sL[pk(i, i)]);
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const int i = t;
        if (i < n) {
            const double v = i >= j ? sX[pk(i, j)] : 0.0;
            X[i + (i64)j * ldx] = v;
            if (Lout != nullptr && (i > j || (i == j && !unit))) Lout[i + (i64)j * ldo] = v;
        }
    }
}

__global__ __launch_bounds__(256) void trtrmm_unb_kernel(int n, const double* __restrict__ Xin, i64 ldx,
                                                         double* __restrict__ C, i64 ldc,
                                                         const int* __restrict__ info) {
#pragma clang fp contract(off)
    __shared__ double sX[kInvPacked];
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x;
    for (int j = 0; j < n; j += 2) {
        const int jj = j + (t >> 7), i = jj + (t & 127);
        if (jj < n && i < n) sX[pk(i, jj)] = Xin[i + (i64)jj * ldx];
    }
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += 2) {
        const int k = k0 + (t >> 7), i = k + (t & 127);
        if (k >= n || i >= n) continue;
        const double xii = sX[pk(i, i)];
        double v = (i == k) ? xii * xii : sX[pk(i, k)] * xii;
        for (int j = i + 1; j < n; ++j) v = v + sX[pk(j, i)] * sX[pk(j, k)];
        C[i + (i64)k * ldc] = v;
    }
}

hipError_t launch_trtri_unb_f64(int n, const double* L, i64 ldl, bool unit, double* X, i64 ldx, double* Lout,
                                i64 ldo, const int* info, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > kInvN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(trtri_unb_kernel, dim3(1), dim3(128), 0, s, n, L, ldl, unit ? 1 : 0, X, ldx, Lout, ldo, info);
    return hipGetLastError();
}

hipError_t launch_trtrmm_unb_f64(int n, const double* X, i64 ldx, double* C, i64 ldc, const int* info,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > kInvN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(trtrmm_unb_kernel, dim3(1), dim3(256), 0, s, n, X, ldx, C, ldc, info);
    return hipGetLastError();
}

}  // namespace elemamd
