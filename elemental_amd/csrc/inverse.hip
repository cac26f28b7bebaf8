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

namespace elemamd {

// inv(L) of an n x n (n <= 128) lower-triangular block by recursive doubling
// (not the reference's operation order; for the blocked algorithms, whose
// diagonal blocks only need an accurate inverse).  For a block split as
// L = [L11 0; L21 L22] with X11 = inv(L11), X22 = inv(L22) known,
//     X21 = -X22 (L21 X11),
// applied for block sizes 1, 2, 4, ..., 64 over the whole 128 x 128 block
// (padded with the identity past n), so the 128-step substitution chain
// becomes 7 levels of short dot products.  X lives in LDS (128 x 129
// doubles).  At each level the b x b blocks L21 are first staged into the
// upper-right b x b block of their 2b x 2b diagonal block (zero in the
// result), T = L21 X11 is formed in registers, written over them, and
// X21 = -X22 T read from there.  Output: X (ldx >= 128), the full 128 x 128
// inverse with zeros above the diagonal (the identity past n); when Lout is
// not null its lower triangle (strictly lower for a unit diagonal) receives
// inv(L) too.  unit: the diagonal of L is taken as one and never read.
constexpr int kTrtriFastPerThread = 16;  // T elements per thread at the 64-level
__global__ __launch_bounds__(256) void trtri_fast_kernel(int n, const double* L, i64 ldl, int unit,
                                                         double* __restrict__ X, i64 ldx, double* Lout, i64 ldo,
                                                         const int* __restrict__ info) {
    __shared__ double sX[kInvN][kInvN + 1];
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x;
    // padded L element (identity past n and on a unit diagonal)
    auto Lat = [&](int i, int j) -> double {
        const bool in = i < n && j < n && !(unit && i == j);
        const double v = L[in ? i + (i64)j * ldl : 0];
        return in ? v : (i == j ? 1.0 : 0.0);
    };
    for (int e = t; e < kInvN; e += 256) sX[e][e] = 1.0 / Lat(e, e);
    for (int e = t; e < kInvN * kInvN; e += 256) {
        const int i = e & 127, j = e >> 7;
        if (i != j) sX[i][j] = 0.0;
    }
    int lb = 0;
    for (int b = 1; b < kInvN; b *= 2, ++lb) {
        const int ne = 64 * b;  // elements of all the level's b x b blocks (128 / 2b pairs)
        const int bb = 2 * lb;  // log2(b * b)
        __syncthreads();
        // stage L21 of every pair (lanes along rows: coalesced column reads)
        for (int e = t; e < ne; e += 256) {
            const int q = e >> bb, r = e & ((1 << bb) - 1), i = r & (b - 1), j = r >> lb;
            const int r0 = 2 * b * q, rb = r0 + b;
            sX[r0 + i][rb + j] = Lat(rb + i, r0 + j);
        }
        __syncthreads();
        // T = L21 X11 (X11 lower: m >= j), kept in registers until every
        // thread is done reading the staged L21
        double tv[kTrtriFastPerThread];
#pragma unroll
        for (int u = 0; u < kTrtriFastPerThread; ++u) {
            const int e = t + 256 * u;
            tv[u] = 0.0;
            if (e < ne) {
                const int q = e >> bb, r = e & ((1 << bb) - 1), i = r & (b - 1), j = r >> lb;
                const int r0 = 2 * b * q, rb = r0 + b;
                double acc = 0.0;
                for (int m = j; m < b; ++m) acc = fma(sX[r0 + i][rb + m], sX[r0 + m][r0 + j], acc);
                tv[u] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kTrtriFastPerThread; ++u) {
            const int e = t + 256 * u;
            if (e < ne) {
                const int q = e >> bb, r = e & ((1 << bb) - 1), i = r & (b - 1), j = r >> lb;
                const int r0 = 2 * b * q, rb = r0 + b;
                sX[r0 + i][rb + j] = tv[u];
            }
        }
        __syncthreads();
        // X21 = -X22 T (X22 lower: m <= i)
        for (int e = t; e < ne; e += 256) {
            const int q = e >> bb, r = e & ((1 << bb) - 1), i = r & (b - 1), j = r >> lb;
            const int r0 = 2 * b * q, rb = r0 + b;
            double acc = 0.0;
            for (int m = 0; m <= i; ++m) acc = fma(sX[rb + i][rb + m], sX[r0 + m][rb + j], acc);
            sX[rb + i][r0 + j] = -acc;
        }
        __syncthreads();
        // the T blocks back to zero (the upper triangle of the result)
        for (int e = t; e < ne; e += 256) {
            const int q = e >> bb, r = e & ((1 << bb) - 1), i = r & (b - 1), j = r >> lb;
            const int r0 = 2 * b * q, rb = r0 + b;
            sX[r0 + i][rb + j] = 0.0;
        }
    }
    __syncthreads();
    for (int e = t; e < kInvN * kInvN; e += 256) {
        const int i = e & 127, j = e >> 7;
        const double v = sX[i][j];
        X[i + (i64)j * ldx] = v;
        if (Lout != nullptr && i < n && (i > j || (i == j && !unit))) Lout[i + (i64)j * ldo] = v;
    }
}

hipError_t launch_trtri_fast_f64(int n, const double* L, i64 ldl, bool unit, double* X, i64 ldx, double* Lout,
                                 i64 ldo, const int* info, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > kInvN || ldx < kInvN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(trtri_fast_kernel, dim3(1), dim3(256), 0, s, n, L, ldl, unit ? 1 : 0, X, ldx, Lout, ldo,
                       info);
    return hipGetLastError();
}

}  // namespace elemamd
