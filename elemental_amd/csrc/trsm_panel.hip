// Panel triangular solve of the blocked Cholesky:
//   B := B * inv(L)^T,  L nb x nb lower (nb <= 128), B m x nb.
// This is Trsm(RIGHT, LOWER, ADJOINT, NON_UNIT, 1, A11, A21)
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:113) which the reference
// hands to dtrsm_('R','L','T','N') (src/core/imports/blas.cpp:1203-1210).
//
// Rows of B are independent.  Each group of four lanes owns one row; lane q
// of the group keeps columns q, q+4, q+8, ... in registers (32 doubles).
// L and the reciprocals of its diagonal live in LDS.  Columns are finalised
// in increasing k exactly as in the column-oriented reference BLAS loop
// (x_k := (1/l_kk) * b_k, then b_j := b_j - l_jk * x_k for j > k), with
// separately rounded products, so the panel is bit-identical to that loop.
// The finished x_k is broadcast inside the lane group with DPP, so the whole
// solve runs without a single barrier after the initial load.
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

constexpr int kTrsmMaxN = 128;
constexpr int kTrsmRowsPerBlock = 64;

// One column k = 4*K4 + R of the right-looking sweep.  All register indices
// are compile-time constants (the sweep is instantiated per column), so the
// row slice stays in VGPRs.
template <int K4, int R, bool FULL>
__device__ __forceinline__ void rlt_column(double (&b)[32], int q, int n, const double* sL, const double* sRinv) {
#pragma clang fp contract(off)
    constexpr int k = 4 * K4 + R;
    // A uniform branch per column (also when n == 128 is known to the caller)
    // ends the basic block, so each column is scheduled on its own and the
    // sweep's LDS reads are not all hoisted into registers.
    if (k >= n) return;
    const double xv = sRinv[k] * b[K4];
    const double x = quad_bcast<R>(xv);
    if (q == R) b[K4] = x;
#pragma unroll
    for (int s = K4; s < 32; ++s) {
        const int j = 4 * s + q;
        if (s > K4 || q > R) {
            if (FULL || j < n) b[s] = b[s] - sL[j + k * kTrsmMaxN] * x;
        }
    }
}

template <int K4, bool FULL>
struct RltSweep {
    static __device__ __forceinline__ void run(double (&b)[32], int q, int n, const double* sL, const double* sRinv) {
        rlt_column<K4, 0, FULL>(b, q, n, sL, sRinv);
        rlt_column<K4, 1, FULL>(b, q, n, sL, sRinv);
        rlt_column<K4, 2, FULL>(b, q, n, sL, sRinv);
        rlt_column<K4, 3, FULL>(b, q, n, sL, sRinv);
        RltSweep<K4 + 1, FULL>::run(b, q, n, sL, sRinv);
    }
};
template <bool FULL>
struct RltSweep<32, FULL> {
    static __device__ __forceinline__ void run(double (&)[32], int, int, const double*, const double*) {}
};

template <typename T, bool FULL>
__global__ __launch_bounds__(256) void trsm_rlt_kernel(int m, int nb, const T* __restrict__ L, i64 ldl,
                                                       T* __restrict__ B, i64 ldb, const int* __restrict__ info,
                                                       int limit) {
#pragma clang fp contract(off)
    __shared__ double sL[kTrsmMaxN * kTrsmMaxN];
    __shared__ double sRinv[kTrsmMaxN];
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x;
    const int n = nb;
    for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        sL[i + j * kTrsmMaxN] = (i >= j) ? (double)L[i + (i64)j * ldl] : 0.0;
    }
    for (int k = t; k < n; k += 256) sRinv[k] = 1.0 / (double)L[k + (i64)k * ldl];
    __syncthreads();

    const int q = t & 3;
    const i64 row = (i64)blockIdx.x * kTrsmRowsPerBlock + (t >> 2);
    const bool active = row < m;
    double b[32];
#pragma unroll
    for (int s = 0; s < 32; ++s) {
        const int j = 4 * s + q;
        b[s] = (active && (FULL || j < n)) ? (double)B[row + (i64)j * ldb] : 0.0;
    }
    RltSweep<0, FULL>::run(b, q, n, sL, sRinv);
    if (active) {
#pragma unroll
        for (int s = 0; s < 32; ++s) {
            const int j = 4 * s + q;
            if (FULL || j < n) B[row + (i64)j * ldb] = (T)b[s];
        }
    }
}

template <typename T>
hipError_t launch_trsm_rlt(int m, int nb, const T* L, i64 ldl, T* B, i64 ldb, const int* info, int limit,
                           hipStream_t s) {
    if (m <= 0 || nb <= 0) return hipSuccess;
    dim3 grid((unsigned)ceil_div(m, kTrsmRowsPerBlock));
    if (nb == kTrsmMaxN)
        hipLaunchKernelGGL((trsm_rlt_kernel<T, true>), grid, dim3(256), 0, s, m, nb, L, ldl, B, ldb, info, limit);
    else
        hipLaunchKernelGGL((trsm_rlt_kernel<T, false>), grid, dim3(256), 0, s, m, nb, L, ldl, B, ldb, info, limit);
    return hipGetLastError();
}

template hipError_t launch_trsm_rlt<double>(int, int, const double*, i64, double*, i64, const int*, int, hipStream_t);
template hipError_t launch_trsm_rlt<float>(int, int, const float*, i64, float*, i64, const int*, int, hipStream_t);

}  // namespace elemamd
