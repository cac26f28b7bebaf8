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
#include "panel_dev.hpp"

namespace elemamd {

template <typename T, bool FULL>
__global__ __launch_bounds__(256) void trsm_rlt_kernel(int m, int nb, const T* __restrict__ L, i64 ldl,
                                                       T* __restrict__ B, i64 ldb, const int* __restrict__ info,
                                                       int limit) {
    trsm_rlt_rows<T, FULL>(m, nb, L, ldl, B, ldb, info, limit, blockIdx.x);
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
