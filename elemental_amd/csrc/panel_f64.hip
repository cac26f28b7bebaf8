// Panel solve of the FP64 blocked factorization, B := B L11^{-T}: the
// Trsm(RIGHT, LOWER, ADJOINT, NON_UNIT, 1, A11, A21) of cholesky::LVar3
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:113), which the
// reference hands to dtrsm_('R','L','T','N') (src/core/imports/blas.cpp:1203-1210).
//
// One 64-row slab per workgroup, 16 rows per wave; a wave never waits for
// another after L11 is staged.  The rows stay in registers in the FP64
// MFMA accumulator layout (lane (c, g) = (lane % 16, lane / 16) holds rows
// g + 4v, v = 0..3, of column c of a 16-column tile), 128 columns as four
// 32-column pieces of two tiles each.  Piece by piece, right-looking:
//   * substitution inside the piece: for each column k, x = b_k / l_kk (by
//     the reciprocal, as the reference BLAS loop), broadcast along the
//     16-lane row with DPP row_newbcast (the column's owner lane is lane k
//     of every row), then b_j -= l_jk x for the later columns j;
//   * the solved piece X_q goes through a 4 KB wave-private LDS buffer into
//     MFMA operand order and updates every later piece p at once,
//     B_p -= X_q L_pq^T (FP64 MFMA 16x16x4, the accumulators are the rows);
//     the L_pq operands come straight from memory (L2), loaded before the
//     piece's substitution so their latency hides under it.
// Same operations as the column sweep in the blocked order of a BLAS dtrsm
// (exact on integer problems); FMA contractions, so not bit-identical to the
// reference loop (the bit-exact column sweep is trsm_rlt_kernel).
//
// LDS: the four diagonal pieces' strict lower triangles packed two to an
// array (16 KB), the reciprocal diagonal and the four waves' buffers: 34 KB,
// so a workgroup fits on a CU beside two workgroups of the FP64 trailing
// update (whose launches on the look-ahead stream reserve 56 KB each).
#include "chain64_dev.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

using namespace chain64;

__global__ __launch_bounds__(256) void panel_f64_kernel(int m, int b, const double* __restrict__ L, i64 ldl,
                                                        double* __restrict__ B, i64 ldb, const int* __restrict__ info,
                                                        int limit) {
    __shared__ double sDg[2][32][32];  // diagonal pieces, strict lower triangles (PackedDiag)
    __shared__ double sRinv[128];
    __shared__ double sXq[4][16][34];  // per wave: the piece just solved, row-major
    // the factorization's chain: its waves issue ahead of the trailing
    // update's waves sharing the SIMD (wave priority, not queue priority)
    __builtin_amdgcn_s_setprio(3);
    if (info != nullptr && *info != 0 && *info <= limit) return;
    panel64_slab<double>(m, b, L, ldl, B, ldb, blockIdx.x, sDg, sRinv, sXq);
}

hipError_t launch_panel_f64(int m, int b, const double* L, i64 ldl, double* B, i64 ldb, const int* info, int limit,
                            hipStream_t s) {
    if (m <= 0 || b <= 0) return hipSuccess;
    if (b > 128) return hipErrorInvalidValue;
    hipLaunchKernelGGL(panel_f64_kernel, dim3((unsigned)ceil_div(m, 64)), dim3(256), 0, s, m, b, L, ldl, B, ldb, info,
                       limit);
    return hipGetLastError();
}

}  // namespace elemamd
