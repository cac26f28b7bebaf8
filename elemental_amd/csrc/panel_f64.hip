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
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, c = lane & 15, g = lane >> 4;
    const int np = (b + 31) / 32;
    const i64 rw = (i64)blockIdx.x * 64 + 16 * w;  // this wave's first row
    // the wave's rows (zero past m and past b), loads in flight while L11 is staged
    d4_t a[4][2];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const i64 r = rw + g + 4 * v;
                const int col = 32 * p + 16 * h + c;
                const bool ok = r < m && col < b;
                const double x = B[(ok ? r : 0) + (i64)(ok ? col : 0) * ldb];
                a[p][h][v] = ok ? x : 0.0;
            }
    // the diagonal pieces of L11: rows past b are zero, the padded diagonal
    // is the identity
    for (int e = t; e < 4 * 1024; e += 256) {
        const int q = e >> 10, cc = e & 31, k = (e >> 5) & 31;
        if (cc <= k) continue;
        const int r = 32 * q + cc, col = 32 * q + k;
        const bool ok = r < b && col < b;
        const double x = L[(ok ? r : 0) + (i64)(ok ? col : 0) * ldl];
        if (q & 1) sDg[q >> 1][31 - k][31 - cc] = ok ? x : 0.0;
        else sDg[q >> 1][k][cc] = ok ? x : 0.0;
    }
    if (t < 128) sRinv[t] = t < b ? 1.0 / L[t + (i64)t * ldl] : 1.0;
    __syncthreads();
    double(*sx)[34] = sXq[w];  // X_q, row-major
    const int i16 = lane & 15;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q >= np) break;
        // operands of the updates this piece applies, L_pq[16h + i16][4 ks + g]
        // (rows past b are zero; the columns, below 32 p <= b, all exist)
        double y[3][2][8];
#pragma unroll
        for (int p = q + 1; p < 4; ++p) {
            if (p >= np) break;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int r = 32 * p + 16 * h + i16;
#pragma unroll
                for (int ks = 0; ks < 8; ++ks) {
                    const double x = L[(r < b ? r : 0) + (i64)(32 * q + 4 * ks + g) * ldl];
                    y[p - q - 1][h][ks] = r < b ? x : 0.0;
                }
            }
        }
        if (q & 1) Substitute<0, PackedDiag<true>>::run(a[q], c, PackedDiag<true>{sDg[q >> 1]}, sRinv + 32 * q);
        else Substitute<0, PackedDiag<false>>::run(a[q], c, PackedDiag<false>{sDg[q >> 1]}, sRinv + 32 * q);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                sx[g + 4 * v][16 * h + c] = a[q][h][v];
                const i64 r = rw + g + 4 * v;
                const int col = 32 * q + 16 * h + c;
                if (r < m && col < b) B[r + (i64)col * ldb] = a[q][h][v];
            }
        if (q + 1 >= np) break;
        // the buffer holds X_q for every lane of the wave (the wave's LDS
        // accesses are performed in order; the barrier keeps the compiler's)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int p = q + 1; p < 4; ++p)
            if (p < np) sub_product_regs(a[p], sx, 0, y[p - q - 1][0], y[p - q - 1][1], lane);
        mfma_retire();
        __builtin_amdgcn_wave_barrier();
    }
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
