// Device code of the panel triangular solve B := B inv(L)^T, shared by the
// single-matrix and batched launches.
#pragma once

#include "common.hpp"

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

// Rows [64*row_block, 64*row_block + 64) of the panel by the calling
// 256-thread workgroup.
template <typename T, bool FULL>
__device__ __forceinline__ void trsm_rlt_rows(int m, int nb, const T* __restrict__ L, i64 ldl, T* __restrict__ B,
                                              i64 ldb, const int* __restrict__ info, int limit, int row_block) {
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
    const i64 row = (i64)row_block * kTrsmRowsPerBlock + (t >> 2);
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



// ------------------------------------------------------------------------
// Blocked variant used inside the factorization: the nb columns are taken in
// sub-blocks of 32.  For sub-block c the contributions of the solved
// sub-blocks are applied as one MFMA product,
//     B_c := B_c - X_{0..c-1} L(c, 0..c-1)^T        (FP64 MFMA 16x16x4),
// then X_c := B_c L_cc^{-T} by right-looking substitution, one lane per row
// with the 32 values of the row in registers.  The whole 64 x nb slab of B
// stays in LDS; L is staged one 32-row block row at a time.  Same operations
// as the column sweep, different summation order (the blocked order of a
// BLAS dtrsm); arithmetic is double for double and float storage.
constexpr int kPanelSub = 32;
constexpr int kPanelLd = kTrsmMaxN + 1;

template <typename T>
__device__ __forceinline__ void trsm_rlt_blocked_rows(int m, int nb, const T* __restrict__ L, i64 ldl,
                                                      T* __restrict__ B, i64 ldb, const int* __restrict__ info,
                                                      int limit, int row_block) {
    __shared__ double sX[kTrsmRowsPerBlock][kPanelLd];  // B, then X, of the 64-row slab
    __shared__ double sLr[kPanelSub][kPanelLd];         // block row c of L (columns 0 .. 32c+31)
    __shared__ double sRinv[kPanelSub];
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const i64 r0 = (i64)row_block * kTrsmRowsPerBlock;
    const int nc = (nb + kPanelSub - 1) / kPanelSub;
    const int ncol = nc * kPanelSub;  // padded width
    // slab of B (rows past m and padded columns are zero)
    for (int idx = t; idx < kTrsmRowsPerBlock * ncol; idx += 256) {
        const int r = idx % kTrsmRowsPerBlock, j = idx / kTrsmRowsPerBlock;
        sX[r][j] = (r0 + r < m && j < nb) ? (double)B[r0 + r + (i64)j * ldb] : 0.0;
    }
    for (int c = 0; c < nc; ++c) {
        const int c0 = c * kPanelSub;
        const int kw = c0 + kPanelSub;  // columns of this block row of L that are used
        __syncthreads();                 // previous substitution done with sLr / sX
        for (int idx = t; idx < kPanelSub * kw; idx += 256) {
            const int i = idx % kPanelSub, j = idx / kPanelSub;
            const int gi = c0 + i;
            double v;
            if (gi < nb && j < nb) v = (j <= gi) ? (double)L[gi + (i64)j * ldl] : 0.0;
            else v = (gi == j) ? 1.0 : 0.0;  // identity on the padding
            sLr[i][j] = v;
        }
        if (t < kPanelSub) {
            const int gi = c0 + t;
            sRinv[t] = (gi < nb) ? 1.0 / (double)L[gi + (i64)gi * ldl] : 1.0;
        }
        __syncthreads();
        if (c > 0) {
            // wave w: rows 16w..16w+15, columns c0..c0+31 as two 16x16 tiles
            d4_t acc0, acc1;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * w + (lane >> 4) + 4 * v;
                acc0[v] = sX[r][c0 + (lane & 15)];
                acc1[v] = sX[r][c0 + 16 + (lane & 15)];
            }
            const int ksteps = c0 / 4;
            for (int ks = 0; ks < ksteps; ++ks) {
                const int kk = 4 * ks + (lane >> 4);
                const double a = -sX[16 * w + (lane & 15)][kk];
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sLr[lane & 15][kk], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sLr[16 + (lane & 15)][kk], acc1, 0, 0, 0);
            }
            // dependent zero-operand MFMAs before the accumulators are read (the
            // compiler's wait states after a loop back edge are short for FP64 MFMA)
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc1, 0, 0, 0);
            __syncthreads();  // every wave has read its operands of sX
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * w + (lane >> 4) + 4 * v;
                sX[r][c0 + (lane & 15)] = acc0[v];
                sX[r][c0 + 16 + (lane & 15)] = acc1[v];
            }
            __syncthreads();
        }
        if (w == 0) {
            double b[kPanelSub];
#pragma unroll
            for (int j = 0; j < kPanelSub; ++j) b[j] = sX[lane][c0 + j];
#pragma unroll
            for (int k = 0; k < kPanelSub; ++k) {
                b[k] = b[k] * sRinv[k];
#pragma unroll
                for (int j = k + 1; j < kPanelSub; ++j) b[j] = b[j] - sLr[j][c0 + k] * b[k];
            }
#pragma unroll
            for (int j = 0; j < kPanelSub; ++j) sX[lane][c0 + j] = b[j];
        }
    }
    __syncthreads();
    for (int idx = t; idx < kTrsmRowsPerBlock * nb; idx += 256) {
        const int r = idx % kTrsmRowsPerBlock, j = idx / kTrsmRowsPerBlock;
        if (r0 + r < m) B[r0 + r + (i64)j * ldb] = (T)sX[r][j];
    }
}

}  // namespace elemamd
