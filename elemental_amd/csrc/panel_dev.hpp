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
    // stage L11 (lower, zero above) with all loads of a thread in flight at once:
    // thread t covers rows i = t % 128 of columns j = t / 128 + 2s
    {
        const int i = t & (kTrsmMaxN - 1), j0 = t >> 7;
#pragma unroll
        for (int h = 0; h < 4; ++h) {  // four batches of 16 loads in flight
            T g[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int j = j0 + 2 * (16 * h + s);
                const bool ok = i < n && j < n && i >= j;
                g[s] = L[ok ? i + (i64)j * ldl : 0];
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int j = j0 + 2 * (16 * h + s);
                sL[i + j * kTrsmMaxN] = (i < n && j < n && i >= j) ? (double)g[s] : 0.0;
            }
        }
    }
    if (t < n) sRinv[t] = 1.0 / (double)L[t + (i64)t * ldl];
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
    __shared__ double sLd[kPanelSub][kPanelSub];  // diagonal block, transposed
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const i64 r0 = (i64)row_block * kTrsmRowsPerBlock;
    const int nc = (nb + kPanelSub - 1) / kPanelSub;
    const int ncol = nc * kPanelSub;  // padded width
    // slab of B (rows past m and padded columns are zero)
    {  // thread: row r = t % 64 of columns t / 64 + 4s.  Loads from clamped
       // addresses without branches, all in flight before the first LDS store.
        const int r = t & 63, j0 = t >> 6;
        const bool rin = r0 + r < m;
        T g[kTrsmMaxN / 4];
#pragma unroll
        for (int s = 0; s < kTrsmMaxN / 4; ++s) {
            const int j = j0 + 4 * s;
            const bool ok = rin && j < nb;
            g[s] = B[ok ? r0 + r + (i64)j * ldb : 0];
        }
#pragma unroll
        for (int s = 0; s < kTrsmMaxN / 4; ++s) {
            const int j = j0 + 4 * s;
            if (j < ncol) sX[r][j] = (rin && j < nb) ? (double)g[s] : 0.0;
        }
    }
    for (int c = 0; c < nc; ++c) {
        const int c0 = c * kPanelSub;
        const int kw = c0 + kPanelSub;  // columns of this block row of L that are used
        __syncthreads();                 // previous substitution done with sLr / sX
        {  // thread: row i = t % 32 of columns t / 32 + 8s; loads first, then LDS
            const int i = t & 31, jj = t >> 5;
            const int gi = c0 + i;
            T g[kTrsmMaxN / 8];
#pragma unroll
            for (int s = 0; s < kTrsmMaxN / 8; ++s) {
                const int j = jj + 8 * s;
                const bool ok = gi < nb && j <= gi;
                g[s] = L[ok ? gi + (i64)j * ldl : 0];
            }
#pragma unroll
            for (int s = 0; s < kTrsmMaxN / 8; ++s) {
                const int j = jj + 8 * s;
                double v;
                if (gi < nb && j < nb) v = (j <= gi) ? (double)g[s] : 0.0;
                else v = (gi == j) ? 1.0 : 0.0;  // identity on the padding
                if (j < kw) sLr[i][j] = v;
            }
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
            // diagonal block transposed, sLd[k][j] = L(c0+j, c0+k): the column a
            // step needs is contiguous, read as broadcast 16-byte loads
            if (lane < kPanelSub) {
#pragma unroll
                for (int k = 0; k < kPanelSub; ++k) sLd[k][lane] = sLr[lane][c0 + k];
            }
            double b[kPanelSub];
#pragma unroll
            for (int j = 0; j < kPanelSub; ++j) b[j] = sX[lane][c0 + j];
            // right-looking sweep; column k+1 is fetched while column k is applied
            double lk[kPanelSub], ln[kPanelSub];
#pragma unroll
            for (int j = 0; j < kPanelSub; ++j) lk[j] = sLd[0][j];
#pragma unroll
            for (int k = 0; k < kPanelSub; ++k) {
                if (k + 1 < kPanelSub) {
#pragma unroll
                    for (int j = k + 2; j < kPanelSub; ++j) ln[j] = sLd[k + 1][j];
                }
                b[k] = b[k] * sRinv[k];
#pragma unroll
                for (int j = k + 1; j < kPanelSub; ++j) b[j] = b[j] - lk[j] * b[k];
#pragma unroll
                for (int j = k + 2; j < kPanelSub; ++j) lk[j] = ln[j];
            }
#pragma unroll
            for (int j = 0; j < kPanelSub; ++j) sX[lane][c0 + j] = b[j];
        }
    }
    __syncthreads();
    {
        const int r = t & 63, j0 = t >> 6;
        if (r0 + r < m) {
#pragma unroll
            for (int s = 0; s < kTrsmMaxN / 4; ++s) {
                const int j = j0 + 4 * s;
                if (j < nb) B[r0 + r + (i64)j * ldb] = (T)sX[r][j];
            }
        }
    }
}

}  // namespace elemamd
