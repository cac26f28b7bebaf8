// Device code of the panel triangular solve B := B inv(L)^T, shared by the
// single-matrix and batched launches.
#pragma once

#include "common.hpp"

namespace elemamd {

constexpr int kTrsmMaxN = 128;
constexpr int kTrsmRowsPerBlock = 64;

// One column k = 4*K4 + R of the right-looking sweep (NM = the largest
// block width, the leading dimension of sL).  All register indices are
// compile-time constants (the sweep is instantiated per column), so the
// row slice stays in VGPRs.
template <int K4, int R, bool FULL, int NM>
__device__ __forceinline__ void rlt_column(double (&b)[NM / 4], int q, int n, const double* sL, const double* sRinv) {
#pragma clang fp contract(off)
    constexpr int k = 4 * K4 + R;
    // A uniform branch per column (also when n == NM is known to the caller)
    // ends the basic block, so each column is scheduled on its own and the
    // sweep's LDS reads are not all hoisted into registers.
    if (k >= n) return;
    const double xv = sRinv[k] * b[K4];
    const double x = quad_bcast<R>(xv);
    if (q == R) b[K4] = x;
#pragma unroll
    for (int s = K4; s < NM / 4; ++s) {
        const int j = 4 * s + q;
        if (s > K4 || q > R) {
            if (FULL || j < n) b[s] = b[s] - sL[j + k * NM] * x;
        }
    }
}

template <int K4, bool FULL, int NM>
struct RltSweep {
    static __device__ __forceinline__ void run(double (&b)[NM / 4], int q, int n, const double* sL,
                                               const double* sRinv) {
        rlt_column<K4, 0, FULL, NM>(b, q, n, sL, sRinv);
        rlt_column<K4, 1, FULL, NM>(b, q, n, sL, sRinv);
        rlt_column<K4, 2, FULL, NM>(b, q, n, sL, sRinv);
        rlt_column<K4, 3, FULL, NM>(b, q, n, sL, sRinv);
        RltSweep<K4 + 1, FULL, NM>::run(b, q, n, sL, sRinv);
    }
};
template <bool FULL, int NM>
struct RltSweep<NM / 4, FULL, NM> {
    static __device__ __forceinline__ void run(double (&)[NM / 4], int, int, const double*, const double*) {}
};

// Rows [64*row_block, 64*row_block + 64) of the panel by the calling
// 256-thread workgroup; nb <= NM (NM = 128, or 64 where LDS is short).
template <typename T, bool FULL, int NM = kTrsmMaxN>
__device__ __forceinline__ void trsm_rlt_rows(int m, int nb, const T* __restrict__ L, i64 ldl, T* __restrict__ B,
                                              i64 ldb, const int* __restrict__ info, int limit, int row_block) {
#pragma clang fp contract(off)
    __shared__ double sL[NM * NM];
    __shared__ double sRinv[NM];
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x;
    const int n = nb;
    // stage L11 (lower, zero above) with all loads of a thread in flight at once:
    // thread t covers rows i = t % NM of columns j = t / NM + (256/NM) s
    {
        constexpr int JS = 256 / NM, PER = NM * NM / 256;
        const int i = t & (NM - 1), j0 = t / NM;
#pragma unroll
        for (int h = 0; h < PER / 16; ++h) {  // batches of 16 loads in flight
            T g[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int j = j0 + JS * (16 * h + s);
                const bool ok = i < n && j < n && i >= j;
                g[s] = L[ok ? i + (i64)j * ldl : 0];
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int j = j0 + JS * (16 * h + s);
                sL[i + j * NM] = (i < n && j < n && i >= j) ? (double)g[s] : 0.0;
            }
        }
    }
    if (t < n) sRinv[t] = 1.0 / (double)L[t + (i64)t * ldl];
    __syncthreads();

    const int q = t & 3;
    const i64 row = (i64)row_block * kTrsmRowsPerBlock + (t >> 2);
    const bool active = row < m;
    double b[NM / 4];
#pragma unroll
    for (int s = 0; s < NM / 4; ++s) {
        const int j = 4 * s + q;
        b[s] = (active && (FULL || j < n)) ? (double)B[row + (i64)j * ldb] : 0.0;
    }
    RltSweep<0, FULL, NM>::run(b, q, n, sL, sRinv);
    if (active) {
#pragma unroll
        for (int s = 0; s < NM / 4; ++s) {
            const int j = 4 * s + q;
            if (FULL || j < n) B[row + (i64)j * ldb] = (T)b[s];
        }
    }
}

}  // namespace elemamd
