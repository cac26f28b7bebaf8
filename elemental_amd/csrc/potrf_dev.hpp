// Device code of the diagonal-block Cholesky (LVar3Unb), shared by the
// single-matrix and batched launches.
#pragma once

#include "common.hpp"

namespace elemamd {

constexpr int kPotrfMaxN = 128;
constexpr int kPotrfThreads = 512;

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Columns j = 16*A + r, r = 0..15, with A fixed at compile time so that
// the owner's register indices (a = A/2, b = A) are static.
template <int A>
__device__ __forceinline__ bool potrf_col_group(double (&e)[4][8], int n, int tr, int tc, int wave, int lane,
                                                double (*col)[kPotrfMaxN], int* sfail) {
#pragma clang fp contract(off)
    constexpr int a_own = A / 2;  // row block of the pivot (j / 32)
    constexpr int b_own = A;      // column block of column j (j / 16)
    for (int r = 0; r < 16; ++r) {
        const int j = 16 * A + r;
        if (j >= n) return true;
        double* cj = col[j & 1];
        // owners of column j: tc == r, threads 32r .. 32r+31, wave r/2
        if (wave == (r >> 1)) {
            const int own_lane = ((r & 1) << 5) + (j & 31);  // thread tr = j%32, tc = r
            const double ajj = readlane_d(e[a_own][b_own], own_lane);
            if (tc == r) {
                if (ajj <= 0.0) {
                    if (tr == 0) *sfail = j + 1;
                } else {
                    const double alpha = sqrt(ajj);
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int i = tr + 32 * a;
                        if (i > j && i < n) {
                            e[a][b_own] = e[a][b_own] / alpha;
                            cj[i] = e[a][b_own];
                        } else if (i == j) {
                            e[a][b_own] = alpha;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (*sfail != 0) return false;
        // rank-1 update of the elements with j < k <= i < n.  All column-j
        // values this thread needs are read from LDS first (one wait), then
        // the masked products are applied from registers.
        double ri[4], ck[8];
#pragma unroll
        for (int a = 0; a < 4; ++a) ri[a] = cj[tr + 32 * a];
#pragma unroll
        for (int b = A; b < 8; ++b) ck[b] = cj[tc + 16 * b];
#pragma unroll
        for (int b = A; b < 8; ++b) {
            const int k = tc + 16 * b;
            if (16 * b + 15 <= j) continue;  // whole column block finished (uniform)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (32 * a + 31 <= j) continue;  // whole row block finished (uniform)
                const int i = tr + 32 * a;
                if (k > j && i >= k && i < n) e[a][b] = e[a][b] - ri[a] * ck[b];
            }
        }
    }
    return true;
}

// One n x n (n <= 128) diagonal block by the calling 512-thread workgroup.
template <typename T>
__device__ __forceinline__ void potrf_unb_block(int n, T* __restrict__ A, i64 lda, int* __restrict__ info,
                                                int col_offset) {
#pragma clang fp contract(off)
    __shared__ double col[2][kPotrfMaxN];
    __shared__ int sfail;
    if (info != nullptr && *info != 0 && *info <= col_offset) return;  // an earlier block failed
    const int t = threadIdx.x;
    const int tr = t & 31, tc = t >> 5;
    const int wave = t >> 6, lane = t & 63;
    if (t == 0) sfail = 0;
    double e[4][8];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int i = tr + 32 * a, k = tc + 16 * b;
            e[a][b] = (i < n && k <= i) ? (double)A[i + (i64)k * lda] : 0.0;
        }
    __syncthreads();
    bool ok = potrf_col_group<0>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<1>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<2>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<3>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<4>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<5>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<6>(e, n, tr, tc, wave, lane, col, &sfail);
    if (ok) ok = potrf_col_group<7>(e, n, tr, tc, wave, lane, col, &sfail);
    // Write back the lower triangle.  On failure at column j this is the
    // state the reference leaves behind when LVar3Unb throws: columns < j
    // factored and the trailing triangle updated by them.
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int i = tr + 32 * a, k = tc + 16 * b;
            if (i < n && k <= i) A[i + (i64)k * lda] = (T)e[a][b];
        }
    if (!ok && t == 0 && info != nullptr) *info = col_offset + sfail;
}


}  // namespace elemamd
