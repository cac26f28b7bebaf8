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

// Columns j = NTC*A + r, r < NTC, with A fixed at compile time so that the
// owner's register indices (a = j/32, b = A) are static.  NTC = threads/32
// column groups: thread (tr, tc) = (t % 32, t / 32) owns the elements
// (i, k) = (tr + 32a, tc + NTC b), a < 4, b < 128/NTC.
template <int A, int NTC>
__device__ __forceinline__ bool potrf_col_group(double (&e)[4][128 / NTC], int n, int tr, int tc, int wave,
                                                double (*col)[kPotrfMaxN], int* sfail) {
#pragma clang fp contract(off)
    constexpr int NBC = 128 / NTC;
    constexpr int a_own = NTC * A / 32;  // row block of the pivots (j / 32)
    constexpr int b_own = A;             // column block of the columns (j / NTC)
    for (int r = 0; r < NTC; ++r) {
        const int j = NTC * A + r;
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
        double ri[4], ck[NBC];
#pragma unroll
        for (int a = 0; a < 4; ++a) ri[a] = cj[tr + 32 * a];
#pragma unroll
        for (int b = A; b < NBC; ++b) ck[b] = cj[tc + NTC * b];
#pragma unroll
        for (int b = A; b < NBC; ++b) {
            const int k = tc + NTC * b;
            if (NTC * b + NTC - 1 <= j) continue;  // whole column block finished (uniform)
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

template <int A, int NTC>
struct PotrfGroups {
    static __device__ __forceinline__ bool run(double (&e)[4][128 / NTC], int n, int tr, int tc, int wave,
                                               double (*col)[kPotrfMaxN], int* sfail) {
        if (!potrf_col_group<A, NTC>(e, n, tr, tc, wave, col, sfail)) return false;
        return PotrfGroups<A + 1, NTC>::run(e, n, tr, tc, wave, col, sfail);
    }
};
template <int NTC>
struct PotrfGroups<128 / NTC, NTC> {
    static __device__ __forceinline__ bool run(double (&)[4][128 / NTC], int, int, int, int, double (*)[kPotrfMaxN],
                                               int*) {
        return true;
    }
};

// One n x n (n <= 128) diagonal block by the calling workgroup of THREADS
// (512 or 256) threads.  The operation order per element is LVar3Unb's
// whatever the thread count, so the result is the same bit for bit.
template <typename T, int THREADS = kPotrfThreads>
__device__ __forceinline__ void potrf_unb_block(int n, T* __restrict__ A, i64 lda, int* __restrict__ info,
                                                int col_offset) {
#pragma clang fp contract(off)
    constexpr int NTC = THREADS / 32, NBC = 128 / NTC;
    __shared__ double col[2][kPotrfMaxN];
    __shared__ int sfail;
    if (info != nullptr && *info != 0 && *info <= col_offset) return;  // an earlier block failed
    const int t = threadIdx.x;
    const int tr = t & 31, tc = t >> 5;
    const int wave = t >> 6;
    if (t == 0) sfail = 0;
    double e[4][NBC];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NBC; ++b) {
            const int i = tr + 32 * a, k = tc + NTC * b;
            e[a][b] = (i < n && k <= i) ? (double)A[i + (i64)k * lda] : 0.0;
        }
    __syncthreads();
    const bool ok = PotrfGroups<0, NTC>::run(e, n, tr, tc, wave, col, &sfail);
    // Write back the lower triangle.  On failure at column j this is the
    // state the reference leaves behind when LVar3Unb throws: columns < j
    // factored and the trailing triangle updated by them.
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NBC; ++b) {
            const int i = tr + 32 * a, k = tc + NTC * b;
            if (i < n && k <= i) A[i + (i64)k * lda] = (T)e[a][b];
        }
    if (!ok && t == 0 && info != nullptr) *info = col_offset + sfail;
}

}  // namespace elemamd
