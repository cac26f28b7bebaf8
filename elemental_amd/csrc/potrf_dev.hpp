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

// Rank-1 update of the elements with j < k <= i < n by the finished column
// j published in cj (all of it read first, one wait), skipping the column
// block b == SKIPB of the threads with tc == SKIPTC (elements already
// updated by the column's owners; -1: none).
template <int A, int NTC>
__device__ __forceinline__ void potrf_rank1(double (&e)[4][128 / NTC], int n, int tr, int tc, int j, const double* cj,
                                            int skip_tc) {
#pragma clang fp contract(off)
    constexpr int NBC = 128 / NTC;
    double ri[4], ck[NBC];
#pragma unroll
    for (int a = 0; a < 4; ++a) ri[a] = cj[tr + 32 * a];
#pragma unroll
    for (int b = A; b < NBC; ++b) ck[b] = cj[tc + NTC * b];
#pragma unroll
    for (int b = A; b < NBC; ++b) {
        const int k = tc + NTC * b;
        if (NTC * b + NTC - 1 <= j) continue;  // whole column block finished (uniform)
        if (b == A && tc == skip_tc) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (32 * a + 31 <= j) continue;  // whole row block finished (uniform)
            const int i = tr + 32 * a;
            if (k > j && i >= k && i < n) e[a][b] = e[a][b] - ri[a] * ck[b];
        }
    }
}

// Columns j = NTC*A + r, r < NTC, two per barrier, with A fixed at compile
// time so that the owners' register indices (a = j/32, b = A) are static.
// NTC = threads/32 column groups: thread (tr, tc) = (t % 32, t / 32) owns
// the elements (i, k) = (tr + 32a, tc + NTC b), a < 4, b < 128/NTC.
// Columns j and j+1 (r even) belong to one wave (tc = r: lanes 0..31,
// tc = r+1: lanes 32..63): it factors column j, applies it to column j+1,
// factors column j+1 and publishes both; after the barrier every thread
// applies column j, then column j+1.  Each element sees the reference's
// operations in the reference's order, so the result is unchanged bit for
// bit; half the barriers.
template <int A, int NTC>
__device__ __forceinline__ bool potrf_col_group(double (&e)[4][128 / NTC], int n, int tr, int tc, int wave,
                                                double (*col)[kPotrfMaxN], int* sfail) {
#pragma clang fp contract(off)
    constexpr int a_own = NTC * A / 32;  // row block of the pivots (j / 32)
    for (int r = 0; r < NTC; r += 2) {
        const int j = NTC * A + r;
        if (j >= n) return true;
        double* cj = col[((j >> 1) & 1) * 2];   // two pairs of buffers by step parity: this
        double* cj1 = cj + kPotrfMaxN;          // step's writers never meet the last step's readers
        const bool two = j + 1 < n;
        if (wave == (r >> 1)) {
            // column j: owners tc == r (lanes 0..31)
            const double ajj = readlane_d(e[a_own][A], j & 31);
            const bool okj = ajj > 0.0;
            if (tc == r) {
                if (!okj) {
                    if (tr == 0) *sfail = j + 1;
                } else {
                    const double alpha = sqrt(ajj);
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int i = tr + 32 * a;
                        if (i > j && i < n) {
                            e[a][A] = e[a][A] / alpha;
                            cj[i] = e[a][A];
                        } else if (i == j) {
                            e[a][A] = alpha;
                        }
                    }
                }
            }
            // the column's entries are read by the other half of this wave
            __builtin_amdgcn_wave_barrier();
            if (okj && two) {
                // column j+1 (owners tc == r + 1, lanes 32..63): column j's
                // update first, then its own pivot
                const double lkj = cj[j + 1];
                if (tc == r + 1) {
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        if (32 * a + 31 <= j) continue;  // uniform
                        const int i = tr + 32 * a;
                        if (i >= j + 1 && i < n) e[a][A] = e[a][A] - cj[i] * lkj;
                    }
                }
                const double a11 = readlane_d(e[a_own][A], 32 + ((j + 1) & 31));
                if (tc == r + 1) {
                    if (!(a11 > 0.0)) {
                        if (tr == 0) *sfail = j + 2;
                    } else {
                        const double alpha = sqrt(a11);
#pragma unroll
                        for (int a = 0; a < 4; ++a) {
                            const int i = tr + 32 * a;
                            if (i > j + 1 && i < n) {
                                e[a][A] = e[a][A] / alpha;
                                cj1[i] = e[a][A];
                            } else if (i == j + 1) {
                                e[a][A] = alpha;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        const int f = *sfail;
        if (f != 0 && f <= j + 1) return false;  // column j failed: nothing of it is applied
        // column j (the owners already applied it to column j+1)
        potrf_rank1<A, NTC>(e, n, tr, tc, j, cj, two ? r + 1 : -1);
        if (f != 0) return false;  // column j+1 failed after column j's update
        if (two) potrf_rank1<A, NTC>(e, n, tr, tc, j + 1, cj1, -1);
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
    __shared__ double col[4][kPotrfMaxN];
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
