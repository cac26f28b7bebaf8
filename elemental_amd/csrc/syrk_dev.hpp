// Device code of the trailing-update tiles (FP64, FP32, split-BF16x3), shared by
// the single-matrix and batched launches.
#pragma once

#include "common.hpp"

namespace elemamd {

constexpr int kSyrkTile = 128;
constexpr int kSyrkKC = 16;
constexpr int kSyrkLdp = kSyrkTile + 16;  // 144 doubles: k-rows start 128 B apart in bank space

// Tile coordinates of this workgroup; false if the tile lies above the diagonal.
__device__ inline bool syrk_tile(int m, int n, int& ti, int& tj);

__device__ inline void tri_tile(int t, int& ti, int& tj) {
    // t = ti*(ti+1)/2 + tj, 0 <= tj <= ti
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((i64)(r + 1) * (r + 2) / 2 <= t) ++r;
    while ((i64)r * (r + 1) / 2 > t) --r;
    ti = r;
    tj = t - (int)((i64)r * (r + 1) / 2);
}

__device__ inline bool syrk_tile(int m, int n, int& ti, int& tj) {
    if (m == n) {
        tri_tile(blockIdx.x, ti, tj);
        return true;
    }
    ti = blockIdx.x;
    tj = blockIdx.y;
    return ti >= tj;
}

// ---------------------------------------------------------------- FP64
// K chunks of 8, double-buffered: 36 KB of LDS per workgroup, so the
// factorization's chain kernels fit on a CU beside two of them (the launches
// beside the chain reserve extra LDS so that no third one is placed: see
// launch_syrk_ln_f64; chunks of 32 with one workgroup per CU measured 10-15%
// slower).  With beta == 1 (the factorization) the accumulators start from
// C, read at the start under the first chunk's loads, and the tile is only
// stored at the end.
constexpr int kSyrkKC64 = 8;

// LDS from the caller: sI and sJ, each 2 x kSyrkKC64 x kSyrkLdp doubles.
__device__ __forceinline__ void syrk_tile_f64_lds(int m, int n, int k, double alpha,
        const double* __restrict__ A, i64 lda, const double* __restrict__ B, i64 ldb, double beta,
        double* __restrict__ C, i64 ldc, int ti, int tj, double (*sI)[kSyrkKC64][kSyrkLdp],
        double (*sJ)[kSyrkKC64][kSyrkLdp]) {
    constexpr int KC = kSyrkKC64;
    const int i0 = ti * kSyrkTile, j0 = tj * kSyrkTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wj = (w >> 1) * 64;  // MFMA M offset = C column offset
    const int wi = (w & 1) * 64;   // MFMA N offset = C row offset
    const bool skip = (ti == tj) && (wj > wi);
    const bool acc_c = (beta == 1.0);

    // global -> register staging: thread owns tile row lr, k rows lk + 2q
    const int lr = tid & 127;
    const int lk = tid >> 7;
    const bool rowI = (i0 + lr) < m, rowJ = (j0 + lr) < n;
    const double* pI = A + (i64)(i0 + lr);
    const double* pJ = B + (i64)(j0 + lr);
    double gI[KC / 2], gJ[KC / 2];

    auto gload = [&](int kc0) {
#pragma unroll
        for (int q = 0; q < KC / 2; ++q) {
            const int kk = kc0 + lk + 2 * q;
            const bool kin = kk < k;
            gI[q] = (rowI && kin) ? pI[(i64)kk * lda] : 0.0;
            gJ[q] = (rowJ && kin) ? pJ[(i64)kk * ldb] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < KC / 2; ++q) {
            sI[buf][lk + 2 * q][lr] = gI[q];
            sJ[buf][lk + 2 * q][lr] = gJ[q];
        }
    };

    d4_t acc[4][4];
    gload(0);
    if (acc_c && !skip) {
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = j0 + wj + 16 * mb + (lane >> 4) + 4 * v;  // f64 MFMA: row = lane/16 + 4*v
                    acc[mb][nb][v] = (i < m && j < n && j <= i) ? C[(i64)i + (i64)j * ldc] : 0.0;
                }
            }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = (d4_t){0.0, 0.0, 0.0, 0.0};
    }
    const double ascale = acc_c ? alpha : 1.0;
    sstore(0);
    __syncthreads();
    const int nchunks = (int)ceil_div(k, KC);
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = (c + 1) < nchunks;
        if (more) gload((c + 1) * KC);
        if (!skip) {
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const int kr = ks * 4 + (lane >> 4);
                double fa[4], fb[4];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) fa[mb] = ascale * sJ[buf][kr][wj + 16 * mb + (lane & 15)];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) fb[nb] = sI[buf][kr][wi + 16 * nb + (lane & 15)];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                    for (int nb = 0; nb < 4; ++nb)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[mb], fb[nb], acc[mb][nb], 0, 0, 0);
            }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
    }
    if (skip) return;
    // dependent zero-operand MFMAs before the accumulators are read (the
    // compiler's wait states after a loop back edge are short for FP64 MFMA)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc[mb][nb], 0, 0, 0);

#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wj + 16 * mb + (lane >> 4) + 4 * v;  // f64 MFMA: row = lane/16 + 4*v
                if (i < m && j < n && j <= i) {
                    double* p = C + (i64)i + (i64)j * ldc;
                    if (acc_c) *p = acc[mb][nb][v];
                    else *p = (beta == 0.0) ? alpha * acc[mb][nb][v] : beta * (*p) + alpha * acc[mb][nb][v];
                }
            }
        }
    }
}


__device__ __forceinline__ void syrk_tile_f64(int m, int n, int k, double alpha,
        const double* __restrict__ A, i64 lda, const double* __restrict__ B, i64 ldb, double beta,
        double* __restrict__ C, i64 ldc, int ti, int tj) {
    __shared__ double sI[2][kSyrkKC64][kSyrkLdp];
    __shared__ double sJ[2][kSyrkKC64][kSyrkLdp];
    syrk_tile_f64_lds(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj, sI, sJ);
}

// ---------------------------------------------------------------- FP32
constexpr int kSyrkLdpF = kSyrkTile + 16;  // 144 floats: k-rows 16 banks apart

// LDS from the caller: sI and sJ, each 2 x kSyrkKC x kSyrkLdpF floats.
__device__ __forceinline__ void syrk_tile_f32_lds(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, int ti, int tj, float (*sI)[kSyrkKC][kSyrkLdpF],
        float (*sJ)[kSyrkKC][kSyrkLdpF]) {
    const int i0 = ti * kSyrkTile, j0 = tj * kSyrkTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wj = (w >> 1) * 64, wi = (w & 1) * 64;
    const bool skip = (ti == tj) && (wj > wi);
    const int lr = tid & 127, lk = tid >> 7;
    const bool rowI = (i0 + lr) < m, rowJ = (j0 + lr) < n;
    const float* pI = A + (i64)(i0 + lr);
    const float* pJ = B + (i64)(j0 + lr);
    float gI[8], gJ[8];
    auto gload = [&](int kc0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int kk = kc0 + lk + 2 * q;
            const bool kin = kk < k;
            gI[q] = (rowI && kin) ? pI[(i64)kk * lda] : 0.0f;
            gJ[q] = (rowJ && kin) ? pJ[(i64)kk * ldb] : 0.0f;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            sI[buf][lk + 2 * q][lr] = gI[q];
            sJ[buf][lk + 2 * q][lr] = gJ[q];
        }
    };
    f4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
    gload(0);
    sstore(0);
    __syncthreads();
    const int nchunks = (int)ceil_div(k, kSyrkKC);
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = (c + 1) < nchunks;
        if (more) gload((c + 1) * kSyrkKC);
        if (!skip) {
#pragma unroll
            for (int ks = 0; ks < kSyrkKC / 4; ++ks) {
                const int kr = ks * 4 + (lane >> 4);
                float fa[4], fb[4];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) fa[mb] = sJ[buf][kr][wj + 16 * mb + (lane & 15)];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) fb[nb] = sI[buf][kr][wi + 16 * nb + (lane & 15)];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                    for (int nb = 0; nb < 4; ++nb)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[mb], fb[nb], acc[mb][nb], 0, 0, 0);
            }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
    }
    if (skip) return;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wj + 16 * mb + 4 * (lane >> 4) + v;
                if (i < m && j < n && j <= i) {
                    float* p = C + (i64)i + (i64)j * ldc;
                    const float upd = alpha * acc[mb][nb][v];
                    *p = (beta == 0.0f) ? upd : beta * (*p) + upd;
                }
            }
        }
}

// ------------------------------------------------------- split-BF16x3
// Each float x of the panel is split as hi = bf16(x), lo = bf16(x - hi);
// products use hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-16 relative, is
// dropped), each a v_mfma_f32_16x16x32_bf16 with fp32 accumulation.
// LDS holds row-major [row][k] planes so one lane fetches its 8 consecutive
// k values (16 bytes) per operand.
__device__ __forceinline__ void syrk_tile_f32(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, int ti, int tj) {
    __shared__ float sI[2][kSyrkKC][kSyrkLdpF];
    __shared__ float sJ[2][kSyrkKC][kSyrkLdpF];
    syrk_tile_f32_lds(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj, sI, sJ);
}

constexpr int kBfKC = 32;
constexpr int kBfLd = kBfKC + 8;  // 40 bf16 = 80 B per row: 8 lanes of 16 B hit disjoint banks

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// LDS from the caller: four kSyrkTile x kBfLd bf16 arrays, 16-byte aligned.
__device__ __forceinline__ void syrk_tile_bf16x3_lds(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, int ti, int tj, unsigned short (*sIh)[kBfLd], unsigned short (*sIl)[kBfLd],
        unsigned short (*sJh)[kBfLd], unsigned short (*sJl)[kBfLd]) {
    const int i0 = ti * kSyrkTile, j0 = tj * kSyrkTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wj = (w >> 1) * 64, wi = (w & 1) * 64;
    const bool skip = (ti == tj) && (wj > wi);
    const int lr = tid & 127, kh = tid >> 7;  // row, k half (16 values)
    const bool rowI = (i0 + lr) < m, rowJ = (j0 + lr) < n;
    const float* pI = A + (i64)(i0 + lr);
    const float* pJ = B + (i64)(j0 + lr);
    float gI[16], gJ[16];
    auto gload = [&](int kc0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int kk = kc0 + kh * 16 + q;
            const bool kin = kk < k;
            gI[q] = (rowI && kin) ? pI[(i64)kk * lda] : 0.0f;
            gJ[q] = (rowJ && kin) ? pJ[(i64)kk * ldb] : 0.0f;
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const unsigned short hI = f32_to_bf16_rne(gI[q]);
            const unsigned short hJ = f32_to_bf16_rne(gJ[q]);
            sIh[lr][kh * 16 + q] = hI;
            sIl[lr][kh * 16 + q] = f32_to_bf16_rne(gI[q] - bf16_to_f32(hI));
            sJh[lr][kh * 16 + q] = hJ;
            sJl[lr][kh * 16 + q] = f32_to_bf16_rne(gJ[q] - bf16_to_f32(hJ));
        }
    };
    f4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
    const int nchunks = (int)ceil_div(k, kBfKC);
    gload(0);
    for (int c = 0; c < nchunks; ++c) {
        __syncthreads();  // previous chunk's LDS reads are done
        sstore();
        __syncthreads();
        if (c + 1 < nchunks) gload((c + 1) * kBfKC);
        if (!skip) {
            const int kofs = 8 * (lane >> 4);
            s8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const int r = wj + 16 * mb + (lane & 15);
                ah[mb] = *reinterpret_cast<const s8_t*>(&sJh[r][kofs]);
                al[mb] = *reinterpret_cast<const s8_t*>(&sJl[r][kofs]);
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int r = wi + 16 * nb + (lane & 15);
                bh[nb] = *reinterpret_cast<const s8_t*>(&sIh[r][kofs]);
                bl[nb] = *reinterpret_cast<const s8_t*>(&sIl[r][kofs]);
            }
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    f4_t a = acc[mb][nb];
                    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, al[mb]),
                                                                __builtin_bit_cast(bf8_t, bh[nb]), a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, ah[mb]),
                                                                __builtin_bit_cast(bf8_t, bl[nb]), a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, ah[mb]),
                                                                __builtin_bit_cast(bf8_t, bh[nb]), a, 0, 0, 0);
                    acc[mb][nb] = a;
                }
        }
    }
    if (skip) return;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wj + 16 * mb + 4 * (lane >> 4) + v;
                if (i < m && j < n && j <= i) {
                    float* p = C + (i64)i + (i64)j * ldc;
                    const float upd = alpha * acc[mb][nb][v];
                    *p = (beta == 0.0f) ? upd : beta * (*p) + upd;
                }
            }
        }
}


__device__ __forceinline__ void syrk_tile_bf16x3(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, int ti, int tj) {
    __shared__ __attribute__((aligned(16))) unsigned short sIh[kSyrkTile][kBfLd];
    __shared__ __attribute__((aligned(16))) unsigned short sIl[kSyrkTile][kBfLd];
    __shared__ __attribute__((aligned(16))) unsigned short sJh[kSyrkTile][kBfLd];
    __shared__ __attribute__((aligned(16))) unsigned short sJl[kSyrkTile][kBfLd];
    syrk_tile_bf16x3_lds(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj, sIh, sIl, sJh, sJl);
}

// Split-BF16 tile from pre-split panels.  The panel is stored as NP bf16
// planes, plane p at planes + p*ps: plane 0 = bf16(x), plane p = bf16 of
// what planes 0..p-1 leave of x (row-major, ldk >= k, k contiguous, rows of
// the panel = rows of C).  NP = 2 is split-BF16x3 (hi*hi + hi*lo + lo*hi,
// the terms down to 2^-16 of a product), NP = 3 split-BF16x6 (the six terms
// p + q <= 2, down to 2^-24).  The bf16 MFMA sums its 32 products without
// rounding bits below the largest product's 24-bit window (bits of smaller
// products there are dropped, the accumulator add itself rounds to nearest),
// so split-BF16x6 lands between split-BF16x3 and FP32 MFMA accuracy; the
// drop is biased toward zero per product and grows with k.  The split is done once per
// pass over the trailing matrix instead of once per tile; the tile only
// moves 16-byte groups of 8 bf16 from memory to LDS.  Smallest terms first.
template <int NP>
__device__ __forceinline__ void syrk_tile_bf16_planes(int m, int n, int k, float alpha,
        const unsigned short* __restrict__ planes, i64 ps, i64 ldk, float beta, float* __restrict__ C, i64 ldc,
        int ti, int tj) {
    __shared__ __attribute__((aligned(16))) unsigned short sI[NP][kSyrkTile][kBfLd];
    __shared__ __attribute__((aligned(16))) unsigned short sJ[NP][kSyrkTile][kBfLd];
    const int i0 = ti * kSyrkTile, j0 = tj * kSyrkTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wj = (w >> 1) * 64, wi = (w & 1) * 64;
    const bool skip = (ti == tj) && (wj > wi);
    // staging: four threads per row read its 32-wide chunk (64 contiguous
    // bytes per plane), rows lrow and lrow + 64
    const int lq = tid & 3, lrow = tid >> 2;
    bool rowI[2], rowJ[2];
    const unsigned short *pI[2], *pJ[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = lrow + 64 * h;
        rowI[h] = (i0 + r) < m;
        rowJ[h] = (j0 + r) < n;
        pI[h] = planes + (i64)(rowI[h] ? i0 + r : 0) * ldk + 8 * lq;
        pJ[h] = planes + (i64)(rowJ[h] ? j0 + r : 0) * ldk + 8 * lq;
    }
    s8_t gI[NP][2], gJ[NP][2];  // [plane][row pass]
    const s8_t zero = {0, 0, 0, 0, 0, 0, 0, 0};
    auto gload = [&](int kc0) {
        const bool kin = kc0 + 8 * lq < k;  // ldk pads to a multiple of 32 with zeros
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                gI[p][h] = (rowI[h] && kin) ? *reinterpret_cast<const s8_t*>(pI[h] + p * ps + kc0) : zero;
                gJ[p][h] = (rowJ[h] && kin) ? *reinterpret_cast<const s8_t*>(pJ[h] + p * ps + kc0) : zero;
            }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                *reinterpret_cast<s8_t*>(&sI[p][lrow + 64 * h][8 * lq]) = gI[p][h];
                *reinterpret_cast<s8_t*>(&sJ[p][lrow + 64 * h][8 * lq]) = gJ[p][h];
            }
    };
    // C - P P^T (the factorization): accumulate onto C, read while the first
    // chunk loads, with the J operands negated (sign bits flipped, exact)
    const bool acc_c = (beta == 1.0f && alpha == -1.0f);
    f4_t acc[4][4];
    if (acc_c && !skip) {
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = j0 + wj + 16 * mb + 4 * (lane >> 4) + v;
                    acc[mb][nb][v] = (i < m && j < n && j <= i) ? C[(i64)i + (i64)j * ldc] : 0.0f;
                }
            }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
    }
    const s8_t neg = {(short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000,
                      (short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000};
    const int nchunks = (int)ceil_div(k, kBfKC);
    gload(0);
    for (int c = 0; c < nchunks; ++c) {
        __syncthreads();  // previous chunk's LDS reads are done
        sstore();
        __syncthreads();
        if (c + 1 < nchunks) gload((c + 1) * kBfKC);
        if (!skip) {
            const int kofs = 8 * (lane >> 4);
            s8_t fa[NP][4], fb[NP][4];
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) {
                    fa[p][mb] = *reinterpret_cast<const s8_t*>(&sJ[p][wj + 16 * mb + (lane & 15)][kofs]);
                    if (acc_c) fa[p][mb] ^= neg;  // flip the bf16 sign bits
                }
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
                    fb[p][nb] = *reinterpret_cast<const s8_t*>(&sI[p][wi + 16 * nb + (lane & 15)][kofs]);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    f4_t a = acc[mb][nb];
#pragma unroll
                    for (int o = NP - 1; o >= 0; --o)  // term order o = p + q, smallest first
#pragma unroll
                        for (int p = o; p >= 0; --p)
                            a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, fa[p][mb]),
                                                                        __builtin_bit_cast(bf8_t, fb[o - p][nb]), a,
                                                                        0, 0, 0);
                    acc[mb][nb] = a;
                }
        }
    }
    if (skip) return;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wj + 16 * mb + 4 * (lane >> 4) + v;
                if (i < m && j < n && j <= i) {
                    float* p = C + (i64)i + (i64)j * ldc;
                    if (acc_c) *p = acc[mb][nb][v];
                    else *p = (beta == 0.0f) ? alpha * acc[mb][nb][v] : beta * (*p) + alpha * acc[mb][nb][v];
                }
            }
        }
}

}  // namespace elemamd
