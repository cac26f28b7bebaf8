// The panel solve B := B L11^{-T} of the blocked factorizations (single
// matrix FP64, batched): rows kept in the FP64 MFMA accumulator layout,
// substitution inside 32-column pieces, MFMA products between pieces (the
// method is described in panel_f64.hip).
//
// Layout of a 16-row x 32-column piece in one wave: a[h] (h = 0, 1) is the
// 16 x 16 tile of columns 16h .. 16h+15; lane (c, g) = (lane % 16, lane / 16)
// holds rows g + 4v (v = 0..3) of column c, which is the output layout of
// v_mfma_f64_16x16x4_f64, so a product with the piece accumulates into it.
#pragma once

#include "common.hpp"
#include "solve_dev.hpp"

namespace elemamd {
namespace chain64 {

// acc += a * b, one FP64 16x16x4 MFMA step with the accumulator tied to the
// destination; read the accumulator only after mfma_retire().
__device__ inline void mfma_acc_f64(d4_t& acc, double a, double b) {
    asm volatile("s_nop 1\n\tv_mfma_f64_16x16x4_f64 %0, %1, %2, %0" : "+v"(acc) : "v"(a), "v"(b));
}

// Lane N of every 16-lane row, broadcast to its row.
template <int N>
__device__ inline double row_bcast(double v) {
    constexpr int ctrl = 0x150 | N;  // DPP row_newbcast:N
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Strict lower triangles of two 32 x 32 diagonal pieces sharing one array:
// the even piece at [k][c] (c > k, upper half), the odd piece at
// [31-k][31-c] (lower half).  l_ck of the piece, or 0 when c <= k.
template <bool ODD>
struct PackedDiag {
    const double (*d)[32];
    __device__ double operator()(int c, int k) const {
        const double v = ODD ? d[31 - k][31 - c] : d[k][c];
        return c > k ? v : 0.0;
    }
};

// Column K of the substitution X = B L^{-T} inside one piece: x = b_K / l_KK
// (by the reciprocal, as the reference BLAS loop), broadcast along the
// 16-lane row (the column's owner is lane K % 16 of every row), then
// b_j -= l_jK x for the later columns j.
template <int K, typename Diag>
__device__ __forceinline__ void sub_step(d4_t (&a)[2], int c, const Diag& L, const double* rinv) {
    constexpr int T = K / 16, KC = K % 16;
    const double rk = rinv[K];
    const double l1 = L(16 + c, K);
    double x[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = row_bcast<KC>(a[T][v] * rk);
    if constexpr (T == 0) {
        const double l0 = L(c, K);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            a[0][v] = (c == KC) ? x[v] : fma(-l0, x[v], a[0][v]);
            a[1][v] = fma(-l1, x[v], a[1][v]);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) a[1][v] = (c == KC) ? x[v] : fma(-l1, x[v], a[1][v]);
    }
}

template <int K, typename Diag>
struct Substitute {
    static __device__ __forceinline__ void run(d4_t (&a)[2], int c, const Diag& L, const double* rinv) {
        sub_step<K, Diag>(a, c, L, rinv);
        Substitute<K + 1, Diag>::run(a, c, L, rinv);
    }
};
template <typename Diag>
struct Substitute<32, Diag> {
    static __device__ __forceinline__ void run(d4_t (&)[2], int, const Diag&, const double*) {}
};

// a (16 rows x 32 columns) -= X Y^T, FP64 MFMA: X (16 x 32) rows xr ..
// xr+15 of a row-major LDS array, Y (32 x 32) with its operands already in
// registers, y0[ks] = Y[lane % 16][4 ks + lane / 16] and y1[ks] =
// Y[16 + lane % 16][4 ks + lane / 16].
template <int LX>
__device__ __forceinline__ void sub_product_regs(d4_t (&a)[2], const double (*X)[LX], int xr, const double (&y0)[8],
                                                 const double (&y1)[8], int lane) {
    const int i = lane & 15, g = lane >> 4;
    double fa[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) fa[ks] = -X[xr + i][4 * ks + g];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        mfma_acc_f64(a[0], fa[ks], y0[ks]);
        mfma_acc_f64(a[1], fa[ks], y1[ks]);
    }
}

// Rows [64 slab, 64 slab + 64) of B (m x b, b <= 128) := B L^{-T} by the
// calling 256-thread workgroup, storage T (double, or float computed in
// double).  LDS from the caller: sDg 2 x 32 x 32 and sXq 4 x 16 x 34
// doubles, sRinv 128 doubles (34 KB).
template <typename T>
__device__ __forceinline__ void panel64_slab(int m, int b, const T* __restrict__ L, i64 ldl, T* __restrict__ B,
                                             i64 ldb, int slab, double (*sDg)[32][32], double* sRinv,
                                             double (*sXq)[16][34]) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, c = lane & 15, g = lane >> 4;
    const int np = (b + 31) / 32;
    const i64 rw = (i64)slab * 64 + 16 * w;  // this wave's first row
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
                const double x = (double)B[(ok ? r : 0) + (i64)(ok ? col : 0) * ldb];
                a[p][h][v] = ok ? x : 0.0;
            }
    // the diagonal pieces of L11: rows past b are zero, the padded diagonal
    // is the identity
    for (int e = t; e < 4 * 1024; e += 256) {
        const int q = e >> 10, cc = e & 31, k = (e >> 5) & 31;
        if (cc <= k) continue;
        const int r = 32 * q + cc, col = 32 * q + k;
        const bool ok = r < b && col < b;
        const double x = (double)L[(ok ? r : 0) + (i64)(ok ? col : 0) * ldl];
        if (q & 1) sDg[q >> 1][31 - k][31 - cc] = ok ? x : 0.0;
        else sDg[q >> 1][k][cc] = ok ? x : 0.0;
    }
    if (t < 128) sRinv[t] = t < b ? 1.0 / (double)L[t + (i64)t * ldl] : 1.0;
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
                    const double x = (double)L[(r < b ? r : 0) + (i64)(32 * q + 4 * ks + g) * ldl];
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
                if (r < m && col < b) B[r + (i64)col * ldb] = (T)a[q][h][v];
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

}  // namespace chain64
}  // namespace elemamd
