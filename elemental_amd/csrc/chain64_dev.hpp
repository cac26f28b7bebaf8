// Device helpers of the FP64 factorization's panel solve: rows kept in the
// FP64 MFMA accumulator layout, substitution inside 32-column pieces, MFMA
// products between pieces.
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

}  // namespace chain64
}  // namespace elemamd
