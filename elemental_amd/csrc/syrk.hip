// Trailing update of the blocked Cholesky (the dominant kernel):
//   C := beta*C + alpha*A*B^T on the lower trapezoid of C (entries with
//   j <= i), C m x n with n <= m, A m x k, B n x k.  With B = A and m = n this is
//   the lower SYRK; with n = nb it is the look-ahead update of the next block
//   column, whose top nb x nb block is itself lower triangular.
// This is Herk(LOWER, NORMAL, -1, A21, 1, A22)
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:114) which the reference
// hands to dsyrk_('L','N') (include/elemental/blas-like/level3/Syrk.hpp:45-48,
// src/core/imports/blas.cpp:1043-1050).  The strictly upper triangle of C is
// never read or written.
//
// CDNA4 mapping: one 256-thread workgroup per 128x128 lower tile of C (for a
// square C the grid enumerates only the T(T+1)/2 lower tiles; otherwise it is
// Tm x Tn and tiles above the diagonal exit), four waves each owning a 64x64
// quadrant as 4x4 MFMA blocks
// of 16x16.  K is streamed through double-buffered LDS in chunks of 16.
// MFMA "rows" index C's columns and MFMA "columns" index C's rows, so each
// lane's four accumulator values land in four columns at one row and the 16
// lanes of a row-group write 128 contiguous bytes.
//
// Three arithmetic modes share the tiling:
//   F64    : v_mfma_f64_16x16x4_f64 on double storage
//   F32    : v_mfma_f32_16x16x4_f32 on float storage
//   BF16X3 : split-bf16 (hi + lo) of the float panel, three products
//            hi*hi + hi*lo + lo*hi with v_mfma_f32_16x16x32_bf16, fp32 accumulate
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "syrk_dev.hpp"
#include "solve_dev.hpp"

namespace elemamd {

// Number of leading panel columns still to apply after a factorization
// failure: the panels are global columns [kcol0, kcol0 + k) of the factor,
// and a failing pivot at column c (info = c + 1) voids the nbp-block holding
// c and every later one, exactly the panels the reference never computes
// before it throws.  Without info (or info == 0) all k columns apply.
__device__ inline int syrk_valid_k(const int* info, int k, int kcol0, int nbp) {
    if (info == nullptr) return k;
    const int v = *info;
    if (v == 0) return k;
    const int lim = ((v - 1) / nbp) * nbp - kcol0;
    return lim < 0 ? 0 : (lim < k ? lim : k);
}

__global__ __launch_bounds__(256, 2) void syrk_ln_f64_kernel(int m, int n, int k, double alpha,
        const double* __restrict__ A, i64 lda, const double* __restrict__ B, i64 ldb, double beta,
        double* __restrict__ C, i64 ldc, const int* __restrict__ info, int kcol0, int nbp, int chain) {
    // a launch the factorization's chain waits for issues ahead of the
    // concurrent bulk update's waves (wave priority, not queue priority)
    if (chain) __builtin_amdgcn_s_setprio(3);
    const int ke = syrk_valid_k(info, k, kcol0, nbp);
    if (info != nullptr && ke == 0) return;
    int ti, tj;
    if (!syrk_tile(m, n, ti, tj)) return;
    syrk_tile_f64(m, n, ke, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj);
}

__global__ __launch_bounds__(256, 2) void syrk_ln_f32_kernel(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, const int* __restrict__ info, int kcol0, int nbp) {
    const int ke = syrk_valid_k(info, k, kcol0, nbp);
    if (info != nullptr && ke == 0) return;
    int ti, tj;
    if (!syrk_tile(m, n, ti, tj)) return;
    syrk_tile_f32(m, n, ke, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj);
}

__global__ __launch_bounds__(256, 2) void syrk_ln_bf16x3_kernel(int m, int n, int k, float alpha,
        const float* __restrict__ A, i64 lda, const float* __restrict__ B, i64 ldb, float beta,
        float* __restrict__ C, i64 ldc, const int* __restrict__ info, int kcol0, int nbp) {
    const int ke = syrk_valid_k(info, k, kcol0, nbp);
    if (info != nullptr && ke == 0) return;
    int ti, tj;
    if (!syrk_tile(m, n, ti, tj)) return;
    syrk_tile_bf16x3(m, n, ke, alpha, A, lda, B, ldb, beta, C, ldc, ti, tj);
}

static dim3 syrk_grid(int m, int n) {
    const i64 Tm = ceil_div(m, kSyrkTile), Tn = ceil_div(n, kSyrkTile);
    if (m == n) return dim3((unsigned)(Tm * (Tm + 1) / 2));
    return dim3((unsigned)Tm, (unsigned)Tn);
}

hipError_t launch_syrk_ln_f64(int m, int n, int k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                              double beta, double* C, i64 ldc, const int* info, int kcol0, int nbp, hipStream_t s,
                              bool beside_chain) {
    if (m <= 0 || n <= 0) return hipSuccess;
    // beside the chain: 56 KB of LDS per workgroup (36 KB used), so at most
    // two share a CU and 48 KB stay free for the chain's kernels; every other
    // launch of the factorization is on its chain
    const unsigned pad = beside_chain ? 20u * 1024u : 0u;
    hipLaunchKernelGGL(syrk_ln_f64_kernel, syrk_grid(m, n), dim3(256), pad, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc, info, kcol0, nbp, (info != nullptr && !beside_chain) ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_syrk_ln_f32(bool bf16x3, int m, int n, int k, float alpha, const float* A, i64 lda, const float* B,
                              i64 ldb, float beta, float* C, i64 ldc, const int* info, int kcol0, int nbp, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (bf16x3)
        hipLaunchKernelGGL(syrk_ln_bf16x3_kernel, syrk_grid(m, n), dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta,
                           C, ldc, info, kcol0, nbp);
    else
        hipLaunchKernelGGL(syrk_ln_f32_kernel, syrk_grid(m, n), dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                           ldc, info, kcol0, nbp);
    return hipGetLastError();
}


// np bf16 planes (row-major, k contiguous, ldk a multiple of 32, zero padded;
// plane p at planes + p*ps) of a float panel P (m x k, column-major): plane 0
// = bf16(x), plane p = bf16 of what the earlier planes leave of x.
// Thread = (row, 8 k).
__global__ __launch_bounds__(256) void split_bf16_kernel(int np, int m, int k, const float* __restrict__ P, i64 ldp,
                                                         unsigned short* __restrict__ planes, i64 ps, i64 ldk) {
    const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
    const int k0 = blockIdx.y * 8;
    if (r >= m) return;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = k0 + q < k ? P[r + (i64)(k0 + q) * ldp] : 0.0f;
    for (int p = 0; p < np; ++p) {
        s8_t h;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned short hb = f32_to_bf16_rne(x[q]);
            h[q] = (short)hb;
            x[q] -= bf16_to_f32(hb);
        }
        *reinterpret_cast<s8_t*>(planes + p * ps + r * ldk + k0) = h;
    }
}

template <int NP>
__global__ __launch_bounds__(256, 2) void syrk_ln_bf16p_kernel(int m, int n, int k, float alpha,
        const unsigned short* __restrict__ planes, i64 ps, i64 ldk, float beta, float* __restrict__ C, i64 ldc,
        const int* __restrict__ info, int kcol0, int nbp, int chain) {
    if (chain) __builtin_amdgcn_s_setprio(3);  // the factorization's chain waits for this launch
    const int ke = syrk_valid_k(info, k, kcol0, nbp);
    if (info != nullptr && ke == 0) return;
    int ti, tj;
    if (!syrk_tile(m, n, ti, tj)) return;
    syrk_tile_bf16_planes<NP>(m, n, ke, alpha, planes, ps, ldk, beta, C, ldc, ti, tj);
}

hipError_t launch_split_bf16(int np, int m, int k, const float* P, i64 ldp, unsigned short* planes, i64 ps, i64 ldk,
                             hipStream_t s) {
    if (m <= 0 || k <= 0) return hipSuccess;
    if (np < 1 || np > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)(ldk / 8)), dim3(256), 0, s, np,
                       m, k, P, ldp, planes, ps, ldk);
    return hipGetLastError();
}

hipError_t launch_syrk_ln_bf16_planes(int np, int m, int n, int k, float alpha, const unsigned short* planes, i64 ps,
                                      i64 ldk, float beta, float* C, i64 ldc, const int* info, int kcol0, int nbp,
                                      hipStream_t s, int chain) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (np == 2)
        hipLaunchKernelGGL(syrk_ln_bf16p_kernel<2>, syrk_grid(m, n), dim3(256), 0, s, m, n, k, alpha, planes, ps, ldk,
                           beta, C, ldc, info, kcol0, nbp, chain);
    else if (np == 3)
        hipLaunchKernelGGL(syrk_ln_bf16p_kernel<3>, syrk_grid(m, n), dim3(256), 0, s, m, n, k, alpha, planes, ps, ldk,
                           beta, C, ldc, info, kcol0, nbp, chain);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// The diagonal block alone: C (ncol x ncol, lower) -= P P^T, P = the first
// ncol rows of the planes, k columns -- the update the panel step waits for.
// One 16 x 16 output tile per workgroup (the lower ones), its four waves
// splitting the k chunks (each wave issues the loads of up to 4 chunks at
// once, then multiplies), partial sums added in LDS in a fixed order.
template <int NP>
__global__ __launch_bounds__(256) void syrk_diag_bf16p_kernel(int ncol, int k, const unsigned short* __restrict__ planes,
                                                              i64 ps, i64 ldk, float* __restrict__ C, i64 ldc,
                                                              const int* __restrict__ info, int kcol0, int nbp) {
    __builtin_amdgcn_s_setprio(3);  // chain kernel: issues ahead of the trailing update's waves
    const int ke = syrk_valid_k(info, k, kcol0, nbp);
    if (info != nullptr && ke == 0) return;
    __shared__ f4_t part[3][64];
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m16 = lane & 15, g = lane >> 4;
    const int ra = 16 * ti + m16 < ncol ? 16 * ti + m16 : ncol - 1;  // clamped rows: their outputs are dropped
    const int rb = 16 * tj + m16 < ncol ? 16 * tj + m16 : ncol - 1;
    const unsigned short* pa = planes + (i64)ra * ldk + 8 * g;
    const unsigned short* pb = planes + (i64)rb * ldk + 8 * g;
    constexpr int kQ = 4;  // chunks per wave in flight
    const int nchunks = (ke + 31) / 32;  // ldk pads with zeros to a multiple of 32
    f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = w; c0 < nchunks; c0 += 4 * kQ) {
        s8_t fa[kQ][NP], fb[kQ][NP];
        const s8_t zero = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int c = c0 + 4 * q;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                fa[q][p] = c < nchunks ? *reinterpret_cast<const s8_t*>(pa + p * ps + 32 * c) : zero;
                fb[q][p] = c < nchunks ? *reinterpret_cast<const s8_t*>(pb + p * ps + 32 * c) : zero;
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q)
#pragma unroll
            for (int o = NP - 1; o >= 0; --o)  // term order o = p + q, smallest first
#pragma unroll
                for (int p = o; p >= 0; --p) mfma_acc_bf16(acc, fa[q][p], fb[q][o - p]);
    }
    mfma_retire();
    if (w > 0) part[w - 1][lane] = acc;
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float sum = ((acc[v] + part[0][lane][v]) + part[1][lane][v]) + part[2][lane][v];
        const int r = 16 * ti + 4 * g + v, c = 16 * tj + m16;
        if (r < ncol && c <= r) C[r + (i64)c * ldc] -= sum;
    }
}

hipError_t launch_syrk_diag_bf16_planes(int np, int ncol, int k, const unsigned short* planes, i64 ps, i64 ldk,
                                        float* C, i64 ldc, const int* info, int kcol0, int nbp, hipStream_t s) {
    if (ncol <= 0 || k <= 0) return hipSuccess;
    if (ncol > 128) return hipErrorInvalidValue;
    const int T = (int)ceil_div(ncol, 16);
    const dim3 grid((unsigned)(T * (T + 1) / 2));
    if (np == 2)
        hipLaunchKernelGGL(syrk_diag_bf16p_kernel<2>, grid, dim3(256), 0, s, ncol, k, planes, ps, ldk, C, ldc, info,
                           kcol0, nbp);
    else if (np == 3)
        hipLaunchKernelGGL(syrk_diag_bf16p_kernel<3>, grid, dim3(256), 0, s, ncol, k, planes, ps, ldk, C, ldc, info,
                           kcol0, nbp);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace elemamd
