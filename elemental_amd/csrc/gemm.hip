// C := beta*C + alpha * A^T * B with A k x m and B k x n (column-major),
// optionally restricted to the lower or the upper triangle of a square C.
//
// This is the product of Elemental's LocalTrrk / Trrk(LOWER, TRANSPOSE,
// NORMAL, ...) (src/blas-like/Trrk/Local.hpp:733-775 and :931-976, leaf
// kernels :452-536), the triangular rank-k update the DistMatrix route of the
// blocked Cholesky uses for its trailing matrix
// (include/elemental/lapack-like/Cholesky/LVar3Square.hpp).  The reference
// recurses by halves down to 64x64 leaves and then does full square GEMMs
// plus AxpyTriangle; on the GPU the lower tiles are simply enumerated and
// diagonal tiles mask their upper part, which gives the same lower trapezoid.
//
// Tiling as in the SYRK kernel: 128x128 tiles, 4 waves of 64x64,
// v_mfma_f64_16x16x4_f64, double-buffered LDS chunks of 16 along k.
// Operand columns are contiguous in k here, so each 16 lanes of a load
// instruction fetch one 128-byte column segment.
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

constexpr int kGemmTile = 128;
constexpr int kGemmKC = 16;
constexpr int kGemmLdp = kGemmTile + 16;

__device__ inline void gemm_tri_tile(int t, int& ti, int& tj) {
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((i64)(r + 1) * (r + 2) / 2 <= t) ++r;
    while ((i64)r * (r + 1) / 2 > t) --r;
    ti = r;
    tj = t - (int)((i64)r * (r + 1) / 2);
}

__global__ __launch_bounds__(256, 2) void gemm_tn_f64_kernel(int m, int n, int k, double alpha,
                                                             const double* __restrict__ A, i64 lda,
                                                             const double* __restrict__ B, i64 ldb, double beta,
                                                             double* __restrict__ C, i64 ldc, int tri) {
    __shared__ double sA[2][kGemmKC][kGemmLdp];  // sA[l][i] = A(l, i0+i)
    __shared__ double sB[2][kGemmKC][kGemmLdp];  // sB[l][j] = B(l, j0+j)
    // tri: 0 every tile; 1 the lower triangle (tiles ti >= tj, j <= i);
    // 2 the upper triangle (tiles ti <= tj, i <= j)
    int ti, tj;
    if (tri == 1) {
        gemm_tri_tile(blockIdx.x, ti, tj);
    } else if (tri == 2) {
        gemm_tri_tile(blockIdx.x, tj, ti);
    } else {
        ti = blockIdx.x;
        tj = blockIdx.y;
    }
    const int i0 = ti * kGemmTile, j0 = tj * kGemmTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wj = (w >> 1) * 64, wi = (w & 1) * 64;
    const bool skip = (ti == tj) && ((tri == 1 && wj > wi) || (tri == 2 && wi > wj));

    // loader: thread -> index column c = tid >> 1, l half (tid & 1) * 8 .. +8
    const int lc = tid >> 1, lh = (tid & 1) * 8;
    const bool colA = (i0 + lc) < m, colB = (j0 + lc) < n;
    const double* pA = A + (i64)(i0 + lc) * lda;
    const double* pB = B + (i64)(j0 + lc) * ldb;
    double gA[8], gB[8];
    auto gload = [&](int kc0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int l = kc0 + lh + q;
            const bool lin = l < k;
            gA[q] = (colA && lin) ? pA[l] : 0.0;
            gB[q] = (colB && lin) ? pB[l] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            sA[buf][lh + q][lc] = gA[q];
            sB[buf][lh + q][lc] = gB[q];
        }
    };
    d4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (d4_t){0.0, 0.0, 0.0, 0.0};
    gload(0);
    sstore(0);
    __syncthreads();
    const int nchunks = (int)ceil_div(k, kGemmKC);
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = (c + 1) < nchunks;
        if (more) gload((c + 1) * kGemmKC);
        if (!skip) {
#pragma unroll
            for (int ks = 0; ks < kGemmKC / 4; ++ks) {
                const int kr = ks * 4 + (lane >> 4);
                double fa[4], fb[4];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) fa[mb] = sB[buf][kr][wj + 16 * mb + (lane & 15)];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) fb[nb] = sA[buf][kr][wi + 16 * nb + (lane & 15)];
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
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int i = i0 + wi + 16 * nb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wj + 16 * mb + (lane >> 4) + 4 * v;  // f64 MFMA: row = lane/16 + 4*v
                if (i < m && j < n && (tri == 0 || (tri == 1 ? j <= i : i <= j))) {
                    double* p = C + (i64)i + (i64)j * ldc;
                    const double upd = alpha * acc[mb][nb][v];
                    *p = (beta == 0.0) ? upd : beta * (*p) + upd;
                }
            }
        }
}

hipError_t launch_gemm_tn_f64(int m, int n, int k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                              double beta, double* C, i64 ldc, int tri, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (tri != 0) {  // square C
        const i64 T = ceil_div(m, kGemmTile);
        hipLaunchKernelGGL(gemm_tn_f64_kernel, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, m, n, k, alpha, A,
                           lda, B, ldb, beta, C, ldc, tri);
    } else {
        hipLaunchKernelGGL(gemm_tn_f64_kernel, dim3((unsigned)ceil_div(m, kGemmTile), (unsigned)ceil_div(n, kGemmTile)),
                           dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
    }
    return hipGetLastError();
}

}  // namespace elemamd
