// C := beta*C + alpha * op(A) * op(B), FP64, op(A) m x k, op(B) k x n
// (column-major), optionally restricted to the lower or the upper triangle
// of a square C, with triangular operand masks.
//
// Uses:
//  * Elemental's LocalTrrk / Trrk(LOWER, TRANSPOSE, NORMAL, ...)
//    (src/blas-like/Trrk/Local.hpp:733-775 and :931-976, leaf kernels
//    :452-536), the triangular rank-k update of the DistMatrix route of the
//    blocked Cholesky (include/elemental/lapack-like/Cholesky/LVar3Square.hpp).
//    The reference recurses by halves down to 64x64 leaves and then does full
//    square GEMMs plus AxpyTriangle; on the GPU the lower tiles are simply
//    enumerated and diagonal tiles mask their upper part, which gives the same
//    lower trapezoid.
//  * elem::Gemm (include/elemental/blas-like/level3/Gemm.hpp) in every
//    orientation, and the products with a triangular factor or its inverse
//    (Trmm-like) of the HPD inverse, the triangular inverse and the
//    two-sided reductions: a masked operand reads as zero outside its
//    triangle (and as one on a unit diagonal), so the other triangle of the
//    stored matrix is never used.
//
// Tiling: 128x128 tiles, 4 waves of 64x64, v_mfma_f64_16x16x4_f64,
// double-buffered LDS chunks of 16 along k.  Global loads follow the stored
// layout: an operand whose k index runs down its columns (A transposed, B not)
// is read as 8 consecutive k values of one column per thread; one whose m or
// n index runs down the columns (A not transposed, B transposed) as 8
// consecutive rows of one column, 16 threads covering one 128-element column
// segment.  Either way a wave's 64 lanes read contiguous 64-byte pieces.
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

// Stored element (r, c) of a masked operand (kGemmShape*).
__device__ __forceinline__ double gemm_mask(double v, int r, int c, int shape) {
    if (shape == kGemmShapeFull) return v;
    const bool lower = shape == kGemmShapeLower || shape == kGemmShapeLowerUnit;
    const bool unit = shape == kGemmShapeLowerUnit || shape == kGemmShapeUpperUnit;
    if (r == c) return unit ? 1.0 : v;
    return (lower ? r > c : r < c) ? v : 0.0;
}

// One operand tile of 128 (m or n) x 16 (k) into registers, then LDS s[l][x]
// = op(M)(x0 + x, l) with x the m (A) or n (B) index.  KCONTIG: the stored
// matrix holds k down its columns (element (l, x) at M[l + x*ld]); otherwise
// element (x, l) at M[x + l*ld].
template <bool KCONTIG>
struct GemmLoader {
    const double* M;
    i64 ld;
    int ext, k, x0, shape;
    int tid;
    double g[8];
    __device__ __forceinline__ void load(int kc0) {
        if constexpr (KCONTIG) {
            const int x = tid >> 1, lh = (tid & 1) * 8;
            const bool xin = (x0 + x) < ext;
            const double* p = M + (i64)(x0 + x) * ld;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int l = kc0 + lh + q;
                const double v = (xin && l < k) ? p[l] : 0.0;
                g[q] = gemm_mask(v, l, x0 + x, shape);  // stored (row l, column x)
            }
        } else {
            const int l = kc0 + (tid >> 4), xb = (tid & 15) * 8;
            const bool lin = l < k;
            const double* p = M + (i64)l * ld + x0 + xb;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int x = x0 + xb + q;
                const double v = (lin && x < ext) ? p[q] : 0.0;
                g[q] = gemm_mask(v, x, l, shape);  // stored (row x, column l)
            }
        }
    }
    __device__ __forceinline__ void store(double (*s)[kGemmLdp]) const {
        if constexpr (KCONTIG) {
            const int x = tid >> 1, lh = (tid & 1) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) s[lh + q][x] = g[q];
        } else {
            const int l = tid >> 4, xb = (tid & 15) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) s[l][xb + q] = g[q];
        }
    }
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_f64_kernel(int m, int n, int k, double alpha, const double* A,
                                                          i64 lda, int ashape, const double* B, i64 ldb, int bshape,
                                                          double beta, double* C, i64 ldc, int tri,
                                                          const int* __restrict__ info) {
    __shared__ double sA[2][kGemmKC][kGemmLdp];  // sA[l][i] = op(A)(i0+i, l)
    __shared__ double sB[2][kGemmKC][kGemmLdp];  // sB[l][j] = op(B)(l, j0+j)
    if (info != nullptr && *info != 0) return;   // an earlier step failed
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

    // op(A) = A^T: A stored k x m (k down the columns); op(B) = B: B stored k x n
    GemmLoader<TA> la{A, lda, m, k, i0, ashape, tid, {}};
    GemmLoader<!TB> lb{B, ldb, n, k, j0, bshape, tid, {}};
    d4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (d4_t){0.0, 0.0, 0.0, 0.0};
    la.load(0);
    lb.load(0);
    la.store(sA[0]);
    lb.store(sB[0]);
    __syncthreads();
    const int nchunks = (int)ceil_div(k, kGemmKC);
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = (c + 1) < nchunks;
        if (more) {
            la.load((c + 1) * kGemmKC);
            lb.load((c + 1) * kGemmKC);
        }
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
        if (more) {
            la.store(sA[buf ^ 1]);
            lb.store(sB[buf ^ 1]);
        }
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

hipError_t launch_gemm_f64(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, i64 lda, int ashape,
                           const double* B, i64 ldb, int bshape, double beta, double* C, i64 ldc, int tri,
                           const int* info, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (tri != 0 && m != n) return hipErrorInvalidValue;
    const i64 T = ceil_div(m, kGemmTile);
    const dim3 grid = tri != 0 ? dim3((unsigned)(T * (T + 1) / 2))
                               : dim3((unsigned)T, (unsigned)ceil_div(n, kGemmTile));
    k = k < 0 ? 0 : k;
#define ELEMAMD_GEMM_CASE(TA_, TB_)                                                                              \
    hipLaunchKernelGGL((gemm_f64_kernel<TA_, TB_>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, ashape, B, ldb, \
                       bshape, beta, C, ldc, tri, info)
    if (ta && tb) ELEMAMD_GEMM_CASE(true, true);
    else if (ta) ELEMAMD_GEMM_CASE(true, false);
    else if (tb) ELEMAMD_GEMM_CASE(false, true);
    else ELEMAMD_GEMM_CASE(false, false);
#undef ELEMAMD_GEMM_CASE
    return hipGetLastError();
}

hipError_t launch_gemm_tn_f64(int m, int n, int k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                              double beta, double* C, i64 ldc, int tri, hipStream_t s) {
    return launch_gemm_f64(true, false, m, n, k, alpha, A, lda, kGemmShapeFull, B, ldb, kGemmShapeFull, beta, C, ldc,
                           tri, nullptr, s);
}

}  // namespace elemamd
