#include <algorithm>
// Bandwidth-bound helpers: fills, precision conversions, the symmetric
// (lower-stored) product used by the FP64 iterative-refinement residual
// r = b - A*x, and small reductions for the convergence tests.
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

// ------------------------------------------------------------------ fills
__global__ void fill_int_kernel(int* p, int value, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

hipError_t launch_fill_int(int* p, int value, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

// splitmix64 counter stream + Box-Muller: entry t (column-major order over a
// rows x cols matrix) is normal number offset+t of stream `seed`.  Pair p
// uses stream outputs 2p and 2p+1.
__device__ inline uint64_t sm64(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t x) { return ((double)(x >> 11) + 1.0) * 0x1.0p-53; }

__global__ void fill_normal_kernel(double* out, i64 rows, i64 cols, i64 ld, uint64_t seed, i64 offset) {
    const i64 total = rows * cols;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (i64)gridDim.x * blockDim.x) {
        const i64 idx = offset + t;
        const uint64_t p = (uint64_t)(idx >> 1);
        const double u1 = u01(sm64(seed, 2 * p));
        const double u2 = u01(sm64(seed, 2 * p + 1));
        const double r = sqrt(-2.0 * log(u1));
        const double a = 6.283185307179586476925286766559 * u2;
        const double v = (idx & 1) ? r * sin(a) : r * cos(a);
        const i64 i = t % rows, j = t / rows;
        out[i + j * ld] = v;
    }
}

hipError_t launch_fill_normal_f64(double* out, i64 rows, i64 cols, i64 ld, uint64_t seed, i64 offset,
                                  hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    const i64 total = rows * cols;
    const unsigned blocks = (unsigned)std::min<i64>(ceil_div(total, 256), 65536);
    hipLaunchKernelGGL(fill_normal_kernel, dim3(blocks), dim3(256), 0, s, out, rows, cols, ld, seed, offset);
    return hipGetLastError();
}

// ------------------------------------------------------------ transposes
// A := A^T for an n x n matrix in place: the workgroup of tile pair (I, J),
// I >= J, 32 x 32 tiles, loads both tiles (coalesced along columns), then
// writes each one's transpose over the other through LDS.  Every element is
// read and written once; the diagonal tiles transpose within themselves.
// The UPPER factorization runs as the LOWER one on A^T
// (cholesky::UVar3Unb on A is cholesky::LVar3Unb on A^T operation by
// operation for real data).
constexpr int kTT = 32;
template <typename T>
__global__ __launch_bounds__(256) void transpose_inplace_kernel(int n, T* A, i64 lda, int nt) {
    __shared__ T s1[kTT][kTT + 1], s2[kTT][kTT + 1];
    // tile pair from the linear index over the lower triangle of tiles
    const i64 e = blockIdx.x;
    int I = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while ((i64)I * (I + 1) / 2 > e) --I;
    while ((i64)(I + 1) * (I + 2) / 2 <= e) ++I;
    const int J = (int)(e - (i64)I * (I + 1) / 2);
    if (I >= nt) return;
    const int t = threadIdx.x, r = t & 31, c0 = t >> 5;  // 8 columns per pass
    const i64 i0 = (i64)I * kTT, j0 = (i64)J * kTT;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;
        if (i0 + r < n && j0 + c < n) s1[c][r] = A[(i0 + r) + (j0 + c) * lda];  // tile (I, J): s1[c][r] = A(i0+r, j0+c)
        if (j0 + r < n && i0 + c < n) s2[c][r] = A[(j0 + r) + (i0 + c) * lda];  // tile (J, I)
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;
        // new A(i0+r, j0+c) = old A(j0+c, i0+r) = s2[r][c]; new A(j0+r, i0+c) = old A(i0+c, j0+r) = s1[r][c]
        if (i0 + r < n && j0 + c < n) A[(i0 + r) + (j0 + c) * lda] = s2[r][c];
        if (I != J && j0 + r < n && i0 + c < n) A[(j0 + r) + (i0 + c) * lda] = s1[r][c];
    }
}

template <typename T>
hipError_t launch_transpose_inplace(int n, T* A, i64 lda, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    const int nt = (int)ceil_div(n, kTT);
    const i64 pairs = (i64)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(transpose_inplace_kernel<T>, dim3((unsigned)pairs), dim3(256), 0, s, n, A, lda, nt);
    return hipGetLastError();
}
template hipError_t launch_transpose_inplace<double>(int, double*, i64, hipStream_t);
template hipError_t launch_transpose_inplace<float>(int, float*, i64, hipStream_t);
hipError_t launch_transpose_inplace_f64(int n, double* A, i64 lda, hipStream_t s) {
    return launch_transpose_inplace<double>(n, A, lda, s);
}

// F (n x n, ldf) := the symmetric matrix stored in A's lower triangle
// (mirror = 1) or A's lower triangle with zeros above (mirror = 0); A's
// strictly upper triangle is never read.  One workgroup per lower tile pair.
__global__ __launch_bounds__(256) void from_lower_kernel(int n, const double* __restrict__ A, i64 lda,
                                                         double* __restrict__ F, i64 ldf, int nt, int mirror) {
    __shared__ double s1[kTT][kTT + 1];
    const i64 e = blockIdx.x;
    int I = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while ((i64)I * (I + 1) / 2 > e) --I;
    while ((i64)(I + 1) * (I + 2) / 2 <= e) ++I;
    const int J = (int)(e - (i64)I * (I + 1) / 2);
    if (I >= nt) return;
    const int t = threadIdx.x, r = t & 31, c0 = t >> 5;
    const i64 i0 = (i64)I * kTT, j0 = (i64)J * kTT;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;  // s1[c][r] = A(i0+r, j0+c), lower part only
        const bool in = i0 + r < n && j0 + c < n && (I != J || r >= c);
        s1[c][r] = in ? A[(i0 + r) + (j0 + c) * lda] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;
        if (i0 + r < n && j0 + c < n) {  // F(i0+r, j0+c)
            double v = s1[c][r];
            if (I == J && r < c) v = mirror ? s1[r][c] : 0.0;
            F[(i0 + r) + (j0 + c) * ldf] = v;
        }
        if (I != J && j0 + r < n && i0 + c < n)  // F(j0+r, i0+c) = A(i0+c, j0+r) = s1[r][c]
            F[(j0 + r) + (i0 + c) * ldf] = mirror ? s1[r][c] : 0.0;
    }
}

hipError_t launch_symmetrize_lower_f64(int n, const double* A, i64 lda, double* F, i64 ldf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int nt = (int)ceil_div(n, kTT);
    hipLaunchKernelGGL(from_lower_kernel, dim3((unsigned)((i64)nt * (nt + 1) / 2)), dim3(256), 0, s, n, A, lda, F,
                       ldf, nt, 1);
    return hipGetLastError();
}

hipError_t launch_copy_lower_zero_upper_f64(int n, const double* A, i64 lda, double* F, i64 ldf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int nt = (int)ceil_div(n, kTT);
    hipLaunchKernelGGL(from_lower_kernel, dim3((unsigned)((i64)nt * (nt + 1) / 2)), dim3(256), 0, s, n, A, lda, F,
                       ldf, nt, 0);
    return hipGetLastError();
}

// B (n x m) := A^T, A m x n.
__global__ __launch_bounds__(256) void transpose_kernel(int m, int n, const double* __restrict__ A, i64 lda,
                                                        double* __restrict__ B, i64 ldb) {
    __shared__ double sT[kTT][kTT + 1];
    const i64 i0 = (i64)blockIdx.x * kTT, j0 = (i64)blockIdx.y * kTT;
    const int t = threadIdx.x, r = t & 31, c0 = t >> 5;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;
        if (i0 + r < m && j0 + c < n) sT[c][r] = A[(i0 + r) + (j0 + c) * lda];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + 8 * q;
        // B(j0 + r, i0 + c) = A(i0 + c, j0 + r) = sT[r][c]
        if (j0 + r < n && i0 + c < m) B[(j0 + r) + (i0 + c) * ldb] = sT[r][c];
    }
}

hipError_t launch_transpose_f64(int m, int n, const double* A, i64 lda, double* B, i64 ldb, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)ceil_div(m, kTT), (unsigned)ceil_div(n, kTT)), dim3(256), 0,
                       s, m, n, A, lda, B, ldb);
    return hipGetLastError();
}

// ------------------------------------------------------- elementwise 2-D
__global__ void scale_f64_kernel(int m, int n, double alpha, double* A, i64 lda) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) A[i + j * lda] = (alpha == 0.0) ? 0.0 : A[i + j * lda] * alpha;  // alpha = 0 clears (NaN too)
}
hipError_t launch_scale_f64(int m, int n, double alpha, double* A, i64 lda, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_f64_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s, m, n, alpha,
                       A, lda);
    return hipGetLastError();
}

// S(i, j) = float(A(i, j)) for i >= j and j0 <= j < j1.  A workgroup covers
// 256 rows x 8 columns (one row per thread, its 8 loads in flight together);
// workgroups wholly above the diagonal return at once.
__global__ __launch_bounds__(256) void copy_lower_f64_to_f32_kernel(int n, int j0, int j1,
                                                                    const double* __restrict__ A, i64 lda,
                                                                    float* __restrict__ S, i64 lds) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int jb = j0 + blockIdx.y * 8;
    if ((i64)blockIdx.x * 256 + 255 < jb) return;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int j = jb + q;
        v[q] = (i < n && j < j1 && i >= j) ? A[i + (i64)j * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int j = jb + q;
        if (i < n && j < j1 && i >= j) S[i + (i64)j * lds] = (float)v[q];
    }
}
hipError_t launch_copy_lower_f64_to_f32_cols(int n, int j0, int j1, const double* A, i64 lda, float* S, i64 lds,
                                            hipStream_t s) {
    j1 = j1 < n ? j1 : n;
    if (n <= 0 || j0 >= j1) return hipSuccess;
    hipLaunchKernelGGL(copy_lower_f64_to_f32_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)ceil_div(j1 - j0, 8)),
                       dim3(256), 0, s, n, j0, j1, A, lda, S, lds);
    return hipGetLastError();
}
hipError_t launch_copy_lower_f64_to_f32(int n, const double* A, i64 lda, float* S, i64 lds, hipStream_t s) {
    return launch_copy_lower_f64_to_f32_cols(n, 0, n, A, lda, S, lds, s);
}

__global__ void copy_f64_to_f32_scaled_kernel(int m, const double* X, i64 ldx, const double* cs, float* Y, i64 ldy) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) {
        const double sc = cs ? cs[j] : 1.0;
        Y[i + j * ldy] = (float)(sc > 0.0 ? X[i + j * ldx] / sc : 0.0);
    }
}
hipError_t launch_copy_f64_to_f32_scaled(int m, int n, const double* X, i64 ldx, const double* colscale, float* Y,
                                         i64 ldy, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_f64_to_f32_scaled_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s,
                       m, X, ldx, colscale, Y, ldy);
    return hipGetLastError();
}

__global__ void axpy_f32_to_f64_scaled_kernel(int m, const float* D, i64 ldd, const double* cs, double* X, i64 ldx) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) {
        const double sc = cs ? cs[j] : 1.0;
        X[i + j * ldx] += sc * (double)D[i + j * ldd];
    }
}
hipError_t launch_axpy_f32_to_f64_scaled(int m, int n, const float* D, i64 ldd, const double* colscale, double* X,
                                         i64 ldx, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpy_f32_to_f64_scaled_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s,
                       m, D, ldd, colscale, X, ldx);
    return hipGetLastError();
}

__global__ void copy_f64_kernel(int m, const double* X, i64 ldx, double* Y, i64 ldy) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) Y[i + j * ldy] = X[i + j * ldx];
}
hipError_t launch_copy_f64(int m, int n, const double* X, i64 ldx, double* Y, i64 ldy, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_f64_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s, m, X, ldx, Y,
                       ldy);
    return hipGetLastError();
}

// ------------------------------------------- symmetric product (lower read)
// out[p, c] = beta*B[p, c] + alpha * sum_q Asym[p, q] * X[q, c]
// with Asym the symmetric matrix whose lower triangle is stored in A.
//
// A workgroup owns 64 output rows and up to 32 right-hand sides and walks
// the 64x64 tiles of its block row of the symmetric matrix (tiles right of
// the diagonal are read as the transpose of the stored lower tile, the
// diagonal tile is mirrored).  Tiles and the matching 64 x 32 slab of X are
// staged through double-buffered LDS; the product runs on
// v_mfma_f64_16x16x4_f64 with MFMA rows = right-hand sides and MFMA columns =
// output rows, so each 16 lanes store 128 contiguous bytes of a column.
// One LDS buffer (the next tile waits in registers), so three workgroups
// share a CU and hide each other's HBM latency.
// At k = 32 the FP64 work (2 n^2 k flops) and the bytes read (8 n^2) are
// balanced on MI355X, so the kernel sits near both roofs.
// ABS = true reads |A| and X = ones: out[p] = sum_q |Asym[p, q]| (row sums
// for the infinity norm).
constexpr int kSymT = 64;
constexpr int kSymCols = 32;
constexpr int kSymLd = kSymT + 1;

template <bool ABS>
__global__ __launch_bounds__(256, 3) void symm_lower_kernel(int n, int k, double alpha, const double* __restrict__ A,
                                                         i64 lda, const double* __restrict__ X, i64 ldx, double beta,
                                                         const double* __restrict__ B, i64 ldb, double* __restrict__ R,
                                                         i64 ldr, int nsplit, double* __restrict__ part) {
    __shared__ double sM[kSymT][kSymLd];     // sM[q][p] = Asym(p0+p, q0+q)
    // sX[q][c] = X(q0+q, c_base+c); the ABS form multiplies by a column of
    // ones formed in registers and keeps only the tile in LDS (33 KB), which
    // leaves room for the factorization's chain kernels beside it
    __shared__ double sX[ABS ? 1 : kSymT][kSymCols + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int I = blockIdx.x;
    const i64 p0 = (i64)I * kSymT;
    const int c_base = blockIdx.y * kSymCols;
    const int kcl = ABS ? 1 : min(kSymCols, k - c_base);
    const int nT = (int)ceil_div(n, kSymT);
    // this workgroup's share of the tiles of the block row (split over blockIdx.z)
    const int Jb = (int)((i64)nT * blockIdx.z / nsplit), Je = (int)((i64)nT * (blockIdx.z + 1) / nsplit);

    double gm[16], gx[8];
    // stage tile J into registers
    auto gload = [&](int J) {
        const i64 q0 = (i64)J * kSymT;
        if (J < I) {  // stored lower tile: Asym(p, q) = A(p, q); lanes along p
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int qq = w + 4 * s;
                const i64 p = p0 + lane;
                double v = (p < n) ? A[p + (q0 + qq) * lda] : 0.0;
                gm[s] = ABS ? fabs(v) : v;
            }
        } else if (J > I) {  // Asym(p, q) = A(q, p); lanes along q
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int pp = w + 4 * s;
                const i64 q = q0 + lane, p = p0 + pp;
                double v = (q < n && p < n) ? A[q + p * lda] : 0.0;
                gm[s] = ABS ? fabs(v) : v;
            }
        } else {  // diagonal tile, lower part (mirrored at store time)
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int qq = w + 4 * s;
                const i64 p = p0 + lane, q = q0 + qq;
                double v = (p < n && lane >= qq) ? A[p + q * lda] : 0.0;
                gm[s] = ABS ? fabs(v) : v;
            }
        }
        if constexpr (!ABS) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int c = w + 4 * s;
                const i64 q = q0 + lane;
                gx[s] = (q < n && c < kcl) ? X[q + (i64)(c_base + c) * ldx] : 0.0;
            }
        }
    };
    auto sstore = [&](int J) {
        if (J < I) {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[w + 4 * s][lane] = gm[s];
        } else if (J > I) {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[lane][w + 4 * s] = gm[s];
        } else {
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int qq = w + 4 * s;
                if (lane >= qq) {
                    sM[qq][lane] = gm[s];
                    sM[lane][qq] = gm[s];
                }
            }
        }
        if constexpr (!ABS) {
#pragma unroll
            for (int s = 0; s < 8; ++s) sX[lane][w + 4 * s] = gx[s];
        }
    };

    d4_t acc[2];
    acc[0] = (d4_t){0.0, 0.0, 0.0, 0.0};
    acc[1] = (d4_t){0.0, 0.0, 0.0, 0.0};
    gload(Jb);
    sstore(Jb);
    __syncthreads();
    for (int J = Jb; J < Je; ++J) {
        // branch-free body: the last iteration restages its own tile, which
        // is never read again
        gload((J + 1) < Je ? J + 1 : J);
        // wave w: output rows p = 16w .. 16w+15, both 16-column blocks of c
#pragma unroll
        for (int ks = 0; ks < kSymT / 4; ++ks) {
            const int q = 4 * ks + (lane >> 4);
            const double bfrag = sM[q][16 * w + (lane & 15)];
            if constexpr (ABS) {
                // X = ones in column 0 (rows past n meet zero tile entries)
                const double a0 = (lane & 15) == 0 ? 1.0 : 0.0;
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bfrag, acc[0], 0, 0, 0);
            } else {
                const double a0 = sX[q][(lane & 15)];
                const double a1 = sX[q][16 + (lane & 15)];
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bfrag, acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bfrag, acc[1], 0, 0, 0);
            }
        }
        __syncthreads();  // every wave is done with this tile
        sstore((J + 1) < Je ? J + 1 : J);
        __syncthreads();
    }
    // dependent zero-operand MFMAs right before the accumulators are read, so
    // the reads follow an MFMA in the same basic block (the compiler's
    // wait-state padding across the loop back edge is short for FP64 MFMA)
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc[1], 0, 0, 0);
    const i64 p = p0 + 16 * w + (lane & 15);
    if (p < n) {
        // one split: the result; several: this split's partial sum, combined
        // by symm_combine_kernel in a fixed order
        double* P = nsplit > 1 ? part + (i64)blockIdx.z * n * (ABS ? 1 : k) : nullptr;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 16 * cb + (lane >> 4) + 4 * v;  // f64 MFMA: row = lane/16 + 4*v
                if (c < kcl) {
                    const i64 cc = c_base + c;
                    if (nsplit > 1) {
                        P[p + cc * n] = acc[cb][v];
                    } else {
                        const double bv = (beta == 0.0) ? 0.0 : beta * B[p + cc * ldb];
                        R[p + cc * ldr] = bv + alpha * acc[cb][v];
                    }
                }
            }
    }
}

// R = beta*B + alpha * sum_s part[s], s in increasing order
__global__ void symm_combine_kernel(int n, int k, int nsplit, const double* __restrict__ part, double alpha,
                                    double beta, const double* __restrict__ B, i64 ldb, double* __restrict__ R,
                                    i64 ldr) {
    const int c = blockIdx.y;
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int s = 0; s < nsplit; ++s) acc += part[(i64)s * n * k + i + (i64)c * n];
    const double bv = (beta == 0.0) ? 0.0 : beta * B[i + (i64)c * ldb];
    R[i + (i64)c * ldr] = bv + alpha * acc;
}

// Splits of the tile sweep: enough workgroups to keep several tiles in
// flight per CU (the kernel is bound by HBM latency otherwise).
int symm_splits(int n, int k) {
    const i64 rows = ceil_div(n, kSymT) * ceil_div(k, kSymCols);
    i64 s = ceil_div(2048, rows);
    s = std::min<i64>(std::max<i64>(s, 1), 8);
    s = std::min<i64>(s, ceil_div(n, kSymT));
    return (int)s;
}
size_t symm_part_count(int n, int k) {
    const int s = symm_splits(n, k);
    return s > 1 ? (size_t)s * n * k : 0;
}

template <bool ABS>
static hipError_t symm_launch(int n, int k, double alpha, const double* A, i64 lda, const double* X, i64 ldx,
                              double beta, const double* B, i64 ldb, double* R, i64 ldr, double* part, hipStream_t s) {
    if (n <= 0 || k <= 0) return hipSuccess;
    const int ns = part != nullptr ? symm_splits(n, k) : 1;
    dim3 grid((unsigned)ceil_div(n, kSymT), ABS ? 1u : (unsigned)ceil_div(k, kSymCols), (unsigned)ns);
    hipLaunchKernelGGL(symm_lower_kernel<ABS>, grid, dim3(256), 0, s, n, k, alpha, A, lda, X, ldx, beta, B, ldb, R,
                       ldr, ns, part);
    if (ns > 1)
        hipLaunchKernelGGL(symm_combine_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)k), dim3(256), 0, s, n, k,
                           ns, part, alpha, beta, B, ldb, R, ldr);
    return hipGetLastError();
}

hipError_t launch_symm_lower_residual(int n, int k, const double* A, i64 lda, const double* X, i64 ldx,
                                      const double* B, i64 ldb, double* R, i64 ldr, double* part, hipStream_t s) {
    return symm_launch<false>(n, k, -1.0, A, lda, X, ldx, 1.0, B, ldb, R, ldr, part, s);
}

hipError_t launch_symm_lower(int n, int k, const double* A, i64 lda, const double* X, i64 ldx, double* Y, i64 ldy,
                             double* part, hipStream_t s) {
    return symm_launch<false>(n, k, 1.0, A, lda, X, ldx, 0.0, Y, ldy, Y, ldy, part, s);
}

hipError_t launch_sym_lower_abs_rowsum(int n, const double* A, i64 lda, double* rowsum, double* part,
                                       hipStream_t s) {
    return symm_launch<true>(n, 1, 1.0, A, lda, (const double*)nullptr, (i64)1, 0.0, (const double*)nullptr,
                             (i64)1, rowsum, (i64)n, part, s);
}

// ------------------------------------------------------------ reductions
__device__ inline double block_max_256(double v) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Reductions over one column per workgroup of 1024 threads: each thread
// keeps 4 independent loads in flight, then the waves reduce by shuffles.
template <typename Op>
__device__ inline double block_reduce_1024(double v, Op op) {
    __shared__ double part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
    const int t = threadIdx.x;
    if ((t & 63) == 0) part[t >> 6] = v;
    __syncthreads();
    if (t < 64) {
        v = t < 16 ? part[t] : part[0];
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
    }
    return v;  // valid in thread 0
}

__global__ __launch_bounds__(1024) void col_absmax_kernel(int m, const double* X, i64 ldx, double* colmax) {
    const int c = blockIdx.x;
    const double* x = X + (i64)c * ldx;
    auto mx = [](double a, double b) { return (a > b || a != a) ? a : b; };  // propagates NaN
    double v = 0.0;
    for (i64 i0 = threadIdx.x; i0 < m; i0 += 4 * 1024) {
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = i0 + 1024 * u < m ? fabs(x[i0 + 1024 * u]) : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) v = mx(a[u], v);
    }
    v = block_reduce_1024(v, mx);
    if (threadIdx.x == 0) colmax[c] = v;
}
hipError_t launch_col_absmax(int m, int k, const double* X, i64 ldx, double* colmax, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(col_absmax_kernel, dim3((unsigned)k), dim3(1024), 0, s, m, X, ldx, colmax);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void vec_max_kernel(int n, const double* v, double* out) {
    double x = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) x = fmax(x, v[i]);
    x = block_max_256(x);
    if (threadIdx.x == 0) out[0] = x;
}
hipError_t launch_vec_max(int n, const double* v, double* out, hipStream_t s) {
    hipLaunchKernelGGL(vec_max_kernel, dim3(1), dim3(256), 0, s, n, v, out);
    return hipGetLastError();
}


// ------------------------------------------------- batched column kernels
// Right-hand sides are processed as k independent columns; vector i of a
// set lives at V + i*vstride (an n x k block with leading dim ldv).

// out[i + c*nv] = sum_r V_i(r, c) * W(r, c)   (one workgroup per (c, i))
__global__ __launch_bounds__(1024) void batched_dots_kernel(int n, int nv, const double* V, i64 ldv, i64 vstride,
                                                            const double* W, i64 ldw, double* out) {
    const int c = blockIdx.x, i = blockIdx.y;
    const double* v = V + (i64)i * vstride + (i64)c * ldv;
    const double* w = W + (i64)c * ldw;
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (i64 r0 = threadIdx.x; r0 < n; r0 += 4 * 1024) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = r0 + 1024 * u < n;
            a[u] = in ? v[r0 + 1024 * u] : 0.0;
            b[u] = in ? w[r0 + 1024 * u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) sum[u] = fma(a[u], b[u], sum[u]);
    }
    double x = (sum[0] + sum[1]) + (sum[2] + sum[3]);
    x = block_reduce_1024(x, [](double p, double q) { return p + q; });
    if (threadIdx.x == 0) out[i + (i64)c * nv] = x;
}
hipError_t launch_batched_dots(int n, int k, int nv, const double* V, i64 ldv, i64 vstride, const double* W, i64 ldw,
                               double* out, hipStream_t s) {
    if (k <= 0 || nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(batched_dots_kernel, dim3((unsigned)k, (unsigned)nv), dim3(1024), 0, s, n, nv, V, ldv, vstride,
                       W, ldw, out);
    return hipGetLastError();
}

// W(r, c) = beta*W(r, c) + alpha * sum_i h[i + c*nv] * V_i(r, c)
__global__ __launch_bounds__(256) void batched_gemv_kernel(int n, int nv, double alpha, const double* V, i64 ldv,
                                                           i64 vstride, const double* h, double beta, double* W,
                                                           i64 ldw) {
    const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s += h[i + (i64)c * nv] * V[(i64)i * vstride + r + (i64)c * ldv];
    double* w = W + r + (i64)c * ldw;
    *w = (beta == 0.0 ? 0.0 : beta * (*w)) + alpha * s;
}
hipError_t launch_batched_gemv(int n, int k, int nv, double alpha, const double* V, i64 ldv, i64 vstride,
                               const double* h, double beta, double* W, i64 ldw, hipStream_t s) {
    if (n <= 0 || k <= 0) return hipSuccess;
    hipLaunchKernelGGL(batched_gemv_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)k), dim3(256), 0, s, n, nv,
                       alpha, V, ldv, vstride, h, beta, W, ldw);
    return hipGetLastError();
}

// X(:, c) = f(c) * Y(:, c) with f(c) = s[c] (invert = 0) or 1/s[c] (invert = 1; 0 if s[c] == 0)
__global__ void col_scale_kernel(int n, const double* Y, i64 ldy, const double* sc, int invert, double* X, i64 ldx) {
    const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    double f = sc[c];
    if (invert) f = (f != 0.0) ? 1.0 / f : 0.0;
    X[r + (i64)c * ldx] = f * Y[r + (i64)c * ldy];
}
hipError_t launch_col_scale(int n, int k, const double* Y, i64 ldy, const double* sc, int invert, double* X, i64 ldx,
                            hipStream_t s) {
    if (n <= 0 || k <= 0) return hipSuccess;
    hipLaunchKernelGGL(col_scale_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)k), dim3(256), 0, s, n, Y, ldy, sc,
                       invert, X, ldx);
    return hipGetLastError();
}

// S(i, i) += sigma * |A(i, i)| on the float copy (diagonal shift of the low-precision factor)
__global__ void shift_diag_kernel(int n, const double* A, i64 lda, float* S, i64 lds, double sigma) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) S[i + i * lds] = (float)(A[i + i * lda] + sigma * fabs(A[i + i * lda]));
}
hipError_t launch_shift_diag(int n, const double* A, i64 lda, float* S, i64 lds, double sigma, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(shift_diag_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, A, lda, S, lds, sigma);
    return hipGetLastError();
}

// Y(:, c) = colscale[c] * double(D(:, c))
__global__ void copy_f32_to_f64_scaled_kernel(int m, const float* D, i64 ldd, const double* cs, double* Y, i64 ldy) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) Y[i + j * ldy] = (cs ? cs[j] : 1.0) * (double)D[i + j * ldd];
}
hipError_t launch_copy_f32_to_f64_scaled(int m, int n, const float* D, i64 ldd, const double* colscale, double* Y,
                                         i64 ldy, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_f32_to_f64_scaled_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s,
                       m, D, ldd, colscale, Y, ldy);
    return hipGetLastError();
}

// Y += alpha * X
__global__ void axpy_f64_kernel(int m, double alpha, const double* X, i64 ldx, double* Y, i64 ldy) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) Y[i + j * ldy] += alpha * X[i + j * ldx];
}
hipError_t launch_axpy_f64(int m, int n, double alpha, const double* X, i64 ldx, double* Y, i64 ldy, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpy_f64_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s, m, alpha, X,
                       ldx, Y, ldy);
    return hipGetLastError();
}

// ------------------------------------------- preconditioned CG refinement
// Per column c (all n x k blocks with ld n, scalars length k):
//   alpha = rz/pq (0 for an inactive column or a non-positive pq),
//   X += alpha P,  R -= alpha Q.
__global__ void cg_xr_kernel(int m, const double* __restrict__ rz, const double* __restrict__ pq,
                             const int* __restrict__ active, const double* __restrict__ P,
                             const double* __restrict__ Q, double* __restrict__ X, double* __restrict__ R) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m) return;
    const double d = pq[j], a = (active[j] && d > 0.0 && d < INFINITY) ? rz[j] / d : 0.0;
    const i64 e = i + (i64)j * m;
    X[e] += a * P[e];
    R[e] -= a * Q[e];
}
hipError_t launch_cg_xr(int m, int n, const double* rz, const double* pq, const int* active, const double* P,
                        const double* Q, double* X, double* R, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cg_xr_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s, m, rz, pq, active,
                       P, Q, X, R);
    return hipGetLastError();
}

// P := Z + beta P with the flexible (Polak-Ribiere) coefficient
// beta = -z.q / p.q (Q = A P of the step just taken); beta = 0 restarts.
__global__ void cg_p_kernel(int m, const double* __restrict__ zq, const double* __restrict__ pq,
                            const double* __restrict__ Z, double* __restrict__ P) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m) return;
    const double d = pq[j], b = (d > 0.0 && d < INFINITY) ? -zq[j] / d : 0.0;
    const i64 e = i + (i64)j * m;
    P[e] = Z[e] + b * P[e];
}
hipError_t launch_cg_p(int m, int n, const double* zq, const double* pq, const double* Z, double* P, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cg_p_kernel, dim3((unsigned)ceil_div(m, 256), (unsigned)n), dim3(256), 0, s, m, zq, pq, Z, P);
    return hipGetLastError();
}

}  // namespace elemamd
