// Panel step of the low-precision (float-storage) factorization: the
// diagonal block L11 = chol(A11) of cholesky::LVar3
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:83-123, LVar3Unb at :19-49),
// its inverse, and the panel solve L21 := A21 inv(L11)^T done as a product
// with that inverse.  Three kernels, one workgroup each for the first two:
//
// potrf_f32_kernel -- the block in LDS as float, right-looking in 32-column
//   steps: wave 0 factors the 32 x 32 diagonal piece in FP64; waves 1-2 solve
//   the rows below it by substitution (FP64, one row per lane); all waves
//   apply the rank-32 update to the rest of the block (FP32 MFMA).  A block
//   narrower than 128 is padded with the identity.
// trtri_f32_kernel -- inv(L11) of the stored float factor: each wave inverts
//   one 32 x 32 diagonal piece in FP64 (one column per lane), then the
//   off-diagonal blocks follow by distance i - j,
//       X_ij = -X_ii sum_{j <= l < i} L_il X_lj   (FP32 MFMA).
// panel_inv_f32_kernel -- B := B inv(L11)^T per 64-row slab (FP32 MFMA), so
//   the panel no longer runs a 128-step substitution per row; the refinement
//   absorbs the larger rounding error of the inverse (this factor only
//   preconditions).
#include "common.hpp"
#include "kernels.hpp"
#include "solve_dev.hpp"

namespace elemamd {

constexpr int kPiLd = 132;  // LDS row stride (floats): MFMA operand reads are conflict-free

// One wave: factor the 32 x 32 block of sA at (c0, c0) in FP64 into sD
// (transposed) and back into sA as float, with the reciprocal diagonal in sRd.  Returns 0,
// or the 1-based failing column.
// Lane (h, i) = (lane / 32, lane % 32) holds columns 16h .. 16h+15 of row i.
// Column j travels through LDS (the wave's own writes and reads, in order):
// its owners publish the entries below the pivot (zeros above it) and the
// pivot, every lane reads the pivot and the 16 entries of its half, and the
// rank-1 update runs unmasked because the published zeros leave finished
// entries unchanged.  The pivot's reciprocal square root comes from the
// hardware estimate refined by two Newton steps (error far below the float
// the factor is stored in) and is kept, so nothing divides.
__device__ inline int factor32(float (*sA)[kPiLd], double (*sD)[34], double* sCol, double* sRd, int c0, int lane) {
    // an opaque copy of the lane index keeps the compiler from hoisting the
    // per-lane masks of all four calls out of the caller's loop (it ran out
    // of registers doing so)
    int lv = lane;
    asm volatile("" : "+v"(lv));
    const int i = lv & 31, hh = lv >> 5;
    double d[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) d[q] = (double)sA[c0 + i][c0 + 16 * hh + q];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const int jh = j >> 4, jq = j & 15;
        int jv = j;  // opaque per column: the lane masks are made here, not all up front
        asm volatile("" : "+s"(jv));
        const bool own = hh == (jv >> 4), diag = i == jv;
        if (own) {
            sCol[i] = (i > jv) ? d[jq] : 0.0;
            if (diag) sCol[32] = d[jq];
        }
        const double piv = sCol[32];
        if (!(piv > 0.0)) return j + 1;
        const double hp = 0.5 * piv;
        double y = __builtin_amdgcn_rsq(piv);
        y = fma(y, fma(-hp * y, y, 0.5), y);
        y = fma(y, fma(-hp * y, y, 0.5), y);
        const double lij = sCol[i] * y;
        if (own) d[jq] = diag ? piv * y : lij;
#pragma unroll
        for (int q = 0; q < 16; ++q) d[q] = fma(-lij, sCol[16 * hh + q] * y, d[q]);
        if (lv == 0) sRd[j] = y;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = 16 * hh + q;
        sD[k][i] = (k <= i) ? d[q] : 0.0;  // transposed: column k of L is a row of sD
        sA[c0 + i][c0 + k] = (k <= i) ? (float)d[q] : 0.0f;
    }
    return 0;
}

// Waves 1-2: the rows r in [c0 + 32, 128) of the 32 columns below the piece
// factored by factor32, one row per lane, by substitution against L (FP64,
// sD transposed): x_k = a_k / L_kk, then a_c -= x_k L_ck for c > k.
__device__ inline void subpanel32(float (*sA)[kPiLd], const double (*sD)[34], const double* sRd, int c0, int row) {
    if (row >= 128) return;
    double a[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) a[c] = (double)sA[row][c0 + c];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        a[k] *= sRd[k];
#pragma unroll
        for (int c = k + 1; c < 32; ++c) a[c] = fma(-a[k], sD[k][c], a[c]);
        asm volatile("" ::: "memory");  // one column of L in flight at a time (registers)
    }
#pragma unroll
    for (int c = 0; c < 32; ++c) sA[row][c0 + c] = (float)a[c];
}

// One wave: the inverse of the 32 x 32 lower-triangular piece of sL at
// (c0, c0), column c = lane (lanes 0..31), right-looking in FP64:
// x_k *= 1/L_kk, then x_r -= L_rk x_k for r > k; into sX at (c0, c0)
// (zero above the diagonal).
__device__ inline void inverse32(float (*sX)[kPiLd], const float (*sL)[kPiLd], int c0, int lane) {
    const int c = lane & 31;
    double x[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) x[r] = (r == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        const double dk = (double)sL[c0 + k][c0 + k];
        double rk = __builtin_amdgcn_rcp(dk);  // hardware estimate, two Newton steps
        rk = fma(rk, fma(-dk, rk, 1.0), rk);
        rk = fma(rk, fma(-dk, rk, 1.0), rk);
        x[k] *= rk;
#pragma unroll
        for (int r = k + 1; r < 32; ++r) x[r] = fma(-(double)sL[c0 + r][c0 + k], x[k], x[r]);
        asm volatile("" ::: "memory");  // one column of L in flight at a time (registers)
    }
    if (lane < 32) {
#pragma unroll
        for (int r = 0; r < 32; ++r) sX[c0 + r][c0 + c] = (float)x[r];
    }
}

// acc (16 x 16) += P (16 x 32) Q^T: rows pr.. of P from column pc, rows qr..
// of Q from column qc, both row-major in LDS; all 16 operand loads are in
// flight together, then the 8 MFMA steps.
template <int LP, int LQ>
__device__ inline void tile_ab_t(f4_t& acc, const float (*P)[LP], int pr, int pc, const float (*Q)[LQ], int qr, int qc,
                                 int m, int g) {
    float pa[8], qb[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        pa[ks] = P[pr + m][pc + 4 * ks + g];
        qb[ks] = Q[qr + m][qc + 4 * ks + g];
    }
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) mfma_acc(acc, pa[ks], qb[ks]);
}

// One b x b (b <= 128) diagonal block at A (float storage): L11 overwrites
// its lower triangle.  A failing pivot records col_offset + column in *info.
__global__ __launch_bounds__(256) void potrf_f32_kernel(int b, float* __restrict__ A, i64 lda,
                                                        int* __restrict__ info, int col_offset) {
    __shared__ float sA[128][kPiLd];
    __shared__ double sD[32][34];  // the FP64 factor of the current piece, transposed
    __shared__ double sCol[33], sRd[32];
    __shared__ int sfail;
    if (info != nullptr && *info != 0 && *info <= col_offset) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m = lane & 15, g = lane >> 4;
    {
        // element e = t + 256 q: row e % 128, column e / 128; all 64 loads from
        // clamped (in-bounds) addresses in flight together, then selected into LDS
        float v[64];
        const int r = t & 127;
        const i64 rr = r < b ? r : b - 1;
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int c = (t >> 7) + 2 * q;
            v[q] = A[rr + (i64)(c < b ? c : b - 1) * lda];
        }
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int c = (t >> 7) + 2 * q;
            sA[r][c] = (r < b && c < b) ? (c <= r ? v[q] : 0.0f) : (r == c ? 1.0f : 0.0f);
        }
    }
    if (t == 0) sfail = 0;
    __syncthreads();
    for (int p = 0; p < 4; ++p) {
        const int c0 = 32 * p;
        if (w == 0) {
            const int f = factor32(sA, sD, sCol, sRd, c0, lane);
            if (f != 0 && lane == 0) sfail = c0 + f;
        }
        __syncthreads();
        if (sfail != 0) break;
        if (w == 1 || w == 2) subpanel32(sA, sD, sRd, c0, c0 + 32 + 64 * (w - 1) + lane);
        __syncthreads();
        // rank-32 update of the rest of the block (lower 16 x 16 tiles)
        const int nrt = (96 - c0) / 16;
        const int ntl = nrt * (nrt + 1) / 2;
        for (int tt = w; tt < ntl; tt += 4) {
            int ti = 0;
            while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
            const int tj = tt - ti * (ti + 1) / 2;
            const int r0 = c0 + 32 + 16 * ti, q0 = c0 + 32 + 16 * tj;
            f4_t acc = {};
            tile_ab_t(acc, sA, r0, c0, sA, q0, c0, m, g);
            mfma_retire();
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + 4 * g + v, q = q0 + m;
                if (q <= r) sA[r][q] -= acc[v];
            }
        }
        __syncthreads();
    }
    const int fail = sfail;
#pragma unroll 4
    for (int idx = t; idx < 128 * 128; idx += 256) {
        const int r = idx & 127, c = idx >> 7;
        if (r < b && c <= r) A[r + (i64)c * lda] = sA[r][c];
    }
    if (fail != 0 && t == 0 && info != nullptr) *info = col_offset + fail;
}

// Linv (128 x 128, column-major) := inv(L11) of the b x b float factor at L
// (the identity beyond b); skipped, like the factor, after a failure.
__global__ __launch_bounds__(256) void trtri_f32_kernel(int b, const float* __restrict__ L, i64 ldl,
                                                        float* __restrict__ Linv, const int* __restrict__ info,
                                                        int limit) {
    __shared__ float sL[128][kPiLd];
    __shared__ float sX[128][kPiLd];
    __shared__ float sT[3][32][36];  // S_ij^T of the assembly, one per block of a distance
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m = lane & 15, g = lane >> 4;
    {
        float v[64];
        const int r = t & 127;
        const i64 rr = r < b ? r : b - 1;
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int c = (t >> 7) + 2 * q;
            v[q] = L[rr + (i64)(c < b ? c : b - 1) * ldl];
        }
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int c = (t >> 7) + 2 * q;
            sL[r][c] = (r < b && c < b) ? (c <= r ? v[q] : 0.0f) : (r == c ? 1.0f : 0.0f);
            sX[r][c] = 0.0f;
        }
    }
    __syncthreads();
    inverse32(sX, sL, 32 * w, lane);  // wave w: diagonal piece w
    __syncthreads();
    // assembly by distance: S_ij = sum_l L_il X_lj, then X_ij = -X_ii S_ij.
    // sT holds S_ij transposed (row = column of S) so both products read
    // their second operand row-wise.
    for (int dist = 1; dist < 4; ++dist) {
        const int nblk = 4 - dist;
        for (int task = w; task < 4 * nblk; task += 4) {
            const int jb = task >> 2, sub = task & 3, ib = jb + dist;
            const int a = sub >> 1, bq = sub & 1;
            f4_t acc = {};
            for (int l = jb; l < ib; ++l) {
                float pa[8], qb[8];
#pragma unroll
                for (int ks = 0; ks < 8; ++ks) {
                    pa[ks] = sL[32 * ib + 16 * a + m][32 * l + 4 * ks + g];
                    qb[ks] = sX[32 * l + 4 * ks + g][32 * jb + 16 * bq + m];
                }
#pragma unroll
                for (int ks = 0; ks < 8; ++ks) mfma_acc(acc, pa[ks], qb[ks]);
            }
            mfma_retire();
#pragma unroll
            for (int v = 0; v < 4; ++v) sT[jb][16 * bq + m][16 * a + 4 * g + v] = acc[v];
        }
        __syncthreads();
        for (int task = w; task < 4 * nblk; task += 4) {
            const int jb = task >> 2, sub = task & 3, ib = jb + dist;
            const int a = sub >> 1, bq = sub & 1;
            f4_t acc = {};
            tile_ab_t(acc, sX, 32 * ib + 16 * a, 32 * ib, sT[jb], 16 * bq, 0, m, g);
            mfma_retire();
#pragma unroll
            for (int v = 0; v < 4; ++v) sX[32 * ib + 16 * a + 4 * g + v][32 * jb + 16 * bq + m] = -acc[v];
        }
        __syncthreads();
    }
#pragma unroll 4
    for (int idx = t; idx < 128 * 128; idx += 256) Linv[idx] = sX[idx & 127][idx >> 7];
}

// B (m x b, b <= 128) := B inv(L11)^T with Linv from potrf_inv_f32_kernel;
// one 64-row slab per workgroup, the result staged through LDS so the
// column-major stores are contiguous.
__global__ __launch_bounds__(256) void panel_inv_f32_kernel(int m, int b, const float* __restrict__ Linv,
                                                            float* __restrict__ B, i64 ldb, const int* __restrict__ info,
                                                            int limit, unsigned short* __restrict__ Hi,
                                                            unsigned short* __restrict__ Lo, i64 ldk) {
    __shared__ float sX[128][kPiLd];  // sX[c][k] = X(c, k)
    __shared__ float sB[64][kPiLd];
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m16 = lane & 15, g = lane >> 4;
    const i64 r0 = (i64)blockIdx.x * 64;
    {
        float v[16];
        for (int q0 = 0; q0 < 64; q0 += 16) {  // inv(L11): 64 elements per thread
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = Linv[t + 256 * (q0 + q)];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int idx = t + 256 * (q0 + q);
                sX[idx & 127][idx >> 7] = v[q];
            }
        }
        // the slab: row r, columns c_lo + 4q, from clamped addresses
        const int r = t & 63, c_lo = t >> 6;
        const i64 rr = r0 + r < m ? r0 + r : m - 1;
        for (int q0 = 0; q0 < 32; q0 += 16) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = c_lo + 4 * (q0 + q);
                v[q] = B[rr + (i64)(c < b ? c : b - 1) * ldb];
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = c_lo + 4 * (q0 + q);
                sB[r][c] = (r0 + r < m && c < b) ? v[q] : 0.0f;
            }
        }
    }
    __syncthreads();
    f4_t acc[8] = {};
#pragma unroll 8
    for (int ks = 0; ks < 32; ++ks) {
        const float a = sB[16 * w + m16][4 * ks + g];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) mfma_acc(acc[nt], a, sX[16 * nt + m16][4 * ks + g]);
    }
    mfma_retire();
    __syncthreads();  // every wave is done reading sB
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) sB[16 * w + 4 * g + v][16 * nt + m16] = acc[nt][v];
    __syncthreads();
    for (int idx = t; idx < 64 * 128; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        if (r0 + r < m && c < b) B[r0 + r + (i64)c * ldb] = sB[r][c];
    }
    if (Hi != nullptr) {  // bf16 hi/lo planes of the result: thread = (row, 8 columns), 16-byte stores
        for (int e = t; e < 64 * 16; e += 256) {
            const int r = e >> 4, c = 8 * (e & 15);
            if (r0 + r >= m || c >= b) continue;
            s8_t h, l;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float x = (c + q < b) ? sB[r][c + q] : 0.0f;
                const unsigned short hb = f32_to_bf16_rne(x);
                h[q] = (short)hb;
                l[q] = (short)f32_to_bf16_rne(x - bf16_to_f32(hb));
            }
            *reinterpret_cast<s8_t*>(Hi + (r0 + r) * ldk + c) = h;
            *reinterpret_cast<s8_t*>(Lo + (r0 + r) * ldk + c) = l;
        }
    }
}

hipError_t launch_potrf_inv_f32(int b, float* A, i64 lda, float* Linv, int* info, int col_offset, hipStream_t s) {
    if (b <= 0) return hipSuccess;
    hipLaunchKernelGGL(potrf_f32_kernel, dim3(1), dim3(256), 0, s, b, A, lda, info, col_offset);
    hipLaunchKernelGGL(trtri_f32_kernel, dim3(1), dim3(256), 0, s, b, (const float*)A, lda, Linv, (const int*)info,
                       col_offset + b);
    return hipGetLastError();
}

hipError_t launch_panel_inv_f32(int m, int b, const float* Linv, float* B, i64 ldb, const int* info, int limit,
                                unsigned short* Hi, unsigned short* Lo, i64 ldk, hipStream_t s) {
    if (m <= 0 || b <= 0) return hipSuccess;
    hipLaunchKernelGGL(panel_inv_f32_kernel, dim3((unsigned)ceil_div(m, 64)), dim3(256), 0, s, m, b, Linv, B, ldb,
                       info, limit, Hi, Lo, ldk);
    return hipGetLastError();
}

}  // namespace elemamd
