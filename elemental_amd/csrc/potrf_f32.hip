// Panel step of the low-precision (float-storage) factorization: the
// diagonal block L11 = chol(A11) of cholesky::LVar3
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:83-123, LVar3Unb at :19-49)
// and the panel solve L21 := A21 L11^{-T}.  Two kernels:
//
// potrf_f32_kernel -- one workgroup, the block in LDS as float, right-looking
//   in 32-column steps: wave 0 factors the 32 x 32 diagonal piece in FP64;
//   all waves invert it (recursive doubling, FP64) and multiply the rows
//   below by the inverse (FP64 MFMA), then apply the rank-32 update to the
//   rest of the block (FP32 MFMA).  A block narrower than 128 is padded with
//   the identity.
// panel_f32_kernel -- one 64-row slab per workgroup, with the inverses of
//   the four 32 x 32 diagonal pieces of L11 (by the diagonal-block kernel):
//   each wave solves its 16 rows block by block,
//       X_p = (A_p - sum_{q<p} X_q L_pq^T) inv(L_pp)^T   (FP32 MFMA),
//   with no further synchronisation (the rows are independent).  Products
//   with the 32 x 32 inverses replace the 128-step substitution per row; the
//   refinement absorbs their larger rounding error (this factor only
//   preconditions).  With planes requested the result is also written as
//   bf16 hi/lo planes for the split-BF16x3 updates.
#include "common.hpp"
#include "kernels.hpp"
#include "solve_dev.hpp"

namespace elemamd {

constexpr int kPiLd = 132;  // LDS row stride (floats): MFMA operand reads are conflict-free

// One wave: factor the 32 x 32 block of sA at (c0, c0) in FP64 into sD
// (transposed) and back into sA as float, with the reciprocal diagonal in sRd.
// Returns 0, or the 1-based failing column.
// Lane (h, i) = (lane / 32, lane % 32) holds columns 16h .. 16h+15 of row i.
// Column j travels through LDS (the wave's own writes and reads, in order;
// two buffers by column parity so the next column's write need not wait for
// this column's reads): its owners publish the entries below the pivot
// (zeros above it) and the pivot, every lane reads the pivot and the 16
// entries of its half, and the rank-1 update runs unmasked because the
// published zeros leave finished entries unchanged.  Each update starts with
// the next column's entry, so its publication can overlap the rest.  The
// pivot's reciprocal square root is the hardware estimate refined by one
// Newton step (error far below the float the factor is stored in) and is
// kept, so nothing divides.
__device__ inline int factor32(float (*sA)[kPiLd], double (*sD)[34], double (*sCol)[34], double* sRd, int c0,
                               int lane) {
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
        double* col = sCol[j & 1];
        int jv = j;  // opaque per column: the lane masks are made here, not all up front
        asm volatile("" : "+s"(jv));
        const bool own = hh == (jv >> 4), diag = i == jv;
        if (own) col[i] = (i > jv) ? d[jq] : 0.0;
        // the column's readers are other lanes of this wave: the wave
        // barrier keeps every read below the store (without it the compiler
        // may read in the non-owners' branch, ahead of the owners' store)
        __builtin_amdgcn_wave_barrier();
        // the pivot comes straight from its lane's register (no LDS round
        // trip on the test); the 16 column entries are read in one batch
        const int pl = 32 * jh + j;
        const double piv = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(d[jq]), pl),
                                            __builtin_amdgcn_readlane(__double2loint(d[jq]), pl));
        const double cij = col[i];
        double ck[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) ck[q] = col[16 * hh + q];
        if (!(piv > 0.0)) return j + 1;
        double y = __builtin_amdgcn_rsq(piv);
        y = fma(y, fma(-0.5 * piv * y, y, 0.5), y);
        const double lij = cij * y;
        if (own) d[jq] = diag ? piv * y : lij;
        const double sij = lij * y;  // a_ij / a_jj: one multiplier for the whole row
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = (jq + 1 + u) & 15;  // the next column's entry first
            d[q] = fma(-sij, ck[q], d[q]);
        }
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

// All threads: X = inv(L) of the piece factored by factor32 (FP64, L in sD
// transposed, 1/L_kk in sRd) by recursive doubling over diagonal blocks of
// size b = 1, 2, 4, 8, 16:  for L = [L11 0; L21 L22],
//     X21 = -X22 (L21 X11)
// with X11, X22 already inverted; every element of a level is one thread (a
// dot of length <= b), so the 32-step chain of a substitution becomes 10
// short steps.  sX: row-major, zero above the diagonal; sT: scratch.
__device__ inline void invert32(const double (*sD)[34], const double* sRd, double (*sX)[33], double (*sT)[33],
                                int t) {
    for (int e = t; e < 1024; e += 256) {
        const int i = e >> 5, j = e & 31;
        sX[i][j] = (i == j) ? sRd[i] : 0.0;
    }
    __syncthreads();
    for (int b = 1; b < 32; b *= 2) {
        const int ne = 16 * b;  // elements of all the level's X21 blocks
        const int q = t / (b * b), i = (t % (b * b)) / b, j = t % b;
        const int r0 = 2 * b * q, rb = r0 + b;
        if (t < ne) {  // T = L21 X11 (X11 lower: m >= j)
            double acc = 0.0;
            for (int m = j; m < b; ++m) acc = fma(sD[r0 + m][rb + i], sX[r0 + m][r0 + j], acc);
            sT[rb + i][r0 + j] = acc;
        }
        __syncthreads();
        if (t < ne) {  // X21 = -X22 T (X22 lower: m <= i)
            double acc = 0.0;
            for (int m = 0; m <= i; ++m) acc = fma(sX[rb + i][rb + m], sT[rb + m][r0 + j], acc);
            sX[rb + i][r0 + j] = -acc;
        }
        __syncthreads();
    }
}

// All threads: the rows below the piece, r in [c0 + 32, 128), columns
// c0 .. c0 + 31 of sA := A L^{-T} = A X^T with X = inv(L) (FP64 MFMA
// 16x16x4, 16 x 16 output tiles dealt to the waves).  Every wave reads its
// operands, then the workgroup synchronises, then results overwrite sA.
__device__ inline void below_times_xt(float (*sA)[kPiLd], const double (*sX)[33], int c0, int w, int lane) {
    const int m = lane & 15, g = lane >> 4;
    const int ntile = 2 * ((96 - c0) / 16);  // row tiles x 2 column tiles
    double fa[3][8], fb[3][8];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int tt = w + 4 * u, rt = tt >> 1, ct = tt & 1;
        const int ra = tt < ntile ? c0 + 32 + 16 * rt + m : 127;  // clamped when the tile does not exist
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            fa[u][kb] = (double)sA[ra][c0 + 4 * kb + g];  // A[i = m][k = 4 kb + g]
            fb[u][kb] = sX[16 * ct + m][4 * kb + g];      // B[k][j = m] = X[j][k]
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int tt = w + 4 * u, rt = tt >> 1, ct = tt & 1;
        if (tt >= ntile) continue;
        d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][kb], fb[u][kb], acc, 0, 0, 0);
        // a dependent zero-operand MFMA right before the accumulator is read
        // (the wait states between an FP64 MFMA and a read of its result)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) sA[c0 + 32 + 16 * rt + g + 4 * v][c0 + 16 * ct + m] = (float)acc[v];
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
// its lower triangle, the inverses of its four 32 x 32 diagonal pieces go to
// Xd (4 x 32 x 32 floats, row-major).  A failing pivot records
// col_offset + column in *info.
__global__ __launch_bounds__(256) void potrf_f32_kernel(int b, float* __restrict__ A, i64 lda,
                                                        float* __restrict__ Xd, int* __restrict__ info,
                                                        int col_offset) {
    __shared__ float sA[128][kPiLd];
    __shared__ double sD[32][34];  // the FP64 factor of the current piece, transposed
    __shared__ double sCol[2][34], sRd[32];
    __shared__ double sX[32][33], sT[32][33];  // the current piece's inverse, scratch
    __shared__ int sfail;
    // the factorization's chain: its waves issue ahead of the trailing
    // update's waves sharing the SIMD (wave priority, not queue priority)
    __builtin_amdgcn_s_setprio(3);
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
        invert32(sD, sRd, sX, sT, t);
        for (int e = t; e < 1024; e += 256) Xd[p * 1024 + e] = (float)sX[e >> 5][e & 31];
        below_times_xt(sA, sX, c0, w, lane);
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

// Rows rw .. rw+15 of a panel slab (staged in sB) -= P_slab P_top^T: the
// update of the block column from the group's earlier panels, which these
// rows have not received yet.  upd points at the planes' row of the block's
// first (diagonal) row, so P_top = upd rows 0 .. b-1 and the slab's rows are
// upd rows b + r0 + rw + i; K = kk columns (the planes are zero-padded to a
// multiple of 32 columns).  One wave, split-bf16 products (all plane pairs
// p + q < NP), fragments loaded straight from the planes into registers,
// the next 32-wide chunk loading while the current one multiplies.
template <int NP>
__device__ __forceinline__ void panel_slab_update(float (*sB)[kPiLd], const unsigned short* __restrict__ upd, i64 ps,
                                                  i64 ldk, int kk, int b, int m, i64 r0, int rw, int m16, int g) {
    const i64 rows = (i64)b + m;  // rows of the planes block
    const i64 ra = b + r0 + rw + m16 < rows ? b + r0 + rw + m16 : rows - 1;  // clamped; its outputs are dropped
    const unsigned short* a_p = upd + ra * ldk + 8 * g;
    const s8_t neg = {(short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000,
                      (short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000};
    const int nchunks = (kk + 31) / 32;
    // two passes over 64 columns each keep the fragments in few registers
    // (this kernel must stay small enough to be placed beside the trailing
    // update's workgroups)
    for (int h = 0; h < 2; ++h) {
        if (64 * h >= b) break;
        f4_t acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[nt][v] = sB[rw + 4 * g + v][64 * h + 16 * nt + m16];
        constexpr int kR = 3;  // chunks in flight: the loads are L2-latency bound
        s8_t fa[kR][NP], fb[kR][4][NP];  // [buffer][plane], [buffer][column tile][plane]
        auto load = [&](int buf, int kc) __attribute__((always_inline)) {
#pragma unroll
            for (int p = 0; p < NP; ++p) fa[buf][p] = *reinterpret_cast<const s8_t*>(a_p + p * ps + kc) ^ neg;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int c = 64 * h + 16 * nt + m16;
                const i64 rb = (i64)(c < b ? c : b - 1) * ldk + 8 * g + kc;
#pragma unroll
                for (int p = 0; p < NP; ++p) fb[buf][nt][p] = *reinterpret_cast<const s8_t*>(upd + p * ps + rb);
            }
        };
        auto step = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int o = NP - 1; o >= 0; --o)  // term order o = p + q, smallest first
#pragma unroll
                    for (int p = o; p >= 0; --p) mfma_acc_bf16(acc[nt], fa[buf][p], fb[buf][nt][o - p]);
        };
        // chunk c lives in buffer c % kR; chunk c + kR - 1 loads while chunk c multiplies
#pragma unroll
        for (int c = 0; c < kR - 1; ++c)
            if (c < nchunks) load(c, 32 * c);
        for (int c0 = 0; c0 < nchunks; c0 += kR) {
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const int c = c0 + u;
                if (c >= nchunks) break;
                if (c + kR - 1 < nchunks) load((u + kR - 1) % kR, 32 * (c + kR - 1));
                step(u);
            }
        }
        mfma_retire();
        // columns past b stay zero (the padding the solve relies on)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (64 * h + 16 * nt + m16 < b) sB[rw + 4 * g + v][64 * h + 16 * nt + m16] = acc[nt][v];
    }
}

// The same update for two planes with P_top staged through LDS: the four
// waves share each 32-column chunk of P_top's planes (16 KB) instead of
// each loading all of it (the update is bound by that traffic), and chunk
// c + 2's global loads are in flight while chunk c multiplies.  sP: two
// chunk buffers, [buffer][plane][row 0..127][32 columns] bf16 (32 KB).
// All threads of the workgroup call it (it synchronises the workgroup).
__device__ __forceinline__ void panel_slab_update_lds(float (*sB)[kPiLd], unsigned short* sP,
                                                      const unsigned short* __restrict__ upd, i64 ps, i64 ldk, int kk,
                                                      int b, int m, i64 r0, int t, int w, int m16, int g) {
    constexpr int NP = 2, kR = 3;
    const int rw = 16 * w;
    const i64 rows = (i64)b + m;
    const i64 ra = b + r0 + rw + m16 < rows ? b + r0 + rw + m16 : rows - 1;  // clamped; its outputs are dropped
    const unsigned short* a_p = upd + ra * ldk + 8 * g;
    const s8_t neg = {(short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000,
                      (short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000};
    // staging: thread t carries 16-byte segments 2 (t / 128) and 2 (t / 128) + 1
    // of P_top row t % 128 (rows past b repeat row b - 1; their outputs are dropped)
    const int sr = t & 127, sseg = (t >> 7) * 2;
    const unsigned short* s_src = upd + (i64)(sr < b ? sr : b - 1) * ldk + 8 * sseg;
    s8_t st[kR][NP][2], fa[kR][NP];
    auto gload = [&](int slot, int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int h = 0; h < 2; ++h) st[slot][p][h] = *reinterpret_cast<const s8_t*>(s_src + p * ps + kc + 8 * h);
            fa[slot][p] = *reinterpret_cast<const s8_t*>(a_p + p * ps + kc) ^ neg;
        }
    };
    auto sstore = [&](int slot, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<s8_t*>(sP + ((buf * NP + p) * 128 + sr) * 32 + 8 * (sseg + h)) = st[slot][p][h];
    };
    f4_t acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[nt][v] = sB[rw + 4 * g + v][16 * nt + m16];
    const int nchunks = (kk + 31) / 32;
#pragma unroll
    for (int c = 0; c < kR - 1; ++c)
        if (c < nchunks) gload(c, 32 * c);
    for (int c0 = 0; c0 < nchunks; c0 += kR) {
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int c = c0 + u;  // chunk c: registers slot u, LDS buffer c % 2
            if (c >= nchunks) break;
            const int buf = c & 1;
            sstore(u, buf);
            // chunk c is staged, and every wave is done with chunk c - 2,
            // the last user of this buffer
            __syncthreads();
            if (c + kR - 1 < nchunks) gload((u + kR - 1) % kR, 32 * (c + kR - 1));
#pragma unroll
            for (int nt = 0; nt < 8; ++nt) {
                s8_t fb[NP];
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    fb[p] = *reinterpret_cast<const s8_t*>(sP + ((buf * NP + p) * 128 + 16 * nt + m16) * 32 + 8 * g);
#pragma unroll
                for (int o = NP - 1; o >= 0; --o)  // term order o = p + q, smallest first
#pragma unroll
                    for (int p = o; p >= 0; --p) mfma_acc_bf16(acc[nt], fa[u][p], fb[o - p]);
            }
        }
    }
    mfma_retire();
    // columns past b stay zero (the padding the solve relies on)
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (16 * nt + m16 < b) sB[rw + 4 * g + v][16 * nt + m16] = acc[nt][v];
}

// B (m x b, b <= 128) := B L11^{-T}, L11 the b x b float factor at L; one
// 64-row slab per workgroup, the result staged through LDS so the
// column-major stores are contiguous.  With NPU > 0 the slab first takes the
// block column's pending update from the group's earlier panels (NPU planes
// at upd, uk columns; panel_slab_update), so no separate kernel has to
// update the rows below the diagonal block before this one runs.
template <int NPU>
__global__ __launch_bounds__(256) void panel_f32_kernel(int m, int b, const float* __restrict__ L, i64 ldl,
                                                        const float* __restrict__ Xd, float* __restrict__ B, i64 ldb,
                                                        const int* __restrict__ info, int limit,
                                                        unsigned short* __restrict__ planes, int np, i64 ps,
                                                        i64 ldk, const unsigned short* __restrict__ upd, int uk) {
    // LDS kept under half a CU (80 KB) so workgroups of this kernel fit beside
    // two of the trailing update's: L11 only as its six 32 x 32 blocks below
    // the diagonal pieces (block (p, q), q < p, at p (p - 1) / 2 + q; zero
    // outside the b x b block, whose padding is the identity) and the
    // inverses of its four 32 x 32 diagonal pieces, in one area that first
    // stages the slab's pending update
    __shared__ __align__(16) float sU[10 * 32 * 36];
    float(*sLb)[32][36] = reinterpret_cast<float(*)[32][36]>(sU);
    float(*sXd)[32][36] = reinterpret_cast<float(*)[32][36]>(sU + 6 * 32 * 36);
    __shared__ float sB[64][kPiLd];
    __builtin_amdgcn_s_setprio(3);  // chain kernel (see potrf_f32_kernel)
    if (info != nullptr && *info != 0 && *info <= limit) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m16 = lane & 15, g = lane >> 4;
    const i64 r0 = (i64)blockIdx.x * 64;
    const int rw = 16 * w;
    float v[32];
    {  // the slab: row t % 64, columns t / 64 + 4q
        const int rs = t & 63, c_lo = t >> 6;
        const i64 rb = r0 + rs < m ? r0 + rs : m - 1;
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int c = c_lo + 4 * q;
            v[q] = B[rb + (i64)(c < b ? c : b - 1) * ldb];
        }
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int c = c_lo + 4 * q;
            sB[rs][c] = (r0 + rs < m && c < b) ? v[q] : 0.0f;
        }
    }
    // the slab's pending update: wave w on rows 16w .. 16w+15
    if constexpr (NPU > 0) {
        __syncthreads();
        if constexpr (NPU == 2)
            panel_slab_update_lds(sB, reinterpret_cast<unsigned short*>(sU), upd, ps, ldk, uk, b, m, r0, t, w, m16, g);
        else
            panel_slab_update<NPU>(sB, upd, ps, ldk, uk, b, m, r0, rw, m16, g);
        __syncthreads();  // the staging area becomes L11's blocks
    }
    // block e / 1024: row e % 32 (lanes along the column), column (e / 32) % 32
    // (clamped loads, then select)
    for (int e0 = 0; e0 < 6 * 1024; e0 += 32 * 256) {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int e = e0 + t + 256 * q;
            const int blk = e >> 10, rr = e & 31, cc = (e >> 5) & 31;
            const int p = blk < 1 ? 1 : blk < 3 ? 2 : 3, qb = blk - p * (p - 1) / 2;
            const int r = 32 * p + rr, c = 32 * qb + cc;
            v[q] = (e < 6 * 1024) ? L[(r < b ? r : b - 1) + (i64)(c < b ? c : b - 1) * ldl] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int e = e0 + t + 256 * q;
            const int blk = e >> 10, rr = e & 31, cc = (e >> 5) & 31;
            const int p = blk < 1 ? 1 : blk < 3 ? 2 : 3, qb = blk - p * (p - 1) / 2;
            const int r = 32 * p + rr, c = 32 * qb + cc;
            if (e < 6 * 1024) sLb[blk][rr][cc] = (r < b && c < b) ? v[q] : 0.0f;
        }
    }
    for (int e = t; e < 4 * 1024; e += 256) sXd[e >> 10][(e >> 5) & 31][e & 31] = Xd[e];
    __syncthreads();
    // wave w: rows 16w .. 16w+15 of the slab, pieces p = 0..3 in order (no
    // other wave touches these rows)
    for (int p = 0; p < 4; ++p) {
        f4_t acc[2] = {};
        for (int q = 0; q < p; ++q) {  // S = sum_q X_q L_pq^T
            const float(*Lpq)[36] = sLb[p * (p - 1) / 2 + q];
            tile_ab_t(acc[0], sB, rw, 32 * q, Lpq, 0, 0, m16, g);
            tile_ab_t(acc[1], sB, rw, 32 * q, Lpq, 16, 0, m16, g);
        }
        mfma_retire();
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) sB[rw + 4 * g + v][32 * p + 16 * nt + m16] -= acc[nt][v];
        // X_p = T inv(L_pp)^T (T = the updated columns, this wave's rows only)
        f4_t x[2] = {};
        tile_ab_t(x[0], sB, rw, 32 * p, sXd[p], 0, 0, m16, g);
        tile_ab_t(x[1], sB, rw, 32 * p, sXd[p], 16, 0, m16, g);
        mfma_retire();
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) sB[rw + 4 * g + v][32 * p + 16 * nt + m16] = x[nt][v];
    }
    __syncthreads();
    for (int idx = t; idx < 64 * 128; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        if (r0 + r < m && c < b) B[r0 + r + (i64)c * ldb] = sB[r][c];
    }
    if (planes != nullptr) {  // bf16 planes of the result: thread = (row, 8 columns), 16-byte stores
        for (int e = t; e < 64 * 16; e += 256) {
            const int r = e >> 4, c = 8 * (e & 15);
            if (r0 + r >= m || c >= b) continue;
            float x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = (c + q < b) ? sB[r][c + q] : 0.0f;
            for (int p = 0; p < np; ++p) {
                s8_t h;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const unsigned short hb = f32_to_bf16_rne(x[q]);
                    h[q] = (short)hb;
                    x[q] -= bf16_to_f32(hb);
                }
                *reinterpret_cast<s8_t*>(planes + p * ps + (r0 + r) * ldk + c) = h;
            }
        }
    }
}

hipError_t launch_potrf_f32(int b, float* A, i64 lda, float* Xd, int* info, int col_offset, hipStream_t s) {
    if (b <= 0) return hipSuccess;
    hipLaunchKernelGGL(potrf_f32_kernel, dim3(1), dim3(256), 0, s, b, A, lda, Xd, info, col_offset);
    return hipGetLastError();
}

hipError_t launch_panel_f32(int m, int b, const float* L, i64 ldl, const float* Xd, float* B, i64 ldb,
                            const int* info, int limit, unsigned short* planes, int np, i64 ps, i64 ldk,
                            hipStream_t s, const unsigned short* upd, int uk) {
    if (m <= 0 || b <= 0) return hipSuccess;
    const dim3 grid((unsigned)ceil_div(m, 64));
    if (upd == nullptr || uk <= 0)
        hipLaunchKernelGGL(panel_f32_kernel<0>, grid, dim3(256), 0, s, m, b, L, ldl, Xd, B, ldb, info, limit, planes,
                           np, ps, ldk, upd, 0);
    else if (np == 2)
        hipLaunchKernelGGL(panel_f32_kernel<2>, grid, dim3(256), 0, s, m, b, L, ldl, Xd, B, ldb, info, limit, planes,
                           np, ps, ldk, upd, uk);
    else if (np == 3)
        hipLaunchKernelGGL(panel_f32_kernel<3>, grid, dim3(256), 0, s, m, b, L, ldl, Xd, B, ldb, info, limit, planes,
                           np, ps, ldk, upd, uk);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace elemamd
