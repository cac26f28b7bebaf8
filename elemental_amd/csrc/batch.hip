// Batched HPD factor-and-solve for many independent matrices of uneven size
// (no reference counterpart: Elemental factors one matrix per call).
//
// Each matrix is the same computation as the single-matrix path: LVar3 with
// block size nb (the same potrf / panel-solve / trailing-update device code,
// so a matrix gives the same result alone and inside a batch), then the
// forward and backward solves.  Matrices are processed in order of decreasing
// size; step j of the factorization of every matrix that has a step j runs in
// one launch per kernel type, and a workgroup finds its (matrix, tile) from a
// per-step prefix sum of work counts, so small and large matrices share
// launches without idle workgroups.  The whole sequence is captured in a
// hipGraph by the host driver and replayed as one submission.
//
// Layout: matrix i is column-major with ld = n_i at element offset a_off[i] of
// one buffer; its right-hand sides (n_i x k_i, ld = n_i) at b_off[i].
#include <algorithm>

#include "batch.hpp"
#include "chain64_dev.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "panel_dev.hpp"
#include "potrf_dev.hpp"
#include "syrk_dev.hpp"

namespace elemamd {

// index p with prefix[p] <= w < prefix[p+1]
__device__ inline int find_item(const int* prefix, int cnt, int w) {
    int lo = 0, hi = cnt;  // prefix[0] = 0 <= w < prefix[cnt]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------- factorization
template <typename T>
__global__ __launch_bounds__(512) void batch_potrf_kernel(BatchArgs a, T* A, int step, int nb) {
    const int p = blockIdx.x;
    const int n = a.n[p];
    const int c0 = step * nb;
    if (c0 >= n) return;
    if (a.skip != nullptr && a.skip[p]) return;
    const int b = min(nb, n - c0);
    potrf_unb_block<T>(b, A + a.a_off[p] + c0 + (i64)c0 * n, n, a.info + p, c0);
}

template <typename T>
__global__ __launch_bounds__(256) void batch_trsm_kernel(BatchArgs a, T* A, int step, int nb, const int* prefix,
                                                         int cnt) {
    __shared__ double sDg[2][32][32];
    __shared__ double sRinv[128];
    __shared__ double sXq[4][16][34];
    const int p = find_item(prefix, cnt, blockIdx.x);
    const int rb = blockIdx.x - prefix[p];
    const int n = a.n[p];
    const int c0 = step * nb;
    const int m = n - c0 - nb;
    if (m <= 0) return;  // (prefix gives such matrices no workgroups)
    if (a.skip != nullptr && a.skip[p]) return;
    const int* info = a.info + p;
    if (*info != 0 && *info <= c0 + nb) return;
    T* L = A + a.a_off[p] + c0 + (i64)c0 * n;
    chain64::panel64_slab<T>(m, nb, L, n, L + nb, n, rb, sDg, sRinv, sXq);
}

template <typename T, int MODE>
__global__ __launch_bounds__(256, 2) void batch_syrk_kernel(BatchArgs a, T* A, int step, int nb, const int* prefix,
                                                            int cnt) {
    const int p = find_item(prefix, cnt, blockIdx.x);
    const int t = blockIdx.x - prefix[p];
    const int n = a.n[p];
    const int c0 = step * nb;
    const int m = n - c0 - nb;
    if (m <= 0) return;
    if (a.skip != nullptr && a.skip[p]) return;
    const int* info = a.info + p;
    if (*info != 0 && *info <= c0 + nb) return;
    int ti, tj;
    tri_tile(t, ti, tj);
    const T* P = A + a.a_off[p] + (c0 + nb) + (i64)c0 * n;
    T* C = A + a.a_off[p] + (c0 + nb) + (i64)(c0 + nb) * n;
    if constexpr (MODE == 0)
        syrk_tile_f64(m, m, nb, -1.0, (const double*)P, n, (const double*)P, n, 1.0, (double*)C, n, ti, tj);
    else if constexpr (MODE == 1)
        syrk_tile_f32(m, m, nb, -1.0f, (const float*)P, n, (const float*)P, n, 1.0f, (float*)C, n, ti, tj);
    else
        syrk_tile_bf16x3(m, m, nb, -1.0f, (const float*)P, n, (const float*)P, n, 1.0f, (float*)C, n, ti, tj);
}

// ------------------------------------------------------------------ solve
// One workgroup per matrix: forward then backward solve in 64-row blocks,
// 16 right-hand sides at a time.  Contributions of solved blocks are staged
// through LDS and applied with fused multiply-adds; the 64x64 diagonal block
// is solved by substitution (one wave per four right-hand sides, lanes =
// rows), the backward sweep as the forward sweep on reversed indices.
constexpr int kBS = 64;
constexpr int kBC = 16;

template <typename T>
__device__ inline T bcast_lane(T v, int lane);
template <>
__device__ inline double bcast_lane<double>(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
template <>
__device__ inline float bcast_lane<float>(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// LDS from the caller: sL (contraction tile / diagonal block), sX (solved
// block / right-hand sides), sR (block being solved), sRinv.
template <typename T>
__device__ __forceinline__ void batch_solve_one(const BatchArgs& a, const T* L_all, T* B_all, int p,
                                                T (*sL)[kBS + 1], T (*sX)[kBC + 1], T (*sR)[kBC + 1], T* sRinv) {
    if (a.info[p] != 0) return;
    if (a.skip != nullptr && a.skip[p]) return;
    const int n = a.n[p], k = a.k[p];
    const T* L = L_all + a.a_off[p];
    T* B = B_all + a.b_off[p];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nT = (n + kBS - 1) / kBS;
    for (int cs = 0; cs < k; cs += kBC) {
        const int kc = min(kBC, k - cs);
        for (int dir = 0; dir < 2; ++dir) {
            const bool back = dir == 1;
            for (int ii = 0; ii < nT; ++ii) {
                const int I = back ? (nT - 1 - ii) : ii;
                const int r0 = I * kBS;
                const int bs = min(kBS, n - r0);
                // thread (row pl = t % 64, column group cg = t / 64: columns cg, cg+4, cg+8, cg+12)
                const int pl = t & 63, cg = t >> 6;
                T acc[4] = {T(0), T(0), T(0), T(0)};
                for (int jj = 0; jj < ii; ++jj) {
                    const int J = back ? (nT - 1 - jj) : jj;
                    const int q0 = J * kBS;
                    const int qs = min(kBS, n - q0);
                    __syncthreads();
                    // tile M[q][p]: forward L(r0+p, q0+q); backward L(q0+q, r0+p).
                    // All loads of a thread first (clamped addresses), then LDS.
                    {
                        const int x = t & 63, y0 = t >> 6;
                        T g[16], gx[4];
#pragma unroll
                        for (int s2 = 0; s2 < 16; ++s2) {
                            const int y = y0 + 4 * s2;
                            const bool ok = !back ? (x < bs && y < qs) : (x < qs && y < bs);
                            const i64 off = !back ? (r0 + x) + (i64)(q0 + y) * n : (q0 + x) + (i64)(r0 + y) * n;
                            const T v = L[ok ? off : 0];
                            g[s2] = ok ? v : T(0);
                        }
#pragma unroll
                        for (int s2 = 0; s2 < 4; ++s2) {
                            const int c = y0 + 4 * s2;
                            const bool ok = x < qs && c < kc;
                            const T v = B[ok ? (q0 + x) + (i64)(cs + c) * n : 0];
                            gx[s2] = ok ? v : T(0);
                        }
#pragma unroll
                        for (int s2 = 0; s2 < 16; ++s2) {
                            const int y = y0 + 4 * s2;
                            if (!back) sL[y][x] = g[s2];
                            else sL[x][y] = g[s2];
                        }
#pragma unroll
                        for (int s2 = 0; s2 < 4; ++s2) sX[x][y0 + 4 * s2] = gx[s2];
                    }
                    __syncthreads();
                    for (int q = 0; q < qs; ++q) {
                        const T mv = sL[q][pl];
#pragma unroll
                        for (int u = 0; u < 4; ++u) acc[u] += mv * sX[q][cg + 4 * u];
                    }
                }
                __syncthreads();
                // R = B_I - acc
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = cg + 4 * u;
                    sR[pl][c] = (pl < bs && c < kc) ? B[(r0 + pl) + (i64)(cs + c) * n] - acc[u] : T(0);
                }
                // diagonal block in lower-triangular form on (possibly reversed) indices
                {
                    const int x = t & 63, y0 = t >> 6;  // M(x, y) into sL[y][x]
                    T g[16];
#pragma unroll
                    for (int s2 = 0; s2 < 16; ++s2) {
                        const int y = y0 + 4 * s2;
                        const int gi = back ? (kBS - 1 - y) : x;
                        const int gk = back ? (kBS - 1 - x) : y;
                        const bool ok = gi < bs && gk < bs && x >= y;
                        const T v = L[ok ? (r0 + gi) + (i64)(r0 + gk) * n : 0];
                        g[s2] = ok ? v : ((x == y) ? T(1) : T(0));
                    }
#pragma unroll
                    for (int s2 = 0; s2 < 16; ++s2) sL[y0 + 4 * s2][x] = g[s2];
                }
                __syncthreads();
                if (t < kBS) sRinv[t] = T(1) / sL[t][t];
                __syncthreads();
                const int myrow = back ? (kBS - 1 - lane) : lane;
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = sR[myrow][4 * w + u];
                for (int kk = 0; kk < kBS; ++kk) {
                    const T rk = sRinv[kk];
                    const T lik = sL[kk][lane];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const T y = bcast_lane<T>(v[u], kk) * rk;
                        if (lane == kk) v[u] = y;
                        else if (lane > kk) v[u] -= lik * y;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = 4 * w + u;
                    if (myrow < bs && c < kc) B[(r0 + myrow) + (i64)(cs + c) * n] = v[u];
                }
                __threadfence_block();
                __syncthreads();
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void batch_solve_kernel(BatchArgs a, const T* L_all, T* B_all) {
    __shared__ T sL[kBS][kBS + 1];
    __shared__ T sX[kBS][kBC + 1];
    __shared__ T sR[kBS][kBC + 1];
    __shared__ T sRinv[kBS];
    batch_solve_one<T>(a, L_all, B_all, blockIdx.x, sL, sX, sR, sRinv);
}

// ------------------------------------------------------ persistent launch
// One launch per size bucket: every workgroup takes the bucket's matrices in
// processing order (largest first) from a counter and factors each one whole
// with the device code of the stepwise launches above -- LVar3Unb on the
// diagonal block (256 threads, the same operation order), the register
// panel solve, the trailing tiles in the same K order -- so a matrix's
// factor is the same bit for bit in either form; for FP64 it then solves the
// matrix's right-hand sides.  Workgroups never wait on each other.  The
// phases share one LDS arena (the panel, trailing and solve phases never
// overlap), so two workgroups fit on a CU.
constexpr int kArenaPanel = (2 * 32 * 32 + 128 + 4 * 16 * 34) * 8;
constexpr int kArenaSyrk64 = 2 * 2 * kSyrkKC64 * kSyrkLdp * 8;
constexpr int kArenaSyrk32 = 2 * 2 * kSyrkKC * kSyrkLdpF * 4;
constexpr int kArenaBf16 = 4 * kSyrkTile * kBfLd * 2;
constexpr int kArenaSolve = (kBS * (kBS + 1) + 2 * kBS * (kBC + 1) + kBS) * 8;
constexpr int cmax(int x, int y) { return x > y ? x : y; }
constexpr int kArenaBytes = cmax(cmax(kArenaPanel, kArenaSyrk64), cmax(cmax(kArenaSyrk32, kArenaBf16), kArenaSolve));

// The phases are separate (not inlined) functions: each gets the registers
// it needs on its own (inlined together they spilled).
template <typename T>
__device__ __noinline__ void persist_factor_block(int b, T* L, int n, int* info, int c0) {
    potrf_unb_block<T, 256>(b, L, n, info, c0);
}

template <typename T>
__device__ __noinline__ void persist_panel(int m, int b, T* L, int n, char* arena) {
    double(*sDg)[32][32] = reinterpret_cast<double(*)[32][32]>(arena);
    double* sRinv = reinterpret_cast<double*>(arena + 2 * 32 * 32 * 8);
    double(*sXq)[16][34] = reinterpret_cast<double(*)[16][34]>(arena + (2 * 32 * 32 + 128) * 8);
    for (int rb = 0; rb < (m + 63) / 64; ++rb) {
        chain64::panel64_slab<T>(m, b, L, n, L + b, n, rb, sDg, sRinv, sXq);
        __syncthreads();
    }
}

template <typename T, int MODE>
__device__ __noinline__ void persist_trailing(int m, int b, const T* P, T* C, int n, char* arena) {
    const int nt = (m + 127) / 128;
    for (int tt = 0; tt < nt * (nt + 1) / 2; ++tt) {
        int ti, tj;
        tri_tile(tt, ti, tj);
        if constexpr (MODE == 0) {
            auto sI = reinterpret_cast<double(*)[kSyrkKC64][kSyrkLdp]>(arena);
            auto sJ = reinterpret_cast<double(*)[kSyrkKC64][kSyrkLdp]>(arena + kArenaSyrk64 / 2);
            syrk_tile_f64_lds(m, m, b, -1.0, (const double*)P, n, (const double*)P, n, 1.0, (double*)C, n, ti, tj,
                              sI, sJ);
        } else if constexpr (MODE == 1) {
            auto sI = reinterpret_cast<float(*)[kSyrkKC][kSyrkLdpF]>(arena);
            auto sJ = reinterpret_cast<float(*)[kSyrkKC][kSyrkLdpF]>(arena + kArenaSyrk32 / 2);
            syrk_tile_f32_lds(m, m, b, -1.0f, (const float*)P, n, (const float*)P, n, 1.0f, (float*)C, n, ti, tj,
                              sI, sJ);
        } else {
            constexpr int q = kArenaBf16 / 4;
            auto s0 = reinterpret_cast<unsigned short(*)[kBfLd]>(arena);
            auto s1 = reinterpret_cast<unsigned short(*)[kBfLd]>(arena + q);
            auto s2 = reinterpret_cast<unsigned short(*)[kBfLd]>(arena + 2 * q);
            auto s3 = reinterpret_cast<unsigned short(*)[kBfLd]>(arena + 3 * q);
            syrk_tile_bf16x3_lds(m, m, b, -1.0f, (const float*)P, n, (const float*)P, n, 1.0f, (float*)C, n, ti, tj,
                                 s0, s1, s2, s3);
        }
        __syncthreads();
    }
}

__device__ __noinline__ void persist_solve(const BatchArgs& a, const double* A, double* B, int p, char* arena) {
    auto sL = reinterpret_cast<double(*)[kBS + 1]>(arena);
    auto sX = reinterpret_cast<double(*)[kBC + 1]>(arena + kBS * (kBS + 1) * 8);
    auto sR = reinterpret_cast<double(*)[kBC + 1]>(arena + (kBS * (kBS + 1) + kBS * (kBC + 1)) * 8);
    double* sRinv = reinterpret_cast<double*>(arena + (kBS * (kBS + 1) + 2 * kBS * (kBC + 1)) * 8);
    batch_solve_one<double>(a, A, B, p, sL, sX, sR, sRinv);
}

template <typename T, int MODE, bool SOLVE>
__global__ __launch_bounds__(256, 2) void batch_persistent_kernel(BatchArgs a, T* A, double* B, int nb, int* counter) {
    __shared__ __align__(16) char arena[kArenaBytes];
    __shared__ int s_p;
    for (;;) {
        __syncthreads();  // every thread is done with the previous matrix
        if (threadIdx.x == 0) s_p = atomicAdd(counter, 1);
        __syncthreads();
        const int p = s_p;
        if (p >= a.count) return;
        const int n = a.n[p];
        T* M = A + a.a_off[p];
        int* info = a.info + p;
        for (int c0 = 0; c0 < n; c0 += nb) {
            const int b = min(nb, n - c0);
            T* L = M + c0 + (i64)c0 * n;
            persist_factor_block<T>(b, L, n, info, c0);
            __syncthreads();  // the block and info are written (same CU: visible)
            if (*info != 0) break;
            const int m = n - c0 - b;
            if (m <= 0) break;
            persist_panel<T>(m, b, L, n, arena);
            persist_trailing<T, MODE>(m, b, L + b, M + (c0 + b) + (i64)(c0 + b) * n, n, arena);
        }
        if constexpr (SOLVE) {
            __syncthreads();
            persist_solve(a, (const double*)A, B, p, arena);
        }
    }
}

// -------------------------------------------------- refinement helpers
// S_i := float(A_i) (lower triangle), one workgroup row per (matrix, column block)
__global__ void batch_to_float_kernel(BatchArgs a, const double* A, float* S) {
    const int p = blockIdx.y;
    const int n = a.n[p];
    const i64 total = (i64)n * n;
    const double* src = A + a.a_off[p];
    float* dst = S + a.a_off[p];
    for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < total; e += (i64)gridDim.x * 256) {
        const int i = (int)(e % n), j = (int)(e / n);
        if (i >= j) dst[e] = (float)src[e];
    }
}

// R_i := B_i - A_i X_i (A_i symmetric, lower triangle read), FP64.  Work
// items are 64-row blocks of every matrix (prefix sums over the bucket); a
// workgroup streams the 64x64 tiles of its block row of the symmetric matrix
// (tiles above the diagonal are read transposed from the stored lower
// triangle) through LDS, with the right-hand sides in chunks of 8 columns.
// Thread (row r, wave w) accumulates columns w and w + 4 of row r.  Per-matrix
// maxima -- ||A||_inf (row sums of |A|), ||r_c||_inf, ||x_c||_inf,
// ||b_c||_inf -- are reduced per wave and merged with integer atomic max on
// the bit patterns of the non-negative doubles, which is order independent.
// red[p] layout: [0] ||A||_inf, [1 + c] r, [1 + kmax + c] x, [1 + 2 kmax + c] b.
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o);
        v = (u > v || u != u) ? u : v;
    }
    return v;
}

__device__ inline void atomic_max_nonneg(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

__global__ __launch_bounds__(256) void batch_symv_residual_kernel(BatchArgs a, const double* A, const double* X,
                                                                  const double* Bin, double* R, const int* prefix,
                                                                  int cnt, double* red, int kmax) {
    __shared__ double sM[64][65];  // sM[q][p] = Asym(r0 + p, q0 + q)
    __shared__ double sX[64][9];
    const int p = find_item(prefix, cnt, blockIdx.x);
    const int I = blockIdx.x - prefix[p];
    if (a.skip != nullptr && a.skip[p]) return;
    const int n = a.n[p], k = a.k[p];
    const double* Ap = A + a.a_off[p];
    const double* Xp = X + a.b_off[p];
    const double* Bp = Bin + a.b_off[p];
    double* Rp = R + a.b_off[p];
    double* rd = red + (i64)p * (1 + 3 * kmax);
    const int t = threadIdx.x, r = t & 63, w = t >> 6;
    const int r0 = I * 64;
    const int nT = (n + 63) / 64;
    const int row = r0 + r;
    const bool in = row < n;
    for (int c0 = 0; c0 < k; c0 += 8) {
        const int kc = min(8, k - c0);
        double acc0 = 0.0, acc1 = 0.0, asum = 0.0;
        for (int J = 0; J < nT; ++J) {
            const int q0 = J * 64;
            __syncthreads();
            {  // all 16 + 2 loads of a thread in flight (clamped addresses), then LDS
                const int x = t & 63, y0 = t >> 6;
                double g[16], gx[2];
#pragma unroll
                for (int s2 = 0; s2 < 16; ++s2) {
                    const int y = y0 + 4 * s2;
                    bool ok;
                    i64 off;
                    if (J < I) {  // stored tile A(r0 + x, q0 + y): lanes along rows
                        ok = r0 + x < n && q0 + y < n;
                        off = (r0 + x) + (i64)(q0 + y) * n;
                    } else if (J > I) {  // Asym(r0 + y, q0 + x) = A(q0 + x, r0 + y)
                        ok = q0 + x < n && r0 + y < n;
                        off = (q0 + x) + (i64)(r0 + y) * n;
                    } else {  // diagonal tile, lower part (mirrored below)
                        ok = r0 + x < n && x >= y;
                        off = (r0 + x) + (i64)(r0 + y) * n;
                    }
                    const double v = Ap[ok ? off : 0];
                    g[s2] = ok ? v : 0.0;
                }
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) {
                    const int c = y0 + 4 * s2;
                    const bool ok = q0 + x < n && c < kc;
                    const double v = Xp[ok ? (q0 + x) + (i64)(c0 + c) * n : 0];
                    gx[s2] = ok ? v : 0.0;
                }
#pragma unroll
                for (int s2 = 0; s2 < 16; ++s2) {
                    const int y = y0 + 4 * s2;
                    if (J < I) {
                        sM[y][x] = g[s2];
                    } else if (J > I) {
                        sM[x][y] = g[s2];
                    } else if (x >= y) {
                        sM[y][x] = g[s2];
                        sM[x][y] = g[s2];
                    }
                }
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) sX[x][y0 + 4 * s2] = gx[s2];
            }
            __syncthreads();
#pragma unroll 8
            for (int q = 0; q < 64; ++q) {
                const double m = sM[q][r];
                acc0 += m * sX[q][w];
                acc1 += m * sX[q][w + 4];
                asum += fabs(m);
            }
        }
        if (c0 == 0 && w == 0) {
            const double am = wave_max(in ? asum : 0.0);
            if (r == 0) atomic_max_nonneg(&rd[0], am);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = w + 4 * u;
            double rv = 0.0, xv = 0.0, bv = 0.0;
            if (in && c < kc) {
                const i64 e = row + (i64)(c0 + c) * n;
                bv = Bp[e];
                xv = Xp[e];
                rv = bv - (u == 0 ? acc0 : acc1);
                Rp[e] = rv;
            }
            const double rm = wave_max(fabs(rv)), xm = wave_max(fabs(xv)), bm = wave_max(fabs(bv));
            if (r == 0 && c < kc) {
                atomic_max_nonneg(&rd[1 + c0 + c], rm);
                atomic_max_nonneg(&rd[1 + kmax + c0 + c], xm);
                atomic_max_nonneg(&rd[1 + 2 * kmax + c0 + c], bm);
            }
        }
    }
}

// eta_i = max_c ||r_c||_inf / (||A_i||_inf ||x_c||_inf + ||b_c||_inf) (a NaN
// or an infinite ratio gives +inf), rscale_i(c) = ||r_c||_inf.
__global__ void batch_eta_kernel(BatchArgs a, const double* red, double* eta, double* rscale, int kmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.count) return;
    if (a.skip != nullptr && a.skip[p]) return;
    const double* rd = red + (i64)p * (1 + 3 * kmax);
    const double anorm = rd[0];
    double e = 0.0;
    for (int c = 0; c < a.k[p]; ++c) {
        const double rm = rd[1 + c], xm = rd[1 + kmax + c], bm = rd[1 + 2 * kmax + c];
        e = fmax(e, column_backward_error(rm, anorm, xm, bm));
        if (rscale != nullptr) rscale[(i64)p * kmax + c] = rm;
    }
    eta[p] = e;
}

__global__ void batch_fill_kernel(double* p, double v, i64 count) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (i64)gridDim.x * blockDim.x) p[i] = v;
}

// refinement bookkeeping: mark converged matrices, scale residuals to float
__global__ void batch_ir_prepare_kernel(BatchArgs a, const double* eta, double tol, const double* R,
                                        const double* rscale, float* Rf, int kmax, int* iters) {
    const int p = blockIdx.x;
    if (a.skip[p]) return;
    if (eta[p] <= tol) {  // uniform across the workgroup
        if (threadIdx.x == 0) a.skip[p] = 1;
        return;
    }
    if (threadIdx.x == 0) iters[p] += 1;
    const int n = a.n[p], k = a.k[p];
    for (int c = 0; c < k; ++c) {
        const double sc = rscale[(i64)p * kmax + c];
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const i64 e = a.b_off[p] + i + (i64)c * n;
            Rf[e] = (float)(sc > 0.0 ? R[e] / sc : 0.0);
        }
    }
}

__global__ void batch_ir_update_kernel(BatchArgs a, const float* D, const double* rscale, double* X, int kmax) {
    const int p = blockIdx.x;
    if (a.skip[p]) return;
    const int n = a.n[p], k = a.k[p];
    for (int c = 0; c < k; ++c) {
        const double sc = rscale[(i64)p * kmax + c];
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const i64 e = a.b_off[p] + i + (i64)c * n;
            X[e] += sc * (double)D[e];
        }
    }
}

// dst_i := src_i (FP64), fdst_i := float(src_i), or dst_i := 0 when src is null
__global__ void batch_copy_kernel(BatchArgs a, const double* src, double* dst, float* fdst) {
    const int p = blockIdx.x;
    const int n = a.n[p], k = a.k[p];
    const i64 o = a.b_off[p];
    for (i64 e = threadIdx.x; e < (i64)n * k; e += blockDim.x) {
        if (src == nullptr) dst[o + e] = 0.0;
        else if (fdst != nullptr) fdst[o + e] = (float)src[o + e];
        else dst[o + e] = src[o + e];
    }
}

// ------------------------------------------------------------ launchers
hipError_t batch_potrf(const BatchArgs& a, void* A, bool f32, int step, int nb, int cnt, hipStream_t s) {
    if (cnt <= 0) return hipSuccess;
    if (f32) hipLaunchKernelGGL(batch_potrf_kernel<float>, dim3(cnt), dim3(512), 0, s, a, (float*)A, step, nb);
    else hipLaunchKernelGGL(batch_potrf_kernel<double>, dim3(cnt), dim3(512), 0, s, a, (double*)A, step, nb);
    return hipGetLastError();
}

hipError_t batch_trsm(const BatchArgs& a, void* A, bool f32, int step, int nb, const int* prefix, int cnt, int wgs,
                      hipStream_t s) {
    if (wgs <= 0) return hipSuccess;
    if (f32)
        hipLaunchKernelGGL(batch_trsm_kernel<float>, dim3(wgs), dim3(256), 0, s, a, (float*)A, step, nb, prefix, cnt);
    else
        hipLaunchKernelGGL(batch_trsm_kernel<double>, dim3(wgs), dim3(256), 0, s, a, (double*)A, step, nb, prefix,
                           cnt);
    return hipGetLastError();
}

hipError_t batch_syrk(const BatchArgs& a, void* A, int mode, int step, int nb, const int* prefix, int cnt, int wgs,
                      hipStream_t s) {
    if (wgs <= 0) return hipSuccess;
    if (mode == 0)
        hipLaunchKernelGGL((batch_syrk_kernel<double, 0>), dim3(wgs), dim3(256), 0, s, a, (double*)A, step, nb, prefix,
                           cnt);
    else if (mode == 1)
        hipLaunchKernelGGL((batch_syrk_kernel<float, 1>), dim3(wgs), dim3(256), 0, s, a, (float*)A, step, nb, prefix,
                           cnt);
    else
        hipLaunchKernelGGL((batch_syrk_kernel<float, 2>), dim3(wgs), dim3(256), 0, s, a, (float*)A, step, nb, prefix,
                           cnt);
    return hipGetLastError();
}

hipError_t batch_persistent(const BatchArgs& a, void* F, double* B, int mode, int nb, int* counter, int grid,
                            hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    if (nb < 1 || nb > kPersistMaxNb || grid < 1) return hipErrorInvalidValue;
    const dim3 g((unsigned)std::min(grid, a.count));
    if (mode == 0)
        hipLaunchKernelGGL((batch_persistent_kernel<double, 0, true>), g, dim3(256), 0, s, a, (double*)F, B, nb,
                           counter);
    else if (mode == 1)
        hipLaunchKernelGGL((batch_persistent_kernel<float, 1, false>), g, dim3(256), 0, s, a, (float*)F, B, nb,
                           counter);
    else
        hipLaunchKernelGGL((batch_persistent_kernel<float, 2, false>), g, dim3(256), 0, s, a, (float*)F, B, nb,
                           counter);
    return hipGetLastError();
}

hipError_t batch_solve(const BatchArgs& a, const void* L, void* B, bool f32, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    if (f32)
        hipLaunchKernelGGL(batch_solve_kernel<float>, dim3(a.count), dim3(256), 0, s, a, (const float*)L, (float*)B);
    else
        hipLaunchKernelGGL(batch_solve_kernel<double>, dim3(a.count), dim3(256), 0, s, a, (const double*)L,
                           (double*)B);
    return hipGetLastError();
}

hipError_t batch_to_float(const BatchArgs& a, const double* A, float* S, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_to_float_kernel, dim3(16, a.count), dim3(256), 0, s, a, A, S);
    return hipGetLastError();
}

hipError_t batch_residual(const BatchArgs& a, const double* A, const double* X, const double* B, double* R,
                          const int* prefix, int wgs, double* red, double* eta, double* rscale, int kmax,
                          hipStream_t s) {
    if (a.count <= 0 || wgs <= 0) return hipSuccess;
    const i64 nred = (i64)a.count * (1 + 3 * kmax);
    hipLaunchKernelGGL(batch_fill_kernel, dim3((unsigned)std::min<i64>(ceil_div(nred, 256), 1024)), dim3(256), 0, s,
                       red, 0.0, nred);
    hipLaunchKernelGGL(batch_symv_residual_kernel, dim3((unsigned)wgs), dim3(256), 0, s, a, A, X, B, R, prefix,
                       a.count, red, kmax);
    hipLaunchKernelGGL(batch_eta_kernel, dim3((unsigned)ceil_div(a.count, 256)), dim3(256), 0, s, a, red, eta, rscale,
                       kmax);
    return hipGetLastError();
}

hipError_t batch_ir_prepare(const BatchArgs& a, const double* eta, double tol, const double* R, const double* rscale,
                            float* Rf, int kmax, int* iters, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_ir_prepare_kernel, dim3(a.count), dim3(256), 0, s, a, eta, tol, R, rscale, Rf, kmax,
                       iters);
    return hipGetLastError();
}

hipError_t batch_ir_update(const BatchArgs& a, const float* D, const double* rscale, double* X, int kmax,
                           hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_ir_update_kernel, dim3(a.count), dim3(256), 0, s, a, D, rscale, X, kmax);
    return hipGetLastError();
}

hipError_t batch_copy_rhs(const BatchArgs& a, const double* src, double* dst, float* fdst, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_copy_kernel, dim3(a.count), dim3(256), 0, s, a, src, dst, fdst);
    return hipGetLastError();
}

}  // namespace elemamd
