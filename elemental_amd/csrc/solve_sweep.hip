// Block-inverse sweeps: the float correction solves of iterative refinement,
//   forward  y := inv(L) r      backward  x := inv(L^T) y
// with the low-precision Cholesky factor L: the two triangular solves of
// cholesky::SolveAfter (include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75)
// arranged so that the sequential dependence between row blocks is short.
//
// Rows are cut into NB = 512-row blocks b = 0 .. T-1.  With the inverse
// X_b = inv(L_bb) of every diagonal block and the look-ahead products
// P(i)_b = -X_b L_{b,b-i} (i = 1 .. d) computed once per factor, the forward
// sweep is
//     y_b = X_b (r_b - sum_{j < b-d} L_bj y_j) + sum_{i=1..d} P(i)_b y_{b-i}.
// Only the last product waits for the block solved just before, so one sweep
// step costs one 32 x NB x k product per workgroup plus a flag hand-over;
// the older blocks are folded in with d steps of slack.  The backward sweep
// is the same with U = L^T (the strict upper triangle of the factor's
// storage holds the mirror of L, written by the preparation), X_b^T and
// Q(i)_b = -X_b^T U_{b,b+i}, over the blocks in reverse order.
//
// Workgroup (s, rt) owns rows [32 rt, 32 rt + 32) of the block of step s
// and all (<= 32) right-hand sides; its four waves take 16 rows x one half
// of each product's K range, and their partial sums are added at the two
// publication points:
//   1. u = sum over source steps s' < s - d of M(rows, block(s')) y_{s'}
//      (the A operand straight from the factor into registers, y_{s'}
//      straight into LDS);
//   2. z = r - u, published; then the block's other row tiles are awaited;
//   3. y = X z + sum_i P(i) y_{s-i}, published.
// A product's A operand (32 rows x NB, 64 KB) is loaded whole into a second
// set of registers while the previous product multiplies.  A step then
// costs the hand-over latencies (flag poll, the block vector's round trip,
// the publication) plus one product's MFMAs (1.7 us for k = 32).
// Solved blocks, z and the flags move through device-coherent loads and
// stores (solve_dev.hpp); workgroups take tickets in launch order, so every
// wait is on a workgroup that is running or done (the row tiles of one step
// need to be resident together), and every wait is bounded.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "solve_dev.hpp"

namespace elemamd {

constexpr int kSwNB = 512;          // sweep block rows
constexpr int kSwR = kSwNB / 64;    // 64-row diagonal pieces per block (preparation)
constexpr int kSwTR = 32;           // rows per sweep workgroup
constexpr int kSwRT = kSwNB / kSwTR;  // sweep workgroups (row tiles) per block
constexpr int kSwKH = kSwNB / 2;    // K range of one wave in a product
constexpr int kSwLook = 3;          // look-ahead products d
constexpr int kSwKP = 32;           // right-hand sides per sweep (row stride of the block vectors)
constexpr long kSwSpinLimit = 1L << 22;           // polls before a wait is declared lost (seconds)

// Offset of element (a, c) of an NB x NB tile in the fragment order of the
// sweep's A operand: 32-row tile, wave (16-row half + 2 x K half), then per
// lane 4 consecutive K steps as one float4 (MFMA 16x16x4 f32: lane m + 16 g
// holds A[m][4 j + g]).
__host__ __device__ inline i64 sweep_frag(int a, int c) {
    const int rt = a >> 5, w = ((a >> 4) & 1) + 2 * (c / kSwKH), m = a & 15;
    const int j = (c % kSwKH) >> 2, g = c & 3;
    return ((((i64)(rt * 4 + w) * (kSwKH / 16) + (j >> 2)) * 64 + (m + 16 * g)) * 4 + (j & 3));
}

// acc[nt] += sum_k sA[16 w + m][k] sB[k][16 nt + m], k = 0 .. 63 (one wave's
// 16 x 64 slice of a 64 x 64 x 64 product staged in LDS).  The operands are
// read kD K-steps ahead of their MFMAs (see sweep_kernel's product).
__device__ inline void lds_tile_product(f4_t (&acc)[4], const float (*sA)[65], const float (*sB)[65], int w, int m,
                                        int g) {
    constexpr int kD = 2;
    float qa[kD + 1], qb[kD + 1][4];
#pragma unroll
    for (int ks = 0; ks < kD; ++ks) {
        qa[ks] = sA[16 * w + m][4 * ks + g];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) qb[ks][nt] = sB[4 * ks + g][16 * nt + m];
    }
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
        if (ks + kD < 16) {
            const int r = (ks + kD) % (kD + 1);
            qa[r] = sA[16 * w + m][4 * (ks + kD) + g];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) qb[r][nt] = sB[4 * (ks + kD) + g][16 * nt + m];
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) mfma_acc(acc[nt], qa[ks % (kD + 1)], qb[ks % (kD + 1)][nt]);
    }
    mfma_retire();
}

// ------------------------------------------------------------- preparation
// Strict upper triangle of S := mirror of the lower (S(i, j) = S(j, i), i < j).
__global__ __launch_bounds__(256) void sweep_mirror_kernel(int n, float* __restrict__ S, i64 lds) {
    __shared__ float sT[64][65];
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (ti > tj) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) {  // sT[c][r] = S(tj*64 + r, ti*64 + c), lanes along r
        const int c = w + 4 * q;
        const i64 row = (i64)tj * 64 + lane, col = (i64)ti * 64 + c;
        sT[c][lane] = (row < n && col < n) ? S[row + col * lds] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {  // S(ti*64 + c, tj*64 + r) = sT[c][r], lanes along c
        const int r = w + 4 * q;
        const i64 row = (i64)ti * 64 + lane, col = (i64)tj * 64 + r;
        if (row < n && col < n && row < col) S[row + col * lds] = sT[lane][r];
    }
}

// X_b = inv(L_bb) (column-major NB x NB, zero above the diagonal and on rows
// or columns past n), column slab c of it per workgroup, from the inverses of
// the 64-row diagonal blocks:  X_cc = Dinv_c,
//     X_ic = -Dinv_i sum_{c <= l < i} L_il X_lc   (i > c).
__global__ __launch_bounds__(256) void sweep_xinv_kernel(int n, int b0, const float* __restrict__ S, i64 lds,
                                                         const float* __restrict__ Dinv, float* __restrict__ Xp) {
    __shared__ float sA[64][65];  // [row][k]
    __shared__ float sB[64][65];  // [k][col]
    const int c = blockIdx.x, b = b0 + blockIdx.y;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m = lane & 15, g = lane >> 4;
    const int nT64 = (int)ceil_div(n, 64);
    float* X = Xp + (i64)b * kSwNB * kSwNB;
    const i64 base = (i64)b * kSwNB;
    auto dinv = [&](int i, int r, int q) -> float {
        const int I = b * kSwR + i;
        return I < nT64 ? Dinv[(i64)I * 4096 + r + q * 64] : 0.0f;
    };
    for (int idx = t; idx < 64 * c * 64; idx += 256) {
        const int r = idx % (64 * c), cc = idx / (64 * c);
        X[r + (i64)(64 * c + cc) * kSwNB] = 0.0f;
    }
    for (int idx = t; idx < 4096; idx += 256) {
        const int r = idx & 63, q = idx >> 6;
        X[(64 * c + r) + (i64)(64 * c + q) * kSwNB] = dinv(c, r, q);
    }
    __syncthreads();
    // The products (i, l), i = c+1 .. R-1, l = c .. i-1, in that order; the
    // next one's operands load into registers while the current one
    // multiplies.  X(l, c) for l < i is final before row i starts (stored,
    // then a barrier), and the load of X(i-1, c) is issued in row i.
    float ga[16], gb[16];
    auto gload = [&](int i, int l) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int kk = w + 4 * q;
            const i64 row = base + 64 * i + lane, col = base + 64 * l + kk;
            ga[q] = (row < n && col < n) ? S[row + col * lds] : 0.0f;  // L(64i + lane, 64l + kk)
            gb[q] = X[(64 * l + lane) + (i64)(64 * c + kk) * kSwNB];     // X(64l + lane, 64c + kk)
        }
    };
    if (c + 1 < kSwR) gload(c + 1, c);
    for (int i = c + 1; i < kSwR; ++i) {
        f4_t acc[4] = {};
        for (int l = c; l < i; ++l) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int kk = w + 4 * q;
                sA[lane][kk] = ga[q];
                sB[lane][kk] = gb[q];
            }
            __syncthreads();
            // the next product: (i, l+1), else (i+1, c); after the last one
            // the current operands again (in bounds, never stored)
            int ni = i, nl = l + 1;
            if (nl >= i) ni = i + 1, nl = c;
            if (ni >= kSwR) ni = i, nl = l;
            gload(ni, nl);
            lds_tile_product(acc, sA, sB, w, m, g);
            __syncthreads();
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) sB[16 * w + 4 * g + v][16 * nt + m] = acc[nt][v];
        for (int idx = t; idx < 4096; idx += 256) {
            const int r = idx & 63, q = idx >> 6;
            sA[r][q] = dinv(i, r, q);
        }
        __syncthreads();
        f4_t x[4] = {};
        lds_tile_product(x, sA, sB, w, m, g);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                X[(64 * i + 16 * w + 4 * g + v) + (i64)(64 * c + 16 * nt + m) * kSwNB] = -x[nt][v];
        __syncthreads();  // this slab's new tile is read by the next rows (same workgroup)
    }
}

// The sweep operands in fragment order: tiles + ((dir*(d+1) + i)*T + s)*NB*NB
// for direction dir (0 forward, 1 backward), step s, term i (0: the diagonal
// inverse, i >= 1: the look-ahead product with the block i steps back).
// One 64 x 64 sub-tile (ra, rc) per workgroup.
__global__ __launch_bounds__(256) void sweep_tiles_kernel(int n, int b0, int nbk, const float* __restrict__ S,
                                                          i64 lds, const float* __restrict__ Xp,
                                                          float* __restrict__ tiles) {
    __shared__ float sA[64][65];  // [a][q]
    __shared__ float sB[64][65];  // [q][c]
    const int T = (int)ceil_div(n, kSwNB);
    const int ra = blockIdx.x / kSwR, rc = blockIdx.x % kSwR;
    // blocks b0 .. b0+nbk-1 (block b is step b forward, step T-1-b backward)
    const int i = blockIdx.y / nbk, b = b0 + blockIdx.y % nbk, dir = blockIdx.z;
    const int s = dir ? T - 1 - b : b;
    if (i > s) return;  // term i of step s exists for s >= i only
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m = lane & 15, g = lane >> 4;
    const int bb = dir ? b + i : b - i;
    const float* X = Xp + (i64)b * kSwNB * kSwNB;
    float* out = tiles + (((i64)dir * (kSwLook + 1) + i) * T + s) * kSwNB * kSwNB;
    if (i == 0) {
        for (int idx = t; idx < 4096; idx += 256) {
            const int p = idx & 63, q = idx >> 6;  // p runs along the contiguous index of X
            const int a = 64 * ra + (dir ? q : p), c = 64 * rc + (dir ? p : q);
            out[sweep_frag(a, c)] = dir ? X[c + (i64)a * kSwNB] : X[a + (i64)c * kSwNB];
        }
        return;
    }
    // X_b lower triangular: forward X(a, q) needs q < 64(ra+1); backward X(q, a) needs q >= 64 ra.
    // The next 64-wide chunk of both operands loads into registers while
    // the current one multiplies.
    const int q_lo = dir ? 64 * ra : 0, q_hi = dir ? kSwNB : 64 * (ra + 1);
    float ga[16], gb[16];
    auto gload = [&](int q0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = w + 4 * u;
            if (!dir) {
                ga[u] = X[(64 * ra + lane) + (i64)(q0 + e) * kSwNB];  // X(a, q)
                const i64 row = (i64)b * kSwNB + q0 + lane, col = (i64)bb * kSwNB + 64 * rc + e;
                gb[u] = (row < n && col < n) ? S[row + col * lds] : 0.0f;  // L(b*NB + q, bb*NB + c)
            } else {
                ga[u] = X[(q0 + lane) + (i64)(64 * ra + e) * kSwNB];  // X(q, a)
                const i64 row = (i64)bb * kSwNB + 64 * rc + lane, col = (i64)b * kSwNB + q0 + e;
                gb[u] = (row < n && col < n) ? S[row + col * lds] : 0.0f;  // U(b*NB + q, bb*NB + c) = L(bb*NB + c, b*NB + q)
            }
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = w + 4 * u;
            if (!dir) {
                sA[lane][e] = ga[u];
                sB[lane][e] = gb[u];
            } else {
                sA[e][lane] = ga[u];
                sB[e][lane] = gb[u];
            }
        }
    };
    f4_t acc[4] = {};
    gload(q_lo);
    for (int q0 = q_lo; q0 < q_hi; q0 += 64) {
        sstore();
        __syncthreads();
        if (q0 + 64 < q_hi) gload(q0 + 64);
        lds_tile_product(acc, sA, sB, w, m, g);
        __syncthreads();
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            out[sweep_frag(64 * ra + 16 * w + 4 * g + v, 64 * rc + 16 * nt + m)] = -acc[nt][v];
}

// ------------------------------------------------------------------- sweep
// rhs: forward, column-major n x kc (ld ldr); backward, the forward sweep's
// row-major output.  Y: row-major output (T*NB x KP), Z: row-major scratch.
// out (backward only): column-major destination of the solution.  NT
// 16-column tiles of right-hand sides (kc <= 16 NT).
template <int NT, bool BWD>
__global__ __launch_bounds__(256, 2) void sweep_kernel(int n, int kc, const float* __restrict__ S, i64 lds,
                                                    const float* __restrict__ tiles, const float* __restrict__ rhs,
                                                    i64 ldr, float* __restrict__ Y, float* __restrict__ Z,
                                                    float* __restrict__ out, i64 ldo, int* __restrict__ ctrl,
                                                    int* __restrict__ err) {
    // A block vector (NB rows x 16 NT columns) is staged in LDS whole, one
    // coherent round trip per product, straight from memory (global_load_lds,
    // no registers).  Row r, column c sits at r VP + (c ^ 16 ((r >> 1) & 1))
    // for NT = 2 (the 16-column halves swapped on every other row pair), so
    // the B-operand reads of a wave (rows 4j + g, columns 16 nt + m) fall in
    // 64 distinct banks.  35 or 70 KB of LDS and at most 256 registers per
    // lane: two workgroups share a CU, so every workgroup of a sweep of up
    // to 16384 rows is resident from the start.
    constexpr int VP = 16 * NT;
    constexpr int kVecC = kSwNB * VP / 4 / 256;  // 16-byte pieces per thread for one block vector
    constexpr int kJ = kSwKH / 4;                // K steps of a wave per product
    __shared__ float sV[1][kSwNB * VP];
    __shared__ float sRed[2][16][VP + 1];        // the upper K half's partial sums
    __shared__ int s_ord;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, m = lane & 15, g = lane >> 4;
    const int wr = w & 1, kh = w >> 1;  // 16-row half, K half
    const int T = (int)ceil_div(n, kSwNB);
    if (t == 0) s_ord = atomicAdd(&ctrl[0], 1);
    __syncthreads();
    const int ord = s_ord;
    const int s = ord / kSwRT, rt = ord - s * kSwRT;
    int* yflag = ctrl + 1;                // [step][row tile]: solved rows published
    int* zflag = ctrl + 1 + T * kSwRT;    // [step][row tile]: z rows published
    auto blk = [&](int st) { return BWD ? T - 1 - st : st; };
    const int b = blk(s);
    const i64 row0 = (i64)b * kSwNB + kSwTR * rt;
    const i64 arow = row0 + 16 * wr + m < n ? row0 + 16 * wr + m : (i64)n - 1;  // A-operand row (clamped in bounds)
    const i64 orow = row0 + 16 * wr + 4 * g;  // output rows orow + v of this lane

    // The row tiles of one step, then the workgroup.  A wait that expires
    // records an error in *err (kept until the caller clears it); every later
    // wait sees it and gives up at once, so a lost flag costs one expiry, not
    // one per wait.
    auto wait_all = [&](const int* f) {
        if (w == 0) {
            long spins = 0;
            for (;;) {
                const int v = (lane < kSwRT) ? coherent_load(&f[lane]) : (lane == kSwRT ? coherent_load(err) == 0 : 1);
                if (__ballot(v == 0) == 0ull) break;                      // all flags up
                if (__ballot(lane == kSwRT && v == 0) != 0ull) break;     // an earlier wait expired
                if (++spins > kSwSpinLimit) {
                    if (lane == 0) atomicOr(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
    };
    auto vpos = [](int r, int c) { return NT == 2 ? r * VP + (c ^ (((r >> 1) & 1) << 4)) : r * VP + c; };
    // A published block vector (its first 16 NT columns) into sV: each
    // wave instruction fills 1 KB of LDS, lane L the 16 bytes at 16 L, with
    // the row and columns whose place that is; device-coherent (sc1) loads
    // like coherent_load.  The caller's barrier must come first (no wave
    // still reads the buffer), then vec_wait().
    auto vec_fill = [&](const float* src, int bb) __attribute__((always_inline)) {
        const float* p = src + (i64)bb * kSwNB * kSwKP;
#pragma unroll
        for (int q = 0; q < kVecC; ++q) {
            const int chunk = w * kVecC + q;  // wave-uniform
            const int pos = (chunk * 64 + lane) * 4;
            const int r = pos / VP, cl = pos % VP;
            const int c = NT == 2 ? cl ^ (((r >> 1) & 1) << 4) : cl;
            __builtin_amdgcn_global_load_lds(p + (i64)r * kSwKP + c,
                                             (__attribute__((address_space(3))) void*)(sV[0] + chunk * 256), 16, 0,
                                             16 /* sc1 */);
        }
    };
    auto vec_wait = [&]() __attribute__((always_inline)) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    };
    // The A operand of one product, this wave's rows and K half:
    // M(arow, bb*NB + KH kh + 4j + g), L below the diagonal, its mirror
    // above.  Rows and columns past n read clamped (finite) entries: they
    // meet zero rows of the block vectors or produce output rows that are
    // never kept.
    auto a_from_factor = [&](float (&a)[kJ], int bb) __attribute__((always_inline)) {
        const i64 c0 = (i64)bb * kSwNB + kSwKH * kh + g;
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const i64 col = c0 + 4 * j < n ? c0 + 4 * j : (i64)n - 1;
            a[j] = S[arow + col * lds];
        }
    };
    auto a_from_tile = [&](float (&a)[kJ], int i) __attribute__((always_inline)) {
        const float4* p = reinterpret_cast<const float4*>(tiles + ((i64)i * T + s) * kSwNB * kSwNB) +
                          (i64)(rt * 4 + w) * (kSwKH / 16) * 64 + lane;
#pragma unroll
        for (int jq = 0; jq < kJ / 4; ++jq) {
            const float4 v = p[(i64)jq * 64];
            a[4 * jq] = v.x;
            a[4 * jq + 1] = v.y;
            a[4 * jq + 2] = v.z;
            a[4 * jq + 3] = v.w;
        }
    };
    // acc += a x (this wave's K half of the vector in sV[bf]).  The MFMA
    // statements are volatile, so the compiler keeps the LDS reads where
    // the source puts them: they run kD K-steps ahead of their MFMAs through
    // a ring of registers instead of each MFMA waiting out the LDS latency
    // of its own operand.
    auto product = [&](const float (&a)[kJ], int bf, f4_t (&acc)[NT]) __attribute__((always_inline)) {
        constexpr int kD = 6 / NT;
        const float* v = sV[bf];
        const int r0 = kSwKH * kh + g;
        float bq[kD + 1][NT];
#pragma unroll
        for (int j = 0; j < kD; ++j)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bq[j][nt] = v[vpos(r0 + 4 * j, 16 * nt + m)];
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            if (j + kD < kJ)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bq[(j + kD) % (kD + 1)][nt] = v[vpos(r0 + 4 * (j + kD), 16 * nt + m)];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) mfma_acc(acc[nt], a[j], bq[j % (kD + 1)][nt]);
        }
        mfma_retire();
    };
    // the upper K half's partial sums into the lower half's waves
    auto reduce = [&](f4_t (&acc)[NT]) __attribute__((always_inline)) {
        if (kh == 1)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int v = 0; v < 4; ++v) sRed[wr][4 * g + v][16 * nt + m] = acc[nt][v];
        __syncthreads();
        if (kh == 0)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[nt][v] += sRed[wr][4 * g + v][16 * nt + m];
    };

    // own right-hand sides, in the accumulator layout (row orow + v, column 16 nt + m)
    float r[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const i64 row = orow + v;
            const int col = 16 * nt + m;
            if (BWD) r[nt][v] = rhs[row * kSwKP + col];
            else r[nt][v] = (row < n && col < kc) ? rhs[row + col * ldr] : 0.0f;
        }

    // The products, in order: the older blocks p < nU (u), then X z (p ==
    // nU) and the look-ahead terms P(i), i = i_top .. 1 (y).
    const int nU = s > kSwLook ? s - kSwLook : 0;
    const int i_top = s < kSwLook ? s : kSwLook;
    const int np = nU + 1 + i_top;
    auto term = [&](int p) { return p == nU ? 0 : i_top - (p - nU - 1); };
    auto load_a = [&](float (&a)[kJ], int p) __attribute__((always_inline)) {
        if (p < nU) a_from_factor(a, blk(p));
        else a_from_tile(a, term(p));
    };
    f4_t acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = (f4_t){0.0f, 0.0f, 0.0f, 0.0f};
    // product p with its A operand in a; the next one's loads into nxt
    auto step = [&](const float (&a)[kJ], float (&nxt)[kJ], int p) __attribute__((always_inline)) {
        if (p == nU) {
            // z = r - u (rows past n are zero), published; the step's other
            // row tiles awaited
            reduce(acc);
            if (kh == 0)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const i64 row = orow + v;
                        coherent_store(&Z[row * kSwKP + 16 * nt + m], row < n ? r[nt][v] - acc[nt][v] : 0.0f);
                    }
            publish_flag(&zflag[s * kSwRT + rt], t);
            // from here on the step's own products: the critical path, ahead
            // of a workgroup sharing the SIMDs with older blocks to fold in
            __builtin_amdgcn_s_setprio(3);
            wait_all(zflag + s * kSwRT);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = (f4_t){0.0f, 0.0f, 0.0f, 0.0f};
            vec_fill(Z, b);  // wait_all's barrier: the buffer is free
        } else {
            const int src_step = p < nU ? p : s - term(p);
            wait_all(yflag + src_step * kSwRT);
            vec_fill(Y, blk(src_step));
        }
        vec_wait();
        if (p + 1 < np) load_a(nxt, p + 1);
        product(a, 0, acc);
    };
    float a0[kJ], a1[kJ];
    load_a(a0, 0);
    for (int p = 0;;) {
        step(a0, a1, p);
        if (++p == np) break;
        step(a1, a0, p);
        if (++p == np) break;
    }

    // y, published
    reduce(acc);
    if (kh == 0)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const i64 row = orow + v;
                const int col = 16 * nt + m;
                const float yv = row < n ? acc[nt][v] : 0.0f;
                coherent_store(&Y[row * kSwKP + col], yv);
                if (BWD && row < n && col < kc) out[row + col * ldo] = yv;
            }
    publish_flag(&yflag[s * kSwRT + rt], t);
}

// ------------------------------------------------------------------ host
size_t sweep_tiles_count(int n) {
    const i64 T = ceil_div(n, kSwNB);
    return (size_t)2 * (kSwLook + 1) * T * kSwNB * kSwNB;
}
size_t sweep_xinv_count(int n) { return (size_t)ceil_div(n, kSwNB) * kSwNB * kSwNB; }

// Work layout: Yf, Yb, Z (rows x KP floats each), then the error word
// (64 ints reserved), then the forward and backward tickets and flags.
static size_t sweep_ctrl_ints(int n) { return 1 + 2 * (size_t)ceil_div(n, kSwNB) * kSwRT; }
static size_t sweep_vec_bytes(int n) {
    const size_t rows = (size_t)ceil_div(n, kSwNB) * kSwNB;
    return (3 * rows * kSwKP * sizeof(float) + 255) & ~(size_t)255;
}
static int* sweep_err(int n, void* work) { return (int*)((char*)work + sweep_vec_bytes(n)); }

size_t sweep_work_bytes(int n) { return sweep_vec_bytes(n) + (64 + 2 * sweep_ctrl_ints(n)) * sizeof(int) + 256; }

hipError_t launch_sweep_clear_error(int n, void* work, hipStream_t s) {
    return launch_fill_int(sweep_err(n, work), 0, 1, s);
}

hipError_t sweep_read_error(int n, void* work, int* out, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(out, sweep_err(n, work), sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t launch_sweep_mirror(int n, float* S, i64 lds, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned nt64 = (unsigned)ceil_div(n, 64);
    hipLaunchKernelGGL(sweep_mirror_kernel, dim3(nt64, nt64), dim3(256), 0, s, n, S, lds);
    return hipGetLastError();
}

int sweep_block_cols() { return kSwNB; }

hipError_t launch_sweep_prepare_blocks(int n, int b0, int b1, const float* S, i64 lds, const float* Dinv, float* Xp,
                                       float* tiles, hipStream_t s) {
    b1 = (int)std::min<i64>(b1, ceil_div(n, kSwNB));
    if (n <= 0 || b0 >= b1) return hipSuccess;
    const int nbk = b1 - b0;
    hipLaunchKernelGGL(sweep_xinv_kernel, dim3(kSwR, (unsigned)nbk), dim3(256), 0, s, n, b0, S, lds, Dinv, Xp);
    hipLaunchKernelGGL(sweep_tiles_kernel, dim3(kSwR * kSwR, (unsigned)((kSwLook + 1) * nbk), 2), dim3(256), 0, s, n,
                       b0, nbk, S, lds, Xp, tiles);
    return hipGetLastError();
}

hipError_t launch_sweep_prepare(int n, const float* S, i64 lds, const float* Dinv, float* Xp, float* tiles,
                                hipStream_t s) {
    return launch_sweep_prepare_blocks(n, 0, (int)ceil_div(n, kSwNB), S, lds, Dinv, Xp, tiles, s);
}

template <int NT>
static void launch_pair(int n, int kc, const float* S, i64 lds, const float* tiles, float* B, i64 ldb, float* Yf,
                        float* Yb, float* Z, int* ctrl, int* err, hipStream_t s) {
    const unsigned T = (unsigned)ceil_div(n, kSwNB);
    const size_t half = (size_t)(kSwLook + 1) * T * kSwNB * kSwNB;
    hipLaunchKernelGGL((sweep_kernel<NT, false>), dim3(T * kSwRT), dim3(256), 0, s, n, kc, S, lds, tiles, B, ldb, Yf,
                       Z, nullptr, 0, ctrl, err);
    hipLaunchKernelGGL((sweep_kernel<NT, true>), dim3(T * kSwRT), dim3(256), 0, s, n, kc, S, lds, tiles + half, Yf,
                       (i64)kSwKP, Yb, Z, B, ldb, ctrl + sweep_ctrl_ints(n), err);
}

hipError_t launch_sweep_solve(int n, int k, const float* S, i64 lds, const float* tiles, float* B, i64 ldb, void* work,
                              hipStream_t s) {
    if (n <= 0 || k <= 0) return hipSuccess;
    const size_t rows = (size_t)ceil_div(n, kSwNB) * kSwNB;
    float* Yf = (float*)work;
    float* Yb = Yf + rows * kSwKP;
    float* Z = Yb + rows * kSwKP;
    int* err = sweep_err(n, work);
    int* ctrl = err + 64;
    for (int c0 = 0; c0 < k; c0 += kSwKP) {
        const int kc = std::min(kSwKP, k - c0);
        hipError_t e = launch_fill_int(ctrl, 0, (int)(2 * sweep_ctrl_ints(n)), s);
        if (e != hipSuccess) return e;
        if (kc <= 16)
            launch_pair<1>(n, kc, S, lds, tiles, B + (i64)c0 * ldb, ldb, Yf, Yb, Z, ctrl, err, s);
        else
            launch_pair<2>(n, kc, S, lds, tiles, B + (i64)c0 * ldb, ldb, Yf, Yb, Z, ctrl, err, s);
    }
    return hipGetLastError();
}

}  // namespace elemamd
