// Persistent triangular solves with the Cholesky factor, one launch per
// direction:
//   forward  B := inv(L) B     (Trsm(LEFT, LOWER, NORMAL,  NON_UNIT) -> dtrsm_('L','L','N','N'))
//   backward B := inv(L^T) B   (Trsm(LEFT, LOWER, ADJOINT, NON_UNIT) -> dtrsm_('L','L','T','N'))
// as used by cholesky::SolveAfter (include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75).
//
// The rows are cut into 64-row blocks.  Each workgroup owns one block row
// and one 16-column slab of the right-hand sides; it takes a ticket from a
// device counter, so it only ever waits for workgroups that started before
// it.  It accumulates the contributions of the already-solved blocks in the
// order they become available (flags in device memory), each 64x64 tile of L
// staged through double-buffered LDS and multiplied with MFMA; then it solves
// its 64x64 diagonal block, publishes the solved block and raises its flag.
// L is read from HBM once per slab.
//
// Diagonal block, two variants:
//   USE_INV = false (the FP64 solves): substitution, one wave per four
//     right-hand sides, lanes = rows, pivots applied as reciprocal products.
//     Backward stable like the reference BLAS loop, and exact on integer
//     problems whose diagonal is a power of two.
//   USE_INV = true (the FP32 correction solves of iterative refinement):
//     a product with the precomputed inverse of the diagonal block (MFMA),
//     which shortens the sequential critical path; refinement absorbs its
//     larger rounding error.
//
// Every wait is bounded: after kSpinLimit polls the workgroup records an error
// in ctrl[1] and continues, so a lost flag gives a wrong answer and an error
// code, never a hang.
#include "common.hpp"
#include "kernels.hpp"
#include "solve_dev.hpp"

namespace elemamd {

// ------------------------------------------------ inverse of diagonal blocks
// Dinv + I*64*64 (column-major, ld 64) := inv(L(I*64.., I*64..)), upper part 0.
// Column c is built row by row: W(i,c) = -(sum_{c<=k<i} L(i,k) W(k,c)) / L(i,i),
// the dot product split over four thread groups and reduced in LDS.
template <typename T>
__global__ __launch_bounds__(256) void trtri_diag64_kernel(int n, int I0, const T* __restrict__ L, i64 ldl,
                                                           T* __restrict__ Dinv) {
    __shared__ T sL[kSB][kSLd];  // sL[i][k]
    __shared__ T sW[kSB][kSLd];  // sW[k][c]
    __shared__ T red[4][kSB];
    const int I = I0 + blockIdx.x, t = threadIdx.x;
    const i64 r0 = (i64)I * kSB;
    const int bs = (int)min<i64>(kSB, n - r0);
    for (int idx = t; idx < kSB * kSB; idx += 256) {
        const int i = idx % kSB, k = idx / kSB;
        sL[i][k] = (i < bs && k < bs && k <= i) ? L[r0 + i + (r0 + k) * ldl] : (i == k ? T(1) : T(0));
        sW[i][k] = T(0);
    }
    __syncthreads();
    const int c = t & 63, part = t >> 6;
    if (part == 0) sW[c][c] = T(1) / sL[c][c];
    __syncthreads();
    for (int i = 1; i < kSB; ++i) {
        T s = T(0);
        if (c < i)
            for (int k = c + part; k < i; k += 4) s += sL[i][k] * sW[k][c];
        red[part][c] = s;
        __syncthreads();
        if (part == 0 && c < i) sW[i][c] = -(red[0][c] + red[1][c] + red[2][c] + red[3][c]) / sL[i][i];
        __syncthreads();
    }
    T* out = Dinv + (i64)I * kSB * kSB;
    for (int idx = t; idx < kSB * kSB; idx += 256) {
        const int i = idx % kSB, k = idx / kSB;
        out[idx] = (i < bs && k < bs) ? sW[i][k] : T(0);
    }
}

// ------------------------------------------------------- persistent solve
// Cross-workgroup data (the solved blocks and their flags) moves only through
// device-coherent (agent-scope) loads and stores, which bypass the per-XCD
// caches that are not coherent across XCDs.  No cache invalidation or
// write-back is ever needed: the writer waits for its own coherent stores to
// be acknowledged (vmcnt) before raising the flag, and the reader only issues
// its coherent loads of a block after it has seen the block's flag.
template <typename T, bool TRANS, bool USE_INV>
__global__ __launch_bounds__(256) void trsm_persist_kernel(int n, int kc, const T* __restrict__ L, i64 ldl,
                                                           T* __restrict__ B, i64 ldb, const T* __restrict__ Dinv,
                                                           int* __restrict__ ctrl, const int* __restrict__ info) {
    typedef typename Acc4<T>::type acc_t;
    __shared__ T sM[2][kSB][kSLd];  // sM[q][p]: contraction index q, output row p
    __shared__ T sY[2][kSB][kSC + 1];
    __shared__ int s_ord, s_cnt;
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nT = (int)ceil_div(n, kSB);
    const int CY = (int)ceil_div(kc, kSC);
    if (t == 0) s_ord = atomicAdd(&ctrl[0], 1);
    __syncthreads();
    const int ord = s_ord;
    const int cy = ord % CY;
    const int I = TRANS ? (nT - 1 - ord / CY) : (ord / CY);
    const i64 r0 = (i64)I * kSB;
    const int bs = (int)min<i64>(kSB, n - r0);
    const int c_base = cy * kSC;
    const int kcl = min(kSC, kc - c_base);
    int* flags = ctrl + 2;
    // contraction blocks in the order they are solved
    const int nJ = TRANS ? (nT - 1 - I) : I;
    auto blk = [&](int pos) { return TRANS ? nT - 1 - pos : pos; };

    // ---- operands that do not depend on other workgroups, fetched up front
    // own right-hand sides in MFMA output order: row p = 16w + lane%16, column mfma_row(lane, v)
    T gb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = mfma_row<T>(lane, v);
        const i64 row = r0 + 16 * w + (lane & 15);
        gb[v] = (row < n && c < kcl) ? B[row + (i64)(c_base + c) * ldb] : T(0);
    }
    // diagonal operand, staged as sM-style [q][p] rows
    T gd[16];
    if constexpr (USE_INV) {
        const T* Di = Dinv + (i64)I * kSB * kSB;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int qq = w + 4 * s;
            gd[s] = !TRANS ? Di[lane + qq * kSB] : Di[qq + lane * kSB];  // M[q][p] = Dinv(p,q) / Dinv(q,p)
        }
    } else {
        // lower-triangular system M(i, k) at [k][i], identity on padded rows.
        // Forward: M = L_II.  Backward: L_II^T x = r is the same sweep on
        // reversed indices, M(a, b) = L_II(63-b, 63-a), lane a holding row 63-a.
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int kk = w + 4 * s;
            const int gi = TRANS ? (kSB - 1 - kk) : lane;
            const int gk = TRANS ? (kSB - 1 - lane) : kk;
            T v = T(0);
            if (gi < bs && gk < bs) {
                if (lane >= kk) v = L[r0 + gi + (r0 + gk) * ldl];
            } else if (lane == kk) {
                v = T(1);
            }
            gd[s] = v;
        }
    }

    // contraction tile of block J (rows q of the solved block, rows p of ours)
    T gm[16], gy[4];
    auto gload = [&](int J) {
        const i64 q0 = (i64)J * kSB;
        if (!TRANS) {  // M[q][p] = L(r0+p, q0+q), lanes along p
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int qq = w + 4 * s;
                const i64 p = r0 + lane;
                gm[s] = (p < n) ? L[p + (q0 + qq) * ldl] : T(0);
            }
        } else {  // M[q][p] = L(q0+q, r0+p), lanes along q
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int pp = w + 4 * s;
                const i64 q = q0 + lane;
                gm[s] = (q < n) ? L[q + (r0 + pp) * ldl] : T(0);
            }
        }
    };
    auto gload_y = [&](int J) {  // a solved block: device-coherent loads
        const i64 q0 = (i64)J * kSB;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int c = w + 4 * s;
            const i64 q = q0 + lane;
            gy[s] = (q < n && c < kcl)
                        ? __hip_atomic_load(&B[q + (i64)(c_base + c) * ldb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : T(0);
        }
    };
    auto sstore = [&](int buf) {
        if (!TRANS) {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[buf][w + 4 * s][lane] = gm[s];
        } else {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[buf][lane][w + 4 * s] = gm[s];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) sY[buf][lane][w + 4 * s] = gy[s];
    };
    // Length of the run of solved blocks starting at position `done` (>= 1);
    // wave 0 polls up to 64 flags per round trip.  Bounded: past kSpinLimit
    // polls the error is recorded in ctrl[1] and the remaining blocks are
    // taken as they are.
    auto ready_run = [&](int done) -> int {
        if (w == 0) {
            long spins = 0;
            int cnt;
            for (;;) {
                const int pos = done + lane;
                int f = 0;
                if (pos < nJ) f = __hip_atomic_load(&flags[blk(pos) * CY + cy], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long m = __ballot(f != 0);
                cnt = (~m == 0ull) ? 64 : (int)__builtin_ctzll(~m);
                if (cnt > 0) break;
                if (++spins > kSpinLimit) {
                    if (lane == 0) atomicOr(&ctrl[1], 1);
                    cnt = nJ - done;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (lane == 0) s_cnt = cnt;
        }
        __syncthreads();
        return s_cnt;
    };

    // D[c][p]: MFMA rows = right-hand sides, columns = output rows; wave w owns rows 16w..16w+15
    acc_t acc = {};
    int buf = 0;
    if (nJ > 0) gload(blk(0));
    int done = 0;
    while (done < nJ) {
        const int cnt = ready_run(done);
        gload_y(blk(done));
        for (int b = 0; b < cnt; ++b) {
            const int pos = done + b;
            sstore(buf);
            // next tiles (clamped reloads keep the body branch-free; a
            // reloaded tile is overwritten before it is staged again)
            gload(blk(min(pos + 1, nJ - 1)));
            gload_y(blk((b + 1 < cnt) ? pos + 1 : pos));
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kSB / 4; ++ks) {
                const int q = 4 * ks + (lane >> 4);
                acc = mfma16(sY[buf][q][lane & 15], sM[buf][q][16 * w + (lane & 15)], acc);
            }
            buf ^= 1;
        }
        done += cnt;
    }
    // A dependent zero-operand MFMA right before the accumulator is read: the
    // reads below then follow an MFMA in the same basic block, where the
    // compiler's wait-state padding for MFMA results is applied correctly
    // (across a loop back edge it was found to be short for FP64 MFMA).
    acc = mfma16(T(0), T(0), acc);
    __syncthreads();
    // R = B_I - acc into sY[0] as [p][c]; diagonal operand into sM[0]
#pragma unroll
    for (int v = 0; v < 4; ++v) sY[0][16 * w + (lane & 15)][mfma_row<T>(lane, v)] = gb[v] - acc[v];
#pragma unroll
    for (int s = 0; s < 16; ++s) sM[0][w + 4 * s][lane] = gd[s];
    __syncthreads();
    if constexpr (USE_INV) {
        // forward Y(p) = sum_q Dinv(p,q) R(q); backward X(p) = sum_q Dinv(q,p) R(q)
        acc_t y = {};
#pragma unroll
        for (int ks = 0; ks < kSB / 4; ++ks) {
            const int q = 4 * ks + (lane >> 4);
            y = mfma16(sY[0][q][lane & 15], sM[0][q][16 * w + (lane & 15)], y);
        }
        y = mfma16(T(0), T(0), y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = mfma_row<T>(lane, v);
            const i64 row = r0 + 16 * w + (lane & 15);
            if (row < n && c < kcl)
                __hip_atomic_store(&B[row + (i64)(c_base + c) * ldb], y[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else {
        if (t < kSB) sY[1][t][0] = T(1) / sM[0][t][t];
        __syncthreads();
        const int myrow = TRANS ? (kSB - 1 - lane) : lane;
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = sY[0][myrow][4 * w + u];
#pragma unroll 16
        for (int kk = 0; kk < kSB; ++kk) {
            const T rk = sY[1][kk][0];
            const T lik = sM[0][kk][lane];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T y = wave_bcast(v[u], kk) * rk;
                if (lane == kk) v[u] = y;
                else if (lane > kk) v[u] -= lik * y;
            }
        }
        const i64 row = r0 + myrow;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = 4 * w + u;
            if (row < n && c < kcl)
                __hip_atomic_store(&B[row + (i64)(c_base + c) * ldb], v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // every wave waits for the acknowledgement of its coherent stores, then
    // one thread raises the flag with a coherent store
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (t == 0) __hip_atomic_store(&flags[I * CY + cy], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
size_t solve_ctrl_ints(int n, int kc) {
    return 2 + (size_t)ceil_div(n, kSB) * ceil_div(kc, kSC);
}

template <typename T>
hipError_t launch_trtri_diag64_range(int n, int I0, int I1, const T* L, i64 ldl, T* Dinv, hipStream_t s) {
    I1 = (int)std::min<i64>(I1, ceil_div(n, kSB));
    if (n <= 0 || I0 >= I1) return hipSuccess;
    hipLaunchKernelGGL(trtri_diag64_kernel<T>, dim3((unsigned)(I1 - I0)), dim3(256), 0, s, n, I0, L, ldl, Dinv);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_trtri_diag64(int n, const T* L, i64 ldl, T* Dinv, hipStream_t s) {
    return launch_trtri_diag64_range<T>(n, 0, (int)ceil_div(n, kSB), L, ldl, Dinv, s);
}

template <typename T>
hipError_t launch_trsm_persist(bool trans, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv,
                               int* ctrl, const int* info, hipStream_t s) {
    if (n <= 0 || kc <= 0) return hipSuccess;
    const int nctrl = (int)(2 + ceil_div(n, kSB) * ceil_div(kc, kSC));
    hipError_t e = launch_fill_int(ctrl, 0, nctrl, s);
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)(ceil_div(n, kSB) * ceil_div(kc, kSC));
    const bool inv = Dinv != nullptr;
    if (trans) {
        if (inv)
            hipLaunchKernelGGL((trsm_persist_kernel<T, true, true>), dim3(grid), dim3(256), 0, s, n, kc, L, ldl, B, ldb,
                               Dinv, ctrl, info);
        else
            hipLaunchKernelGGL((trsm_persist_kernel<T, true, false>), dim3(grid), dim3(256), 0, s, n, kc, L, ldl, B,
                               ldb, Dinv, ctrl, info);
    } else {
        if (inv)
            hipLaunchKernelGGL((trsm_persist_kernel<T, false, true>), dim3(grid), dim3(256), 0, s, n, kc, L, ldl, B,
                               ldb, Dinv, ctrl, info);
        else
            hipLaunchKernelGGL((trsm_persist_kernel<T, false, false>), dim3(grid), dim3(256), 0, s, n, kc, L, ldl, B,
                               ldb, Dinv, ctrl, info);
    }
    return hipGetLastError();
}

// Solve scratch in ints: the chain solve's tickets and flags, then (256-byte
// aligned) its n x kc partial sums (room for doubles).
size_t solve_ctrl_count(int n, int kc) {
    const size_t flags = ((chain_ctrl_count(n, kc) * sizeof(int) + 255) & ~(size_t)255) / sizeof(int);
    return flags + (size_t)n * kc * 2 + 64;
}
size_t solve_dinv_count(int n) { return (size_t)ceil_div(n, kSB) * kSB * kSB; }

template hipError_t launch_trtri_diag64<double>(int, const double*, i64, double*, hipStream_t);
template hipError_t launch_trtri_diag64<float>(int, const float*, i64, float*, hipStream_t);
template hipError_t launch_trtri_diag64_range<float>(int, int, int, const float*, i64, float*, hipStream_t);
template hipError_t launch_trsm_persist<double>(bool, int, int, const double*, i64, double*, i64, const double*, int*,
                                                const int*, hipStream_t);
template hipError_t launch_trsm_persist<float>(bool, int, int, const float*, i64, float*, i64, const float*, int*,
                                               const int*, hipStream_t);

}  // namespace elemamd
