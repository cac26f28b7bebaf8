// Triangular solves with the Cholesky factor as a chain plus helpers, one
// launch per direction:
//   forward  B := inv(L) B     (Trsm(LEFT, LOWER, NORMAL,  NON_UNIT) -> dtrsm_('L','L','N','N'))
//   backward B := inv(L^T) B   (Trsm(LEFT, LOWER, ADJOINT, NON_UNIT) -> dtrsm_('L','L','T','N'))
// as used by cholesky::SolveAfter (include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75).
//
// Rows are cut into 64-row blocks, solved in order (positions pos = 0, 1, ...;
// reversed blocks for the backward solve).  Block pos needs
//     Y_pos = diag_pos^{-1} (B_pos - sum_{q < pos} M(pos, q) Y_q).
// * One chain workgroup per 16-column slab of right-hand sides walks the
//   positions in order.  It keeps the last D solved blocks in LDS and applies
//   their contributions (the band q = pos-D .. pos-1) itself, adds the partial
//   sum of all older blocks computed by a helper, and solves the diagonal
//   block.  Solved blocks never make a round trip through memory on the
//   chain's own critical path.
// * One helper workgroup per (pos >= D+1, slab) accumulates
//   sum_{q <= pos-D-1} M(pos, q) Y_q as the chain publishes those blocks and
//   hands the partial sum to the chain.  Helpers have D-1 chain steps of slack.
// Cross-workgroup data (solved blocks, partial sums, flags) moves only
// through device-coherent loads and stores; a writer waits for the
// acknowledgement of its stores before it raises a flag (solve_dev.hpp).
// Workgroups take tickets in launch order: the chains first, then helpers by
// increasing position, so every wait is on a workgroup that is running or
// done.  Every wait is bounded (spin_limit polls, then an error in ctrl[1];
// ctrl[1] is cleared once per solve by the caller, not per launch, so a wait
// that gave up in the forward solve is still reported after the backward one).
//
// Diagonal blocks: USE_INV multiplies by the precomputed inverse (float
// solves of iterative refinement), otherwise substitution with reciprocal
// pivots (the FP64 solves; exact on integer problems with power-of-two
// diagonals).
#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
#include "solve_dev.hpp"

namespace elemamd {

// Band width D of the chain: its LDS holds D tiles of L plus D solved blocks.
template <typename T>
struct ChainDepth {
    static constexpr int value = sizeof(T) == 8 ? 3 : 4;
};

// TRANS: the backward direction (blocks in reverse order).  UPPER: the
// factor is stored in the upper triangle (L = U^T, element (i, j) of L at
// (j, i)), so the tiles are read transposed: READ_T = TRANS xor UPPER.
// unit: the diagonal is taken as one and never used (Trsm's UNIT diag).
template <typename T, bool TRANS, bool UPPER, bool USE_INV>
__global__ __launch_bounds__(256) void trsm_chain_kernel(int n, int kc, const T* __restrict__ L, i64 ldl,
                                                         T* __restrict__ B, i64 ldb, const T* __restrict__ Dinv,
                                                         int* __restrict__ ctrl, T* __restrict__ part,
                                                         const int* __restrict__ info, long spin_limit, int unit) {
    static_assert(!(UPPER && USE_INV), "upper-stored factors take the substitution path");
    constexpr bool READ_T = TRANS != UPPER;
    constexpr int D = ChainDepth<T>::value;
    typedef typename Acc4<T>::type acc_t;
    __shared__ T sM[D][kSB][kSLd];      // [q][p] tiles: chain band (then diagonal in sM[0]); helpers: double buffer
    __shared__ T sRing[D][kSB][kSC + 1];  // chain: last D solved blocks [q][c]; helpers: Y double buffer
    __shared__ T sR[kSB][kSC + 1];
    __shared__ T sRinv[kSB];
    __shared__ int s_ord, s_cnt;
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nT = (int)ceil_div(n, kSB);
    const int CY = (int)ceil_div(kc, kSC);
    if (t == 0) s_ord = atomicAdd(&ctrl[0], 1);
    __syncthreads();
    const int ord = s_ord;
    int* flagsX = ctrl + 2;            // by position: solved block published
    int* flagsP = ctrl + 2 + nT * CY;  // by position: helper's partial sum published
    auto blk = [&](int pos) { return TRANS ? nT - 1 - pos : pos; };

    // contraction tile: output block I (rows p), contraction block J (rows q)
    T gm[16];
    auto gload = [&](int I, int J) {
        const i64 r0 = (i64)I * kSB, q0 = (i64)J * kSB;
        // both indices are checked: either block may be the partial last one
        if (!READ_T) {  // M[q][p] = L(r0+p, q0+q), lanes along p
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const i64 p = r0 + lane, q = q0 + w + 4 * s;
                gm[s] = (p < n && q < n) ? L[p + q * ldl] : T(0);
            }
        } else {  // M[q][p] = L(q0+q, r0+p), lanes along q
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const i64 q = q0 + lane, p = r0 + w + 4 * s;
                gm[s] = (q < n && p < n) ? L[q + p * ldl] : T(0);
            }
        }
    };
    auto sstore = [&](int buf) {
        if (!READ_T) {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[buf][w + 4 * s][lane] = gm[s];
        } else {
#pragma unroll
            for (int s = 0; s < 16; ++s) sM[buf][lane][w + 4 * s] = gm[s];
        }
    };
    auto wait_one = [&](int* f) {  // by thread 0, then the workgroup
        if (t == 0) {
            long spins = 0;
            while (coherent_load(f) == 0) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > spin_limit) {
                    atomicOr(&ctrl[1], 1);
                    break;
                }
            }
        }
        __syncthreads();
    };

    if (ord >= CY) {
        // ------------------------------------------------------------ helper
        const int h = ord - CY;
        const int pos = h / CY + D + 1;
        const int cy = h % CY;
        const int c_base = cy * kSC, kcl = min(kSC, kc - c_base);
        const int I = blk(pos);
        const i64 r0 = (i64)I * kSB;
        const int nJ = pos - D;  // positions 0 .. pos-D-1
        T gy[4];
        auto gload_y = [&](int J) {  // a solved block: coherent loads
            const i64 q0 = (i64)J * kSB;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = w + 4 * s;
                const i64 q = q0 + lane;
                gy[s] = (q < n && c < kcl) ? coherent_load(&B[q + (i64)(c_base + c) * ldb]) : T(0);
            }
        };
        auto ready_run = [&](int done) -> int {  // solved positions from `done` on (>= 1)
            if (w == 0) {
                long spins = 0;
                int cnt;
                for (;;) {
                    const int q = done + lane;
                    int f = 0;
                    if (q < nJ) f = coherent_load(&flagsX[q * CY + cy]);
                    const unsigned long long msk = __ballot(f != 0);
                    cnt = (~msk == 0ull) ? 64 : (int)__builtin_ctzll(~msk);
                    if (cnt > 0) break;
                    if (++spins > spin_limit) {
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
        acc_t acc = {};
        int buf = 0;
        gload(I, blk(0));
        int done = 0;
        while (done < nJ) {
            const int cnt = ready_run(done);
            gload_y(blk(done));
            for (int b = 0; b < cnt; ++b) {
                const int q = done + b;
                sstore(buf);
#pragma unroll
                for (int s = 0; s < 4; ++s) sRing[buf][lane][w + 4 * s] = gy[s];
                gload(I, blk(min(q + 1, nJ - 1)));
                gload_y(blk((b + 1 < cnt) ? q + 1 : q));
                __syncthreads();
#pragma unroll
                for (int ks = 0; ks < kSB / 4; ++ks) {
                    const int qq = 4 * ks + (lane >> 4);
                    acc = mfma16(sRing[buf][qq][lane & 15], sM[buf][qq][16 * w + (lane & 15)], acc);
                }
                buf ^= 1;
            }
            done += cnt;
        }
        // dependent zero-operand MFMA before the accumulator is read (the
        // compiler's wait states after a loop back edge are short for FP64 MFMA)
        acc = mfma16(T(0), T(0), acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = mfma_row<T>(lane, v);
            const i64 row = r0 + 16 * w + (lane & 15);
            if (row < n && c < kcl) coherent_store(&part[row + (i64)(c_base + c) * n], acc[v]);
        }
        publish_flag(&flagsP[pos * CY + cy], t);
        return;
    }

    // ----------------------------------------------------------------- chain
    // Every operand of step pos that does not depend on the chain itself (the
    // D band tiles of L, the diagonal operand, the own right-hand sides) is
    // loaded into registers during step pos-1, so the chain never waits for
    // L from memory; only the helper's partial sum is fetched in the step.
    const int cy = ord;
    const int c_base = cy * kSC, kcl = min(kSC, kc - c_base);
    T gt[D][16], gd[16], gb[4];
    auto prefetch = [&](int pos) {
        const int I = blk(pos);
        const i64 r0 = (i64)I * kSB;
        const int bs = (int)min<i64>(kSB, n - r0);
        const int qlo = max(0, pos - D);
#pragma unroll
        for (int t2 = 0; t2 < D; ++t2) {
            const int q = qlo + t2;
            if (q < pos) {  // uniform
                gload(I, blk(q));
#pragma unroll
                for (int s2 = 0; s2 < 16; ++s2) gt[t2][s2] = gm[s2];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = mfma_row<T>(lane, v);
            const i64 row = r0 + 16 * w + (lane & 15);
            gb[v] = (row < n && c < kcl) ? B[row + (i64)(c_base + c) * ldb] : T(0);
        }
        if constexpr (USE_INV) {
            const T* Di = Dinv + (i64)I * kSB * kSB;
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) {
                const int qq = w + 4 * s2;
                gd[s2] = !TRANS ? Di[lane + qq * kSB] : Di[qq + lane * kSB];  // M[q][p] = Dinv(p,q) / Dinv(q,p)
            }
        } else {
            // lower-triangular M(i, k) at [k][i], identity on padded rows; the
            // backward solve is the same sweep on reversed indices
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) {
                const int kk = w + 4 * s2;
                // stored (row, column) of M(lane, kk) within the block: an
                // upper-stored factor swaps them
                const int li = TRANS ? (kSB - 1 - kk) : lane;
                const int lk = TRANS ? (kSB - 1 - lane) : kk;
                const int gi = UPPER ? lk : li;
                const int gk = UPPER ? li : lk;
                T v = T(0);
                if (gi < bs && gk < bs) {
                    if (lane >= kk) v = L[r0 + gi + (r0 + gk) * ldl];
                } else if (lane == kk) {
                    v = T(1);
                }
                gd[s2] = v;
            }
        }
    };
    auto stage = [&](int slot, const T (&g)[16]) {  // a band tile into sM[slot] ([q][p] layout)
        if (!READ_T) {
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) sM[slot][w + 4 * s2][lane] = g[s2];
        } else {
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) sM[slot][lane][w + 4 * s2] = g[s2];
        }
    };

    prefetch(0);
    for (int pos = 0; pos < nT; ++pos) {
        const int I = blk(pos);
        const i64 r0 = (i64)I * kSB;
        const int qlo = max(0, pos - D);
        const int nbt = pos - qlo;  // band tiles of this step
        // Y_{pos-1} is stored (coherently) and no other load is outstanding
        // except the prefetch: publish it, which also orders the LDS reuse below
        if (pos > 0) publish_flag(&flagsX[(pos - 1) * CY + cy], t);
#pragma unroll
        for (int t2 = 0; t2 < D; ++t2)
            if (t2 < nbt) stage(t2, gt[t2]);
        T gdc[16], gbc[4];
#pragma unroll
        for (int s2 = 0; s2 < 16; ++s2) gdc[s2] = gd[s2];
#pragma unroll
        for (int v = 0; v < 4; ++v) gbc[v] = gb[v];
        __syncthreads();
        if (pos + 1 < nT) prefetch(pos + 1);
        // partial sum of the older blocks (helper), fetched while the band runs
        T sp[4] = {T(0), T(0), T(0), T(0)};
        if (pos >= D + 1) {
            wait_one(&flagsP[pos * CY + cy]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = mfma_row<T>(lane, v);
                const i64 row = r0 + 16 * w + (lane & 15);
                sp[v] = (row < n && c < kcl) ? coherent_load(&part[row + (i64)(c_base + c) * n]) : T(0);
            }
        }
        // band: the last D solved blocks, all operands in LDS
        acc_t acc = {};
        for (int t2 = 0; t2 < nbt; ++t2) {
            const int q = qlo + t2;
#pragma unroll
            for (int ks = 0; ks < kSB / 4; ++ks) {
                const int qq = 4 * ks + (lane >> 4);
                acc = mfma16(sRing[q % D][qq][lane & 15], sM[t2][qq][16 * w + (lane & 15)], acc);
            }
        }
        // dependent zero-operand MFMA before the accumulator is read (the
        // compiler's wait states after a loop back edge are short for FP64 MFMA)
        acc = mfma16(T(0), T(0), acc);
        __syncthreads();  // band products are done with sM and sRing
#pragma unroll
        for (int v = 0; v < 4; ++v) sR[16 * w + (lane & 15)][mfma_row<T>(lane, v)] = gbc[v] - acc[v] - sp[v];
#pragma unroll
        for (int s2 = 0; s2 < 16; ++s2) sM[0][w + 4 * s2][lane] = gdc[s2];
        __syncthreads();
        const int slot = pos % D;
        if constexpr (USE_INV) {
            acc_t y = {};
#pragma unroll
            for (int ks = 0; ks < kSB / 4; ++ks) {
                const int qq = 4 * ks + (lane >> 4);
                y = mfma16(sR[qq][lane & 15], sM[0][qq][16 * w + (lane & 15)], y);
            }
            y = mfma16(T(0), T(0), y);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = mfma_row<T>(lane, v);
                const int p = 16 * w + (lane & 15);
                const i64 row = r0 + p;
                sRing[slot][p][c] = y[v];
                if (row < n && c < kcl) coherent_store(&B[row + (i64)(c_base + c) * ldb], y[v]);
            }
        } else {
            if (t < kSB) sRinv[t] = unit ? T(1) : T(1) / sM[0][t][t];
            __syncthreads();
            const int myrow = TRANS ? (kSB - 1 - lane) : lane;
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = sR[myrow][4 * w + u];
            // unrolled so the LDS reads of the coming columns are issued early;
            // the sweep's dependence runs only through v (lane kk's value)
#pragma unroll 16
            for (int kk = 0; kk < kSB; ++kk) {
                const T rk = sRinv[kk];
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
                sRing[slot][myrow][c] = v[u];
                if (row < n && c < kcl) coherent_store(&B[row + (i64)(c_base + c) * ldb], v[u]);
            }
        }
    }
    publish_flag(&flagsX[(nT - 1) * CY + cy], t);
}

size_t chain_ctrl_count(int n, int kc) { return 2 + 2 * (size_t)ceil_div(n, kSB) * ceil_div(kc, kSC); }

template <typename T, bool TRANS, bool UPPER, bool INV>
static void chain_launch(unsigned grid, hipStream_t s, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb,
                         const T* Dinv, int* ctrl, T* part, const int* info, long spin, int unit) {
    hipLaunchKernelGGL((trsm_chain_kernel<T, TRANS, UPPER, INV>), dim3(grid), dim3(256), 0, s, n, kc, L, ldl, B, ldb,
                       Dinv, ctrl, part, info, spin, unit);
}

template <typename T>
hipError_t launch_trsm_chain(bool trans, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv, int* ctrl,
                             T* part, const int* info, hipStream_t s, bool upper, bool unit) {
    if (n <= 0 || kc <= 0) return hipSuccess;
    constexpr int D = ChainDepth<T>::value;
    const int nT = (int)ceil_div(n, kSB), CY = (int)ceil_div(kc, kSC);
    const bool inv = Dinv != nullptr;
    if (inv && (upper || unit)) return hipErrorInvalidValue;
    // ctrl[0] (tickets) and the flags from ctrl[2] on restart at zero; the
    // error word ctrl[1] keeps what earlier launches of the solve recorded
    hipError_t e = launch_fill_int(ctrl, 0, 1, s);
    if (e == hipSuccess) e = launch_fill_int(ctrl + 2, 0, (int)chain_ctrl_count(n, kc) - 2, s);
    if (e != hipSuccess) return e;
    const long spin = solve_spin_limit();
    const unsigned grid = (unsigned)(CY + CY * std::max(0, nT - D - 1));
    const int u = unit ? 1 : 0;
    if (inv) {
        if (trans) chain_launch<T, true, false, true>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
        else chain_launch<T, false, false, true>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
    } else if (!upper) {
        if (trans) chain_launch<T, true, false, false>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
        else chain_launch<T, false, false, false>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
    } else {
        if (trans) chain_launch<T, true, true, false>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
        else chain_launch<T, false, true, false>(grid, s, n, kc, L, ldl, B, ldb, Dinv, ctrl, part, info, spin, u);
    }
    return hipGetLastError();
}

template hipError_t launch_trsm_chain<double>(bool, int, int, const double*, i64, double*, i64, const double*, int*,
                                              double*, const int*, hipStream_t, bool, bool);
template hipError_t launch_trsm_chain<float>(bool, int, int, const float*, i64, float*, i64, const float*, int*,
                                             float*, const int*, hipStream_t, bool, bool);

}  // namespace elemamd
