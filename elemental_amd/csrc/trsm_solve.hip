// Triangular solves of cholesky::SolveAfter(LOWER, NORMAL, L, B)
// (include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75):
//   forward  B := inv(L) * B     Trsm(LEFT, LOWER, NORMAL,  NON_UNIT)  -> dtrsm_('L','L','N','N')
//   backward B := inv(L^T) * B   Trsm(LEFT, LOWER, ADJOINT, NON_UNIT)  -> dtrsm_('L','L','T','N')
// and, for one right-hand side, the Trsv pair (-> dtrsv_, blas.cpp:901-904).
//
// Blocked by 128-row diagonal blocks.  Per block: one kernel solves the
// diagonal triangle (one wave per right-hand side, the triangle in LDS), one
// kernel applies the block's contribution to every remaining row as an
// LDS-staged tile update (the L panel is read once from HBM, coalesced).
// T is the storage and arithmetic type (double for the FP64 path, float for
// the correction solves of the mixed-precision path).
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

constexpr int kSolveBs = 128;
constexpr int kDiagLd = kSolveBs + 1;

__device__ inline double wave_bcast(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ inline float wave_bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Solve the bs x bs diagonal triangle for kc right-hand sides in place.
// TRANS = false: L y = b (forward).  TRANS = true: L^T x = y (backward).
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void trsm_diag_kernel(int bs, int kc, const T* __restrict__ L, i64 ldl,
                                                        T* __restrict__ B, i64 ldb, const int* __restrict__ info) {
    __shared__ T sL[kSolveBs * kDiagLd];
    if (info != nullptr && *info != 0) return;
    const int t = threadIdx.x;
    for (int idx = t; idx < bs * bs; idx += 256) {
        const int i = idx % bs, j = idx / bs;
        sL[i + j * kDiagLd] = (i >= j) ? L[i + (i64)j * ldl] : T(0);
    }
    __syncthreads();
    const int lane = t & 63;
    const int col = blockIdx.x * 4 + (t >> 6);
    if (col >= kc) return;
    T* b = B + (i64)col * ldb;
    T v0 = (lane < bs) ? b[lane] : T(0);
    T v1 = (lane + 64 < bs) ? b[lane + 64] : T(0);
    if (!TRANS) {
        for (int kk = 0; kk < bs; ++kk) {
            const T own = (kk < 64) ? v0 : v1;
            const T y = wave_bcast(own, kk & 63) / sL[kk + kk * kDiagLd];
            if (lane == (kk & 63)) {
                if (kk < 64) v0 = y;
                else v1 = y;
            }
            const int r0 = lane, r1 = lane + 64;
            if (r0 > kk && r0 < bs) v0 -= y * sL[r0 + kk * kDiagLd];
            if (r1 > kk && r1 < bs) v1 -= y * sL[r1 + kk * kDiagLd];
        }
    } else {
        for (int kk = bs - 1; kk >= 0; --kk) {
            const T own = (kk < 64) ? v0 : v1;
            const T x = wave_bcast(own, kk & 63) / sL[kk + kk * kDiagLd];
            if (lane == (kk & 63)) {
                if (kk < 64) v0 = x;
                else v1 = x;
            }
            // rows i < kk use l_{kk,i}
            const int r0 = lane, r1 = lane + 64;
            if (r0 < kk) v0 -= x * sL[kk + r0 * kDiagLd];
            if (r1 < kk) v1 -= x * sL[kk + r1 * kDiagLd];
        }
    }
    if (lane < bs) b[lane] = v0;
    if (lane + 64 < bs) b[lane + 64] = v1;
}

// Tile update out[p, c] -= sum_q M[q, p] * Y[q, c] for p < mrows, q < bs, c < kc.
//   TRANS = false (forward):  M[q, p] = L[p, q]   L points at L(r0, c0)
//   TRANS = true  (backward): M[q, p] = L[q, p]   L points at L(c0, 0)
// Y (bs x kc, leading dim ldy) is the freshly solved block; out has leading
// dim ldo.  A workgroup owns 64 output rows and all kc (<= 32) columns of a
// column slab; the L tile is staged through LDS 16 contraction rows at a time.
constexpr int kUpdRows = 64;
constexpr int kUpdKc = 16;
constexpr int kUpdCols = 32;

template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void trsm_update_kernel(int mrows, int bs, int kc, const T* __restrict__ L, i64 ldl,
                                                          const T* __restrict__ Y, i64 ldy, T* __restrict__ out,
                                                          i64 ldo, const int* __restrict__ info) {
    __shared__ T sY[kSolveBs][kUpdCols + 1];
    if (info != nullptr && *info != 0) return;
    __shared__ T sM[kUpdKc][kUpdRows + 1];
    const int t = threadIdx.x;
    const int c_base = blockIdx.y * kUpdCols;
    const int kcl = min(kUpdCols, kc - c_base);
    for (int idx = t; idx < bs * kUpdCols; idx += 256) {
        const int q = idx % bs, c = idx / bs;
        sY[q][c] = (c < kcl) ? Y[q + (i64)(c_base + c) * ldy] : T(0);
    }
    const i64 p0 = (i64)blockIdx.x * kUpdRows;
    const int pl = t & 63;        // output row within the tile
    const int cg = t >> 6;        // column group: columns cg, cg+4, ..., cg+28
    T acc[kUpdCols / 4];
#pragma unroll
    for (int u = 0; u < kUpdCols / 4; ++u) acc[u] = T(0);

    for (int q0 = 0; q0 < bs; q0 += kUpdKc) {
        __syncthreads();
        if (!TRANS) {
            // 64 rows x 16 columns of L, each wave reads 64 consecutive rows of one column
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int qq = (t >> 6) + 4 * s;
                const i64 p = p0 + (t & 63);
                const int q = q0 + qq;
                sM[qq][t & 63] = (p < mrows && q < bs) ? L[p + (i64)q * ldl] : T(0);
            }
        } else {
            // 16 rows x 64 columns of L, 16 lanes read 16 consecutive rows of a column
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int qq = t & 15;
                const int pp = (t >> 4) + 16 * s;
                const i64 p = p0 + pp;
                const int q = q0 + qq;
                sM[qq][pp] = (p < mrows && q < bs) ? L[q + p * ldl] : T(0);
            }
        }
        __syncthreads();
        const int qmax = min(kUpdKc, bs - q0);
        for (int qq = 0; qq < qmax; ++qq) {
            const T m = sM[qq][pl];
#pragma unroll
            for (int u = 0; u < kUpdCols / 4; ++u) acc[u] += m * sY[q0 + qq][cg + 4 * u];
        }
    }
    const i64 p = p0 + pl;
    if (p < mrows) {
#pragma unroll
        for (int u = 0; u < kUpdCols / 4; ++u) {
            const int c = cg + 4 * u;
            if (c < kcl) out[p + (i64)(c_base + c) * ldo] -= acc[u];
        }
    }
}

template <typename T>
hipError_t launch_trsm_left_lower(bool trans, int n, int kc, const T* L, i64 ldl, T* B, i64 ldb, const int* info,
                                  hipStream_t s) {
    if (n <= 0 || kc <= 0) return hipSuccess;
    const int nblk = (int)ceil_div(n, kSolveBs);
    const dim3 dgrid((unsigned)ceil_div(kc, 4));
    const unsigned cy = (unsigned)ceil_div(kc, kUpdCols);
    if (!trans) {
        for (int b = 0; b < nblk; ++b) {
            const int c0 = b * kSolveBs;
            const int bs = min(kSolveBs, n - c0);
            hipLaunchKernelGGL((trsm_diag_kernel<T, false>), dgrid, dim3(256), 0, s, bs, kc,
                               L + c0 + (i64)c0 * ldl, ldl, B + c0, ldb, info);
            const int rem = n - c0 - bs;
            if (rem > 0) {
                dim3 ugrid((unsigned)ceil_div(rem, kUpdRows), cy);
                hipLaunchKernelGGL((trsm_update_kernel<T, false>), ugrid, dim3(256), 0, s, rem, bs, kc,
                                   L + (c0 + bs) + (i64)c0 * ldl, ldl, B + c0, ldb, B + c0 + bs, ldb, info);
            }
        }
    } else {
        for (int b = nblk - 1; b >= 0; --b) {
            const int c0 = b * kSolveBs;
            const int bs = min(kSolveBs, n - c0);
            hipLaunchKernelGGL((trsm_diag_kernel<T, true>), dgrid, dim3(256), 0, s, bs, kc,
                               L + c0 + (i64)c0 * ldl, ldl, B + c0, ldb, info);
            if (c0 > 0) {
                dim3 ugrid((unsigned)ceil_div(c0, kUpdRows), cy);
                hipLaunchKernelGGL((trsm_update_kernel<T, true>), ugrid, dim3(256), 0, s, c0, bs, kc,
                                   L + c0, ldl, B + c0, ldb, B, ldb, info);
            }
        }
    }
    return hipGetLastError();
}

template hipError_t launch_trsm_left_lower<double>(bool, int, int, const double*, i64, double*, i64, const int*,
                                                   hipStream_t);
template hipError_t launch_trsm_left_lower<float>(bool, int, int, const float*, i64, float*, i64, const int*,
                                                  hipStream_t);

}  // namespace elemamd
