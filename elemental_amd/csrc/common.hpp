// Shared definitions for the CDNA4 (gfx950) kernels and host drivers.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace elemamd {

typedef int64_t i64;
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));  // 8 x bf16 bit patterns

// Status codes of the C-ABI (>0 is a LAPACK-style failing column).
enum : int {
    kOk = 0,
    kErrArg = -1,        // invalid argument (size, leading dimension, pointer)
    kErrHip = -2,        // a HIP runtime call failed
    kErrNotConverged = -3,
    kErrWorkspace = -4,  // workspace missing or too small
};

#define ELEMAMD_HIP_TRY(expr)                                   \
    do {                                                        \
        hipError_t e__ = (expr);                                \
        if (e__ != hipSuccess) return ::elemamd::kErrHip;       \
    } while (0)

__host__ __device__ inline i64 ceil_div(i64 a, i64 b) { return (a + b - 1) / b; }

// Normwise backward error of one column, ||r||_inf / (||A||_inf ||x||_inf +
// ||b||_inf), from the norms: a NaN anywhere (a NaN in A, x or b reaches the
// residual or one of the norms) or a zero denominator with a nonzero
// residual gives +inf, so such a column never counts as converged.
__host__ __device__ inline double column_backward_error(double rnorm, double anorm, double xnorm, double bnorm) {
    const double den = anorm * xnorm + bnorm;
    if (rnorm != rnorm || den != den) return __builtin_inf();
    const double e = den > 0.0 ? rnorm / den : (rnorm > 0.0 ? __builtin_inf() : 0.0);
    return e != e ? __builtin_inf() : e;
}

// Column-major element (i, j) of a matrix with leading dimension ld.
template <typename T>
__host__ __device__ inline T& at(T* A, i64 i, i64 j, i64 ld) { return A[i + j * ld]; }
template <typename T>
__host__ __device__ inline const T& at(const T* A, i64 i, i64 j, i64 ld) { return A[i + j * ld]; }

// Broadcast lane R of each aligned group of 4 lanes to the whole group
// (DPP quad_perm; no LDS round trip, no barrier).
template <int R>
__device__ inline double quad_bcast(double v) {
    constexpr int ctrl = R | (R << 2) | (R << 4) | (R << 6);
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// bf16 helpers for the split-BF16x3 products: hi = round-to-nearest-even
// bf16 of x, lo = bf16 of the remainder x - hi (exact in fp32).
__device__ inline unsigned short f32_to_bf16_rne(float x) {
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf/nan pass-through
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

}  // namespace elemamd
