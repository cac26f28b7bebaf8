// Device helpers shared by the persistent triangular-solve kernels.
#pragma once

#include "common.hpp"

namespace elemamd {

constexpr int kSB = 64;        // block rows
constexpr int kSC = 16;        // right-hand-side columns per workgroup
constexpr int kSLd = kSB + 1;
constexpr long kSpinLimit = 1L << 26;

// MFMA 16x16x4 of the storage type; output row index of accumulator v.
__device__ inline d4_t mfma16(double a, double b, d4_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4_t mfma16(float a, float b, f4_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
template <typename T>
__device__ inline int mfma_row(int lane, int v) {
    return sizeof(T) == 8 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
}
__device__ inline double wave_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ inline float wave_bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <typename T>
struct Acc4;
template <>
struct Acc4<double> {
    typedef d4_t type;
};
template <>
struct Acc4<float> {
    typedef f4_t type;
};

// Wait states for every MFMA already issued to retire (the longest takes 64
// cycles).  Called right after an MFMA product whose accumulators are read
// or rewritten by what follows -- in particular across loop exits, where
// the reads must not rely on the compiler's hazard bookkeeping.  The
// scheduling barriers keep the product's MFMAs on their side of the wait.
__device__ inline void mfma_retire() {
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_nop 15\n\ts_nop 15\n\ts_nop 15\n\ts_nop 15" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// acc += a * b (one 16x16x4 FP32 MFMA step) with the accumulator tied to
// the destination.  The compiler left to itself sometimes feeds an MFMA its
// accumulator from other registers than the ones it writes (renaming inside
// unrolled loops); such code gave wrong products on gfx950, so the products
// of the float kernels issue the instruction themselves.  The leading wait
// states cover a VALU write of an operand just before; the statement is
// volatile so these steps keep their order with mfma_retire(), after which
// the accumulator may be read.
__device__ inline void mfma_acc(f4_t& acc, float a, float b) {
    asm volatile("s_nop 1\n\tv_mfma_f32_16x16x4_f32 %0, %1, %2, %0" : "+v"(acc) : "v"(a), "v"(b));
}

// acc += a * b for one 16x16x32 bf16 MFMA step (8 bf16 per lane and
// operand), accumulator tied like mfma_acc.
__device__ inline void mfma_acc_bf16(f4_t& acc, s8_t a, s8_t b) {
    asm volatile("s_nop 1\n\tv_mfma_f32_16x16x32_bf16 %0, %1, %2, %0" : "+v"(acc) : "v"(a), "v"(b));
}

// Store of a value another workgroup will read (device-coherent, bypassing
// the per-XCD caches that are not coherent across XCDs) and the matching load.
template <typename T>
__device__ inline void coherent_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T coherent_load(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every wave waits until its coherent stores are acknowledged, then one
// thread raises the flag (coherent store): a reader that sees the flag and
// then loads coherently sees the data.
__device__ inline void publish_flag(int* flag, int t) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (t == 0) coherent_store(flag, 1);
}

}  // namespace elemamd
