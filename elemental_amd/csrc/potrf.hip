// Diagonal-block Cholesky: Elemental's cholesky::LVar3Unb
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:19-49) on an n x n block,
// n <= 128, by one 512-thread workgroup.
//
// Rounding order matches the reference loop exactly: per column j the pivot
// is sqrt(a_jj), the column below it is divided by that pivot, and every
// trailing element takes a_ik := a_ik - a_ij*a_kj as a separately rounded
// product and difference, in increasing j.  Contraction into FMA is switched
// off, so the block result is bit-identical to the reference loop run in IEEE
// double on the CPU.
//
// Layout: the block lives in registers.  Thread (tr, tc) = (t % 32, t / 32)
// owns the 32 elements (i, k) = (tr + 32a, tc + 16b), a < 4, b < 8, a 2-D
// cyclic distribution that keeps the shrinking trailing triangle balanced.
// Column j's owners are the 32 consecutive threads with tc = j % 16 (one
// wave): they take the pivot from its owner with a lane read, divide their
// elements and publish the finished column in LDS.  One barrier per column,
// then every thread applies the rank-1 update to its active elements.
// Storage may be double or float (the mixed-precision factor keeps its
// working matrix in float); arithmetic is always double.
#include "common.hpp"
#include "kernels.hpp"
#include "potrf_dev.hpp"

namespace elemamd {

template <typename T>
__global__ __launch_bounds__(kPotrfThreads) void potrf_unb_kernel(int n, T* __restrict__ A, i64 lda,
                                                                  int* __restrict__ info, int col_offset) {
    __builtin_amdgcn_s_setprio(3);  // chain kernel: issues ahead of the trailing update's waves
    potrf_unb_block<T>(n, A, lda, info, col_offset);
}

template <typename T>
hipError_t launch_potrf_unb(int n, T* A, i64 lda, int* info, int col_offset, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(potrf_unb_kernel<T>, dim3(1), dim3(kPotrfThreads), 0, s, n, A, lda, info, col_offset);
    return hipGetLastError();
}

template hipError_t launch_potrf_unb<double>(int, double*, i64, int*, int, hipStream_t);
template hipError_t launch_potrf_unb<float>(int, float*, i64, int*, int, hipStream_t);

}  // namespace elemamd
