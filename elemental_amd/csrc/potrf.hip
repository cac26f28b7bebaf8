// Diagonal-block Cholesky: Elemental's cholesky::LVar3Unb
// (include/elemental/lapack-like/Cholesky/LVar3.hpp:19-49) on an n x n block,
// n <= 128, held entirely in LDS (128 x 128 doubles = 128 KiB of gfx950's
// 160 KiB per CU).
//
// Rounding order matches the reference loop exactly: per column j the pivot
// is sqrt(a_jj), the column below it is divided by that pivot, and every
// trailing element takes a_ik := a_ik - a_ij*a_kj as a separately rounded
// product and difference, in increasing j.  Contraction into FMA is switched
// off in this kernel so that the block result is bit-identical to the
// reference loop run in IEEE double on the CPU.
//
// Storage may be double or float (the mixed-precision factor keeps its
// working matrix in float); arithmetic is always double.
#include "common.hpp"
#include "kernels.hpp"

namespace elemamd {

constexpr int kPotrfMaxN = 128;
constexpr int kPotrfThreads = 512;

template <typename T>
__global__ __launch_bounds__(kPotrfThreads) void potrf_unb_kernel(int n, T* __restrict__ A, i64 lda,
                                                                  int* __restrict__ info, int col_offset) {
#pragma clang fp contract(off)
    __shared__ double s[kPotrfMaxN * kPotrfMaxN];
    if (info != nullptr && *info != 0) return;  // an earlier block failed
    const int t = threadIdx.x;
    const int ld = n;

    for (int idx = t; idx < n * n; idx += kPotrfThreads) {
        const int i = idx % n, j = idx / n;
        if (i >= j) s[i + j * ld] = (double)A[i + (i64)j * lda];
    }
    __syncthreads();

    // Threads 0..127 own rows; the four groups t/128 split the columns k of
    // the rank-1 update round-robin, so a wave touches 64 consecutive rows
    // of one column (conflict-free LDS access, broadcast of a_kj).
    const int r = t & (kPotrfMaxN - 1);
    const int q = t >> 7;
    int fail = 0;
    int done = n;
    for (int j = 0; j < n; ++j) {
        const double ajj = s[j + j * ld];
        if (ajj <= 0.0) {  // uniform: every thread read the same value
            fail = j + 1;
            done = j;
            break;
        }
        const double alpha = sqrt(ajj);
        __syncthreads();  // all reads of a_jj precede its overwrite
        if (t == 0) s[j + j * ld] = alpha;
        for (int i = j + 1 + t; i < n; i += kPotrfThreads) s[i + j * ld] = s[i + j * ld] / alpha;
        __syncthreads();
        const int i = j + 1 + r;
        if (i < n) {
            const double aij = s[i + j * ld];
            for (int k = j + 1 + q; k <= i; k += 4) s[i + k * ld] = s[i + k * ld] - aij * s[k + j * ld];
        }
        __syncthreads();
    }
    (void)done;
    // Write back the lower triangle.  On failure at column j this is the
    // state the reference leaves behind when LVar3Unb throws: columns < j
    // factored and the trailing triangle updated by them.
    for (int idx = t; idx < n * n; idx += kPotrfThreads) {
        const int i = idx % n, j = idx / n;
        if (i >= j) A[i + (i64)j * lda] = (T)s[i + j * ld];
    }
    if (fail && t == 0 && info != nullptr) *info = col_offset + fail;
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
