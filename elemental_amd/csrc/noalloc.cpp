// SolveAfter with a fixed device scratch array: for callers that must not
// allocate (tight serving loops, harnesses that forbid hipMalloc).  The call
// ends by synchronizing the stream to read the solve's error word, as the
// allocating entry point does.  The chain solves' tickets, flags and helper partial
// sums live in `scratch`, one copy per GPU for the life of the process, so
// this entry point calls no allocator.  Calls that overlap on one GPU share
// that GPU's array and race; the caller serialises them.
#include <cstdint>
#include <mutex>

#include "../../include/elemental_amd.h"
#include "kernels.hpp"
#include "lapack.hpp"

using namespace elemamd;

namespace {

// 9 MiB: the helpers' n x k FP64 partial sums (8 n k bytes) plus the flags;
// n k <= 1,130,000 right-hand-side entries (n = 32768 with k = 32 fits).
constexpr size_t kScratchBytes = (size_t)9 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address).  Host code uses
// this address, never the array's name.
void* scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }

}  // namespace

extern "C" {

size_t elemamd_solve_after_f64_noalloc_limit(void) { return kScratchBytes; }

int elemamd_solve_after_f64_noalloc(char uplo, char orientation, int n, int k, const double* A, int64_t lda,
                                    double* B, int64_t ldb, void* stream) {
    uplo = up(uplo);
    const char o = up(orientation);
    if ((uplo != 'L' && uplo != 'U') || (o != 'N' && o != 'T' && o != 'C')) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1)) return ELEMAMD_ERR_ARG;
    if (n == 0 || k == 0) return ELEMAMD_OK;
    // the same ints the allocating entry point carves after its (unused for
    // FP64) diagonal-inverse region: flags, then the partial sums
    if (solve_ctrl_count(n, k) * sizeof(int) > kScratchBytes) return ELEMAMD_ERR_WORKSPACE;
    int* ctrl = (int*)scratch_ptr();
    if (ctrl == nullptr) return ELEMAMD_ERR_HIP;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // solve_apply clears the error word and every launch its tickets and
    // flags (fill kernels, no memset); the partial sums are written before
    // they are read
    const int rc = solve_apply<double>(true, true, n, k, A, lda, B, ldb, nullptr, ctrl, nullptr, s, uplo == 'U', false);
    return rc ? rc : solve_read_error(ctrl, s);
}

}  // extern "C"
