// Per-(device, stream, thread) device scratch for entry points whose C ABI
// takes no workspace (the drop-in Trsm / SolveAfter / HPDSolve calls).
#pragma once

#include "common.hpp"

namespace elemamd {

// Device memory of at least `bytes`, owned by the calling thread and reserved
// for work enqueued on `stream`; grows on demand (the old buffer is released
// only after `stream` has drained).  Returns nullptr on allocation failure.
void* stream_scratch(size_t bytes, hipStream_t stream);

// An auxiliary stream (default priority) and two timing-free events paired
// with `stream` for the look-ahead factorization: work forked onto `aux`
// after an event recorded on `stream` and joined back the same way.
struct StreamPair {
    hipStream_t aux = nullptr;   // bulk trailing updates
    hipStream_t aux2 = nullptr;  // look-ahead block-column updates
    hipStream_t aux3 = nullptr;  // solve preparation beside the factorization
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr, ev_d = nullptr, ev_f = nullptr, ev_g = nullptr,
               ev_h = nullptr;
    double* pinned = nullptr;  // 8 doubles of pinned host memory (device-to-host scalars behind an event)
};
// Owned by the calling thread; nullptr members on failure.
StreamPair* stream_pair(hipStream_t stream);

// Polls a persistent triangular solve makes before it gives up a wait and
// records an error: 2^26, or ELEMAMD_SPIN_LIMIT from the environment (read at
// every launch; tests set a tiny limit to exercise the error path).
long solve_spin_limit();

}  // namespace elemamd
