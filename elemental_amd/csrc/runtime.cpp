#include "runtime.hpp"

#include <map>
#include <utility>

namespace elemamd {

namespace {
struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};
struct Cache {
    std::map<std::pair<int, hipStream_t>, Buf> bufs;
    ~Cache() {
        for (auto& kv : bufs)
            if (kv.second.p) (void)hipFree(kv.second.p);
    }
};
struct PairCache {
    std::map<std::pair<int, hipStream_t>, StreamPair> pairs;
    ~PairCache() {
        for (auto& kv : pairs) {
            if (kv.second.ev_a) (void)hipEventDestroy(kv.second.ev_a);
            if (kv.second.ev_b) (void)hipEventDestroy(kv.second.ev_b);
            if (kv.second.aux) (void)hipStreamDestroy(kv.second.aux);
        }
    }
};
}  // namespace

StreamPair* stream_pair(hipStream_t stream) {
    static thread_local PairCache cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    StreamPair& p = cache.pairs[{dev, stream}];
    if (p.aux == nullptr) {
        if (hipStreamCreateWithFlags(&p.aux, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&p.ev_a, hipEventDisableTiming) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&p.ev_b, hipEventDisableTiming) != hipSuccess) return nullptr;
    }
    return &p;
}

void* stream_scratch(size_t bytes, hipStream_t stream) {
    static thread_local Cache cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    Buf& b = cache.bufs[{dev, stream}];
    if (b.cap >= bytes && b.p) return b.p;
    if (b.p) {
        if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t cap = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    if (hipMalloc(&b.p, cap) != hipSuccess) {
        b.p = nullptr;
        return nullptr;
    }
    b.cap = cap;
    return b.p;
}

}  // namespace elemamd
