#include "runtime.hpp"

#include <cstdlib>
#include <initializer_list>
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
void release(StreamPair& p) {
    for (hipEvent_t* e : {&p.ev_a, &p.ev_b, &p.ev_c, &p.ev_d, &p.ev_f, &p.ev_g, &p.ev_h})
        if (*e) (void)hipEventDestroy(*e), *e = nullptr;
    for (hipStream_t* t : {&p.aux, &p.aux2, &p.aux3})
        if (*t) (void)hipStreamDestroy(*t), *t = nullptr;
    if (p.pinned) (void)hipHostFree(p.pinned), p.pinned = nullptr;
}
struct PairCache {
    std::map<std::pair<int, hipStream_t>, StreamPair> pairs;
    ~PairCache() {
        for (auto& kv : pairs) release(kv.second);
    }
};
}  // namespace

StreamPair* stream_pair(hipStream_t stream) {
    static thread_local PairCache cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    StreamPair& p = cache.pairs[{dev, stream}];
    if (p.aux == nullptr) {  // all or nothing
        bool ok = hipStreamCreateWithFlags(&p.aux, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&p.aux2, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&p.aux3, hipStreamNonBlocking) == hipSuccess;
        for (hipEvent_t* e : {&p.ev_a, &p.ev_b, &p.ev_c, &p.ev_d, &p.ev_f, &p.ev_g, &p.ev_h})
            ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
        ok = ok && hipHostMalloc((void**)&p.pinned, 8 * sizeof(double), hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            release(p);
            return nullptr;
        }
    }
    return &p;
}

long solve_spin_limit() {
    const char* v = std::getenv("ELEMAMD_SPIN_LIMIT");
    if (v != nullptr && v[0] != 0) {
        const long l = std::strtol(v, nullptr, 10);
        if (l > 0) return l;
    }
    return 1L << 26;
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
