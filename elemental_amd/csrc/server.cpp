// Serving loop: continuous batching of factor-and-solve requests of mixed
// size on one GPU (no reference counterpart; each request is the
// HPDSolve(LOWER, NORMAL, A, B) computation of
// include/elemental/lapack-like/HPDSolve.hpp:18-28, plain FP64 or
// low-precision factor + FP64 refinement).
//
// * Slots.  `slots` worker threads, each with its own HIP stream, take
//   requests from one queue; requests on different slots overlap on the
//   device.  A request runs start to end on its slot's stream, so its
//   arithmetic -- and its result -- do not depend on what else is running.
// * Admission by memory.  Every request reserves its device memory before it
//   starts: the factor (resident afterwards) and the transient workspace.  A
//   queued request is admitted, in arrival order, as soon as a slot is idle
//   and the reservation fits in the budget (default: free HBM at creation
//   minus a reserve); later requests that fit may pass one that does not.
//   Unpinned resident factors are evicted least-recently-used to make room.
// * Resident factors.  A factor is kept under the request's key; a request
//   that asks for reuse of a resident key only runs the solves (plus the
//   refinement) with its new right-hand sides.  A reuse request whose key is
//   being factored waits for that factorization instead of repeating it.
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "lapack.hpp"

namespace elemamd {

namespace {

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct Server::Factor {
    int key = 0, n = 0;
    void* mem = nullptr;
    size_t bytes = 0;
    int pins = 0;
    bool ready = false;
    bool failed = false;
    uint64_t last_use = 0;
    MixedFactor mf{};
};

struct Server::Req {
    ServeRequest q;
    ServeResult res{};
    size_t wbytes = 0;           // transient workspace reserved at admission
    hipEvent_t ready = nullptr;  // recorded on the submitter's stream
    bool done = false;
};

Server::Server(const ServerConfig& cfg) : cfg_(cfg), t0_(std::chrono::steady_clock::now()) {}

int Server::start() {
    if (cfg_.slots < 1 || cfg_.slots > 64 || cfg_.nb < 1 || cfg_.nb > 128) return kErrArg;
    if (cfg_.precision < 0 || cfg_.precision > 2) return kErrArg;
    ELEMAMD_HIP_TRY(hipGetDevice(&device_));
    if (cfg_.memory_budget == 0) {
        size_t fr = 0, tot = 0;
        ELEMAMD_HIP_TRY(hipMemGetInfo(&fr, &tot));
        const size_t reserve = std::max<size_t>(tot / 20, (size_t)2 << 30);
        budget_ = fr > reserve ? fr - reserve : fr / 2;
    } else {
        budget_ = cfg_.memory_budget;
    }
    // stream-ordered allocations come from the device's default pool; keep
    // freed blocks in the pool instead of returning them to the driver
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device_) == hipSuccess && pool != nullptr) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    for (int i = 0; i < cfg_.slots; ++i) workers_.emplace_back(&Server::worker, this, i);
    return kOk;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    // factor memory comes from the stream-ordered pool: return it the same way
    for (auto& kv : cache_)
        if (kv.second->mem) (void)hipFreeAsync(kv.second->mem, nullptr);
    (void)hipStreamSynchronize(nullptr);
}

size_t Server::factor_bytes(int n) const {
    return cfg_.precision == 0 ? al256(sizeof(double) * (size_t)n * n) + 256 : mixed_factor_bytes(n);
}

size_t Server::work_bytes(int n, int k, bool factor) const {
    if (cfg_.precision == 0) return al256(solve_work_bytes<double>(n, k)) + 256;
    MixedOpts o;
    o.nb = cfg_.nb;
    o.trailing = cfg_.precision == 2 ? kTrailingBf16x3 : kTrailingF32;
    o.tol = cfg_.tol;
    o.fallback = 0;
    // a request that factors also needs the factorization's plane scratch
    return al256(mixed_refine_bytes(n, k, o)) + (factor ? al256(lowprec_scratch_bytes(n, o.nb, o.trailing)) : 0);
}

int Server::submit(const ServeRequest& q, int64_t* ticket) {
    if (q.n < 1 || q.k < 1 || q.lda < q.n || q.ldb < q.n || q.A == nullptr || q.B == nullptr) return kErrArg;
    if (factor_bytes(q.n) + work_bytes(q.n, q.k, true) > budget_) return kErrWorkspace;  // could never be admitted
    auto r = std::make_shared<Req>();
    r->q = q;
    r->res.t_submit = now_s(t0_);
    {  // order the request after the producer of A and B (null: the default stream)
        if (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) != hipSuccess) return kErrHip;
        if (hipEventRecord(r->ready, q.ready) != hipSuccess) {
            (void)hipEventDestroy(r->ready);
            return kErrHip;
        }
    }
    {
        std::lock_guard<std::mutex> g(mu_);
        r->res.ticket = next_ticket_++;
        reqs_[r->res.ticket] = r;
        queue_.push_back(r);
        if (ticket) *ticket = r->res.ticket;
    }
    cv_.notify_all();
    return kOk;
}

int Server::wait(int64_t ticket, ServeResult* out) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return kErrArg;
    auto r = it->second;
    done_cv_.wait(g, [&] { return r->done; });
    if (out) *out = r->res;
    reqs_.erase(ticket);
    return kOk;
}

int Server::evict(int key) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = cache_.find(key);
    if (it == cache_.end()) return kOk;
    if (it->second->pins > 0 || !it->second->ready) return kErrArg;
    if (it->second->mem) {
        (void)hipFreeAsync(it->second->mem, nullptr);
        (void)hipStreamSynchronize(nullptr);
    }
    used_ -= it->second->bytes;
    cache_.erase(it);
    return kOk;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> g(mu_);
    ServerStats s = stats_;
    s.budget = budget_;
    s.used = used_;
    s.resident = (int)cache_.size();
    return s;
}

// Pick the first queued request that can start now (mu_ held).  On success
// the request's memory is reserved and its factor entry pinned.
std::shared_ptr<Server::Req> Server::admit(std::shared_ptr<Factor>& fac, bool& build, hipStream_t s) {
    for (auto qi = queue_.begin(); qi != queue_.end(); ++qi) {
        auto r = *qi;
        const ServeRequest& q = r->q;
        auto ci = cache_.find(q.key);
        const bool resident = ci != cache_.end() && ci->second->n == q.n && !ci->second->failed;
        if (q.reuse && resident && !ci->second->ready) continue;  // being factored: wait for it
        const bool need_factor = !(q.reuse && resident);
        if (need_factor && ci != cache_.end() && (ci->second->pins > 0 || !ci->second->ready))
            continue;  // the key's current factor is in use: do not replace it now
        const size_t fbytes = need_factor ? factor_bytes(q.n) : 0;
        const size_t wbytes = work_bytes(q.n, q.k, need_factor);
        size_t freed = (need_factor && ci != cache_.end()) ? ci->second->bytes : 0;  // replaced entry
        size_t need = fbytes + wbytes;
        if (used_ - freed + need > budget_) {
            // evict least-recently-used unpinned factors (not the one this request uses)
            std::vector<std::pair<uint64_t, int>> lru;
            for (auto& kv : cache_)
                if (kv.second->pins == 0 && kv.second->ready && kv.first != q.key)
                    lru.push_back({kv.second->last_use, kv.first});
            std::sort(lru.begin(), lru.end());
            size_t can = 0;
            for (auto& e : lru) can += cache_[e.second]->bytes;
            if (used_ - freed - can + need > budget_) continue;  // cannot fit even after evictions
            for (auto& e : lru) {
                if (used_ - freed + need <= budget_) break;
                auto& F = cache_[e.second];
                if (F->mem) (void)hipFreeAsync(F->mem, s);
                used_ -= F->bytes;
                cache_.erase(e.second);
                ++stats_.evictions;
            }
        }
        // reserve
        if (need_factor) {
            if (ci != cache_.end()) {
                if (ci->second->mem) (void)hipFreeAsync(ci->second->mem, s);
                used_ -= ci->second->bytes;
                cache_.erase(ci);
            }
            auto F = std::make_shared<Factor>();
            F->key = q.key;
            F->n = q.n;
            F->bytes = fbytes;
            cache_[q.key] = F;
            fac = F;
            build = true;
        } else {
            fac = ci->second;
            build = false;
            ++stats_.reuse_hits;
        }
        fac->pins++;
        fac->last_use = ++use_clock_;
        used_ += fbytes + wbytes;
        r->wbytes = wbytes;
        stats_.peak_used = std::max(stats_.peak_used, used_);
        queue_.erase(qi);
        return r;
    }
    return nullptr;
}

void Server::worker(int slot) {
    (void)hipSetDevice(device_);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    MixedOpts o;
    o.nb = cfg_.nb;
    o.trailing = cfg_.precision == 2 ? kTrailingBf16x3 : kTrailingF32;
    o.tol = cfg_.tol;
    o.fallback = 0;  // a served request reports its own status
    for (;;) {
        std::shared_ptr<Req> r;
        std::shared_ptr<Factor> F;
        bool build = false;
        {
            std::unique_lock<std::mutex> g(mu_);
            cv_.wait(g, [&] {
                if (stop_ && queue_.empty()) return true;
                r = admit(F, build, s);
                return r != nullptr;
            });
            if (!r) break;
            ++busy_;
        }
        const ServeRequest& q = r->q;
        ServeResult& res = r->res;
        res.t_start = now_s(t0_);
        res.slot = slot;
        res.factored = build ? 1 : 0;
        int rc = s ? kOk : kErrHip;
        if (rc == kOk && r->ready != nullptr && hipStreamWaitEvent(s, r->ready, 0) != hipSuccess) rc = kErrHip;
        // exactly the reserved workspace, stream ordered, released when the request ends
        const size_t wbytes = r->wbytes;
        void* work = nullptr;
        if (rc == kOk && hipMallocAsync(&work, wbytes, s) != hipSuccess) {
            work = nullptr;
            rc = kErrHip;
        }
        if (rc == kOk && build && hipMallocAsync(&F->mem, F->bytes, s) != hipSuccess) {
            F->mem = nullptr;
            rc = kErrHip;
        }
        if (cfg_.precision == 0) {
            double* L = (double*)F->mem;
            int* dinfo = (int*)((char*)work + al256(solve_work_bytes<double>(q.n, q.k)));
            if (rc == kOk && build) {
                // private copy of A, factored in place and kept resident
                if (hipMemcpy2DAsync(L, sizeof(double) * q.n, q.A, sizeof(double) * q.lda, sizeof(double) * q.n, q.n,
                                     hipMemcpyDeviceToDevice, s) != hipSuccess)
                    rc = kErrHip;
                if (rc == kOk) rc = cholesky_lower<double>(q.n, L, q.n, cfg_.nb, kTrailingF64, dinfo, s);
                int info = 0;
                if (rc == kOk && hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
                    rc = kErrHip;
                if (rc == kOk && hipStreamSynchronize(s) != hipSuccess) rc = kErrHip;
                res.info = info;
                if (rc == kOk && info != 0) rc = kServeNotHpd;
            }
            if (rc == kOk) rc = solve_after_lower<double>(q.n, q.k, L, q.n, q.B, q.ldb, nullptr, work, s, false, true);
            // synchronizes; a solve whose bounded wait gave up is an error
            if (rc == kOk && q.n > 0 && q.k > 0) rc = solve_read_error(solve_ctrl<double>(q.n, work), s);
            if (rc == kOk && hipStreamSynchronize(s) != hipSuccess) rc = kErrHip;
        } else {
            if (rc == kOk && build) {
                F->mf = carve_factor(F->mem, q.n);
                // the factorization's plane scratch follows the refinement's vectors
                rc = mixed_factor(q.n, q.A, q.lda, o, F->mf, s, (char*)work + al256(mixed_refine_bytes(q.n, q.k, o)));
                res.info = F->mf.lowprec_info;
            }
            MixedReport rep{};
            if (rc == kOk)
                rc = mixed_refine(q.n, q.k, q.A, q.lda, q.B, q.ldb, o, F->mf, work, mixed_refine_bytes(q.n, q.k, o),
                                  &rep, s);
            res.iterations = rep.iterations;
            res.gmres_iterations = rep.gmres_iterations;
            res.backward_error = rep.backward_error;
        }
        if (work) (void)hipFreeAsync(work, s);
        (void)hipStreamSynchronize(s);
        if (r->ready != nullptr) (void)hipEventDestroy(r->ready);
        res.status = rc;
        res.t_done = now_s(t0_);
        {
            std::lock_guard<std::mutex> g(mu_);
            used_ -= wbytes;
            F->pins--;
            if (build) {
                F->ready = true;
                F->failed = rc != kOk;
            }
            if (F->failed && F->pins == 0) {  // drop a failed factor
                auto ci = cache_.find(F->key);
                if (ci != cache_.end() && ci->second == F) {
                    if (F->mem) (void)hipFreeAsync(F->mem, s);
                    used_ -= F->bytes;
                    cache_.erase(ci);
                }
            }
            ++stats_.completed;
            --busy_;
            r->done = true;
        }
        done_cv_.notify_all();
        cv_.notify_all();  // memory and a slot were released
    }
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
}

}  // namespace elemamd
