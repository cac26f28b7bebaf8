// Serving loop of factor-and-solve requests (see server.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace elemamd {

constexpr int kServeNotHpd = -5;  // the FP64 factorization met a non-positive pivot

struct ServerConfig {
    int slots = 4;             // requests in flight at once (one stream each)
    int precision = 0;         // 0 plain FP64, 1 FP32 factor + FP64 refinement, 2 BF16x3 trailing + refinement
    int nb = 128;
    double tol = 1e-14;        // refinement target (normwise backward error)
    size_t memory_budget = 0;  // bytes; 0 = free HBM at start minus a reserve
};

struct ServeRequest {
    int key;          // factor key: requests with the same key share a resident factor
    int reuse;        // 1: solve with the resident factor of `key` when there is one
    int n, k;
    const double* A;  // device, n x n, lower triangle read (never modified)
    i64 lda;
    double* B;        // device, n x k, overwritten with X
    i64 ldb;
    hipStream_t ready;  // the inputs are ready once this stream's work so far completes (null: default stream)
};

struct ServeResult {
    int64_t ticket;
    int status;        // 0, or a negative error code (kServeNotHpd: info holds the column)
    int info;          // factorization status (1-based failing column, 0 = none)
    int factored;      // 1 if this request computed the factor, 0 if it reused a resident one
    int iterations;    // refinement steps (low-precision modes)
    int gmres_iterations;
    int slot;
    double backward_error;  // low-precision modes: final normwise backward error
    double t_submit, t_start, t_done;  // seconds since the server started
};

struct ServerStats {
    int64_t completed;
    int64_t reuse_hits;
    int64_t evictions;
    size_t peak_used;  // largest reserved total seen
    size_t used;       // currently reserved (resident factors + running requests)
    size_t budget;
    int resident;      // factors held
};

class Server {
public:
    explicit Server(const ServerConfig& cfg);
    ~Server();  // drains the queue, then frees every resident factor
    int start();
    int submit(const ServeRequest& q, int64_t* ticket);
    int wait(int64_t ticket, ServeResult* out);
    int evict(int key);
    ServerStats stats();

private:
    struct Factor;
    struct Req;
    size_t factor_bytes(int n) const;
    size_t work_bytes(int n, int k, bool factor) const;
    std::shared_ptr<Req> admit(std::shared_ptr<Factor>& fac, bool& build, hipStream_t s);
    void worker(int slot);

    ServerConfig cfg_;
    std::chrono::steady_clock::time_point t0_;
    int device_ = 0;
    size_t budget_ = 0, used_ = 0;
    uint64_t use_clock_ = 0;
    int64_t next_ticket_ = 1;
    int busy_ = 0;
    bool stop_ = false;
    ServerStats stats_{};
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Req>> queue_;
    std::map<int64_t, std::shared_ptr<Req>> reqs_;
    std::map<int, std::shared_ptr<Factor>> cache_;
    std::vector<std::thread> workers_;
};

}  // namespace elemamd
