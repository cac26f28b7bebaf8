// Batched HPD factor-and-solve over many matrices of uneven size: the plan
// (size buckets with their own block size, processing order, per-step work
// prefix sums, device descriptors) and its replayable launch sequence.
#pragma once
#include <vector>

#include "common.hpp"

namespace elemamd {

// Per-launch view of a group of matrices (processing order): sizes,
// right-hand-side counts, element offsets into the packed buffers, factor
// status and the refinement's "converged" flags (may be null).
constexpr int kPersistMaxNb = 128;  // largest block size of the persistent launch

struct BatchArgs {
    int count;
    const int* n;
    const int* k;
    const i64* a_off;
    const i64* b_off;
    int* info;
    int* skip;
};

hipError_t batch_potrf(const BatchArgs& a, void* A, bool f32, int step, int nb, int cnt, hipStream_t s);
hipError_t batch_trsm(const BatchArgs& a, void* A, bool f32, int step, int nb, const int* prefix, int cnt, int wgs,
                      hipStream_t s);
hipError_t batch_syrk(const BatchArgs& a, void* A, int mode, int step, int nb, const int* prefix, int cnt, int wgs,
                      hipStream_t s);
hipError_t batch_solve(const BatchArgs& a, const void* L, void* B, bool f32, hipStream_t s);
// The bucket's whole factorization (mode 0 FP64, 1 FP32, 2 BF16x3 trailing;
// F = A, or the float copy) and, for FP64, its solves into B, as one launch
// of `grid` workgroups that take the matrices from *counter (zeroed before).
// nb <= kPersistMaxNb.
hipError_t batch_persistent(const BatchArgs& a, void* F, double* B, int mode, int nb, int* counter, int grid,
                            hipStream_t s);
hipError_t batch_to_float(const BatchArgs& a, const double* A, float* S, hipStream_t s);
// R := B - A X per matrix; eta (and rscale = ||r_c||_inf when non-null) per
// matrix; prefix/wgs: 64-row blocks of the matrices; red: count*(1+3*kmax)
// doubles of reduction scratch.
hipError_t batch_residual(const BatchArgs& a, const double* A, const double* X, const double* B, double* R,
                          const int* prefix, int wgs, double* red, double* eta, double* rscale, int kmax,
                          hipStream_t s);
hipError_t batch_ir_prepare(const BatchArgs& a, const double* eta, double tol, const double* R, const double* rscale,
                            float* Rf, int kmax, int* iters, hipStream_t s);
hipError_t batch_ir_update(const BatchArgs& a, const float* D, const double* rscale, double* X, int kmax,
                           hipStream_t s);
hipError_t batch_copy_rhs(const BatchArgs& a, const double* src, double* dst, float* fdst, hipStream_t s);


struct BatchBucket {
    int index = 0;  // position in the caller's bucket arrays
    int nb = 128;
    int grid = 0;   // 0: one launch per step and kernel; > 0: one persistent launch of grid workgroups
    int first = 0, count = 0;  // processing positions [first, first + count)
    int steps = 0;
    int res_prefix_off = 0, res_wgs = 0;  // 64-row blocks of the residual
    struct Step {
        int active, trsm_wgs, syrk_wgs, trsm_prefix_off, syrk_prefix_off;
    };
    std::vector<Step> step;
};

struct BatchPlan {
    int count = 0;
    int precision = 0;   // 0 FP64, 1 FP32 factor + FP64 refinement, 2 BF16x3 trailing + FP64 refinement
    int kmax = 0;
    int ir_sweeps = 4;   // fixed number of refinement sweeps (graph-replayable)
    double tol = 1e-14;  // converged matrices skip the remaining sweeps
    std::vector<int> n, k;          // caller order
    std::vector<i64> a_off, b_off;  // caller order, packed (ld = n_i)
    i64 a_total = 0, b_total = 0;
    std::vector<int> order;         // processing position -> caller index
    std::vector<BatchBucket> buckets;
    // device descriptors (processing order)
    int *d_n = nullptr, *d_k = nullptr, *d_prefix = nullptr;
    i64 *d_aoff = nullptr, *d_boff = nullptr;
    int *d_info = nullptr, *d_skip = nullptr, *d_iters = nullptr;
    int* d_counters = nullptr;  // one per bucket: the persistent launches' matrix counters
    double *d_eta = nullptr, *d_rscale = nullptr, *d_red = nullptr;
    std::vector<void*> dev_allocs;
    // captured graph and the arguments it was captured with
    hipGraphExec_t graph_exec = nullptr;
    const void *g_A = nullptr, *g_B = nullptr, *g_work = nullptr;
    hipStream_t capture_stream = nullptr;
    ~BatchPlan();
};

// Buckets: matrix i goes to the first bucket b with n_i <= bucket_max_n[b]
// (the last bucket takes every larger matrix); bucket b factors with block
// size bucket_nb[b] (1..128).
int batch_plan_create(int count, const int* n, const int* k, int precision, int nbuckets, const int* bucket_max_n,
                      const int* bucket_nb, BatchPlan** out);
// Launch form per bucket (caller's bucket order): 0 = one launch per step
// and kernel, g > 0 = one persistent launch of g workgroups.
int batch_set_launch(BatchPlan* P, const int* bucket_grid);
size_t batch_workspace_bytes(const BatchPlan* P);
int batch_hpd_solve(BatchPlan* P, double* A, double* B, void* work, size_t work_bytes, bool use_graph,
                    hipStream_t s);
int batch_results(const BatchPlan* P, int* info, double* eta, int* iters, hipStream_t s);
int batch_backward_error(const BatchPlan* P, const double* A, const double* X, const double* B, double* R,
                         double* eta, hipStream_t s);

}  // namespace elemamd
