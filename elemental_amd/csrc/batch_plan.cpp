// Host side of the batched HPD factor-and-solve: the plan (size buckets,
// processing order, per-step work prefix sums, device descriptors), the
// launch sequence, its hipGraph capture and replay, and per-matrix results.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "kernels.hpp"

namespace elemamd {

namespace {

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& v) {
    hipError_t e = hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
    return e;
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// descriptors of one bucket (positions [first, first + count))
BatchArgs bucket_args(const BatchPlan* P, const BatchBucket& b, bool with_skip) {
    return BatchArgs{b.count,
                     P->d_n + b.first,
                     P->d_k + b.first,
                     P->d_aoff + b.first,
                     P->d_boff + b.first,
                     P->d_info + b.first,
                     with_skip ? P->d_skip + b.first : nullptr};
}

}  // namespace

BatchPlan::~BatchPlan() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    for (void* p : dev_allocs) (void)hipFree(p);
}

int batch_plan_create(int count, const int* n, const int* k, int precision, int nbuckets, const int* bucket_max_n,
                      const int* bucket_nb, BatchPlan** out) {
    if (out == nullptr) return kErrArg;
    *out = nullptr;
    if (count < 0 || (count > 0 && (n == nullptr || k == nullptr))) return kErrArg;
    if (precision < 0 || precision > 2) return kErrArg;
    if (nbuckets < 1 || bucket_max_n == nullptr || bucket_nb == nullptr) return kErrArg;
    for (int b = 0; b < nbuckets; ++b) {
        if (bucket_nb[b] < 1 || bucket_nb[b] > 128) return kErrArg;
        if (b > 0 && bucket_max_n[b] <= bucket_max_n[b - 1]) return kErrArg;
    }
    for (int i = 0; i < count; ++i)
        if (n[i] < 1 || k[i] < 1) return kErrArg;
    auto* P = new BatchPlan();
    P->count = count;
    P->precision = precision;
    P->n.assign(n, n + count);
    P->k.assign(k, k + count);
    // packed layout in the caller's order
    P->a_off.resize(count);
    P->b_off.resize(count);
    i64 ao = 0, bo = 0;
    for (int i = 0; i < count; ++i) {
        P->a_off[i] = ao;
        P->b_off[i] = bo;
        ao += (i64)P->n[i] * P->n[i];
        bo += (i64)P->n[i] * P->k[i];
        P->kmax = std::max(P->kmax, P->k[i]);
    }
    P->a_total = ao;
    P->b_total = bo;
    // bucket of each matrix; processing order = by bucket, then by decreasing
    // size, so that the matrices still active at step j form a prefix
    std::vector<int> bk(count);
    for (int i = 0; i < count; ++i) {
        int b = 0;
        while (b < nbuckets - 1 && P->n[i] > bucket_max_n[b]) ++b;
        bk[i] = b;
    }
    P->order.resize(count);
    std::iota(P->order.begin(), P->order.end(), 0);
    std::stable_sort(P->order.begin(), P->order.end(),
                     [&](int a, int b) { return bk[a] != bk[b] ? bk[a] < bk[b] : P->n[a] > P->n[b]; });
    std::vector<int> ns(count), ks(count);
    std::vector<i64> aos(count), bos(count);
    for (int p = 0; p < count; ++p) {
        const int i = P->order[p];
        ns[p] = P->n[i];
        ks[p] = P->k[i];
        aos[p] = P->a_off[i];
        bos[p] = P->b_off[i];
    }
    std::vector<int> prefix_all;
    int pos = 0;
    for (int b = 0; b < nbuckets; ++b) {
        BatchBucket B;
        B.index = b;
        B.nb = bucket_nb[b];
        B.first = pos;
        while (pos < count && bk[P->order[pos]] == b) ++pos;
        B.count = pos - B.first;
        if (B.count == 0) continue;
        const int nb = B.nb;
        B.steps = (ns[B.first] + nb - 1) / nb;
        for (int j = 0; j < B.steps; ++j) {
            BatchBucket::Step st;
            st.active = 0;
            while (st.active < B.count && ns[B.first + st.active] > j * nb) ++st.active;
            std::vector<int> pt(st.active + 1, 0), ps(st.active + 1, 0);
            for (int q = 0; q < st.active; ++q) {
                const int m = ns[B.first + q] - (j + 1) * nb;
                const i64 T = m > 0 ? (m + 127) / 128 : 0;
                pt[q + 1] = pt[q] + (m > 0 ? (m + 63) / 64 : 0);  // panel: 64 rows per workgroup
                ps[q + 1] = ps[q] + (int)(T * (T + 1) / 2);       // trailing: lower 128x128 tiles
            }
            st.trsm_wgs = pt[st.active];
            st.syrk_wgs = ps[st.active];
            st.trsm_prefix_off = (int)prefix_all.size();
            prefix_all.insert(prefix_all.end(), pt.begin(), pt.end());
            st.syrk_prefix_off = (int)prefix_all.size();
            prefix_all.insert(prefix_all.end(), ps.begin(), ps.end());
            B.step.push_back(st);
        }
        std::vector<int> pr(B.count + 1, 0);
        for (int q = 0; q < B.count; ++q) pr[q + 1] = pr[q] + (ns[B.first + q] + 63) / 64;
        B.res_wgs = pr[B.count];
        B.res_prefix_off = (int)prefix_all.size();
        prefix_all.insert(prefix_all.end(), pr.begin(), pr.end());
        P->buckets.push_back(B);
    }
    int *d_n = nullptr, *d_k = nullptr, *d_pref = nullptr, *d_ints = nullptr;
    i64 *d_ao = nullptr, *d_bo = nullptr;
    double* d_dbl = nullptr;
    bool ok = upload(&d_n, ns) == hipSuccess;
    if (d_n) P->dev_allocs.push_back(d_n);
    ok = ok && upload(&d_k, ks) == hipSuccess;
    if (d_k) P->dev_allocs.push_back(d_k);
    ok = ok && upload(&d_ao, aos) == hipSuccess;
    if (d_ao) P->dev_allocs.push_back(d_ao);
    ok = ok && upload(&d_bo, bos) == hipSuccess;
    if (d_bo) P->dev_allocs.push_back(d_bo);
    ok = ok && upload(&d_pref, prefix_all) == hipSuccess;
    if (d_pref) P->dev_allocs.push_back(d_pref);
    ok = ok && hipMalloc(&d_ints, sizeof(int) * (size_t)(3 * count + nbuckets)) == hipSuccess;
    if (d_ints) P->dev_allocs.push_back(d_ints);
    // eta, rscale (kmax per matrix), reduction scratch (1 + 3 kmax per matrix)
    ok = ok && hipMalloc(&d_dbl, sizeof(double) * (size_t)std::max(count * (2 + 4 * P->kmax), 1)) == hipSuccess;
    if (d_dbl) P->dev_allocs.push_back(d_dbl);
    if (!ok) {
        delete P;
        return kErrHip;
    }
    P->d_n = d_n;
    P->d_k = d_k;
    P->d_aoff = d_ao;
    P->d_boff = d_bo;
    P->d_prefix = d_pref;
    P->d_info = d_ints;
    P->d_skip = d_ints + count;
    P->d_iters = d_ints + 2 * count;
    P->d_counters = d_ints + 3 * count;
    P->d_eta = d_dbl;
    P->d_rscale = d_dbl + count;
    P->d_red = d_dbl + (size_t)count * (1 + P->kmax);
    *out = P;
    return kOk;
}

int batch_set_launch(BatchPlan* P, const int* bucket_grid) {
    if (P == nullptr || bucket_grid == nullptr) return kErrArg;
    for (const BatchBucket& bk : P->buckets)
        if (bucket_grid[bk.index] < 0 || (bucket_grid[bk.index] > 0 && bk.nb > kPersistMaxNb)) return kErrArg;
    for (BatchBucket& bk : P->buckets) bk.grid = bucket_grid[bk.index];
    if (P->graph_exec) {  // captured with the previous launch forms
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    return kOk;
}

size_t batch_workspace_bytes(const BatchPlan* P) {
    if (P->precision == 0) return 0;
    // float copy of every A (the factor), FP64 iterate and residual, float correction
    return al256(sizeof(float) * P->a_total) + 2 * al256(sizeof(double) * P->b_total) +
           al256(sizeof(float) * P->b_total);
}

// The launch sequence: kernels only, fixed shapes, so it can be captured.
static int enqueue(BatchPlan* P, double* A, double* B, void* work, hipStream_t s) {
    if (P->count == 0) return kOk;
    // info, skip, iters, the persistent launches' counters
    ELEMAMD_HIP_TRY(launch_fill_int(P->d_info, 0, 3 * P->count + (int)P->buckets.size(), s));
    const bool f32 = P->precision != 0;
    const int mode = P->precision;  // trailing update: 0 FP64, 1 FP32, 2 BF16x3
    float* S = nullptr;
    double *X = nullptr, *R = nullptr;
    float* Rf = nullptr;
    if (f32) {
        char* w = (char*)work;
        S = (float*)w;
        w += al256(sizeof(float) * P->a_total);
        X = (double*)w;
        w += al256(sizeof(double) * P->b_total);
        R = (double*)w;
        w += al256(sizeof(double) * P->b_total);
        Rf = (float*)w;
    }
    for (const BatchBucket& bk : P->buckets) {
        const BatchArgs a = bucket_args(P, bk, false);
        if (f32) ELEMAMD_HIP_TRY(batch_to_float(a, A, S, s));
        void* F = f32 ? (void*)S : (void*)A;
        if (bk.grid > 0) {  // the whole factorization (and FP64 solves) in one launch
            ELEMAMD_HIP_TRY(batch_persistent(a, F, B, mode, bk.nb, P->d_counters + (&bk - P->buckets.data()),
                                             bk.grid, s));
            if (!f32) continue;
        }
        for (int j = 0; bk.grid == 0 && j < bk.steps; ++j) {
            const BatchBucket::Step& st = bk.step[j];
            ELEMAMD_HIP_TRY(batch_potrf(a, F, f32, j, bk.nb, st.active, s));
            ELEMAMD_HIP_TRY(
                batch_trsm(a, F, f32, j, bk.nb, P->d_prefix + st.trsm_prefix_off, st.active, st.trsm_wgs, s));
            ELEMAMD_HIP_TRY(
                batch_syrk(a, F, mode, j, bk.nb, P->d_prefix + st.syrk_prefix_off, st.active, st.syrk_wgs, s));
        }
        if (!f32) {
            ELEMAMD_HIP_TRY(batch_solve(a, A, B, false, s));
            continue;
        }
        // (the refinement below follows either launch form)
        // low-precision factor + FP64-residual refinement from X = 0: a fixed
        // number of sweeps (the sequence is replayable); a converged matrix
        // is skipped by every kernel of the remaining sweeps
        const BatchArgs ai = bucket_args(P, bk, true);
        double* eta = P->d_eta + bk.first;
        double* rscale = P->d_rscale + (i64)bk.first * P->kmax;
        int* iters = P->d_iters + bk.first;
        double* red = P->d_red + (i64)bk.first * (1 + 3 * P->kmax);
        ELEMAMD_HIP_TRY(batch_copy_rhs(a, nullptr, X, nullptr, s));  // X := 0
        for (int it = 0; it < P->ir_sweeps; ++it) {
            ELEMAMD_HIP_TRY(batch_residual(ai, A, X, B, R, P->d_prefix + bk.res_prefix_off, bk.res_wgs, red, eta,
                                           rscale, P->kmax, s));
            ELEMAMD_HIP_TRY(batch_ir_prepare(ai, eta, P->tol, R, rscale, Rf, P->kmax, iters, s));
            ELEMAMD_HIP_TRY(batch_solve(ai, S, Rf, true, s));
            ELEMAMD_HIP_TRY(batch_ir_update(ai, Rf, rscale, X, P->kmax, s));
        }
        // backward error of every matrix of the bucket, then X -> B
        ELEMAMD_HIP_TRY(batch_residual(a, A, X, B, R, P->d_prefix + bk.res_prefix_off, bk.res_wgs, red, eta, nullptr,
                                       P->kmax, s));
        ELEMAMD_HIP_TRY(batch_copy_rhs(a, X, B, nullptr, s));
    }
    return kOk;
}

int batch_hpd_solve(BatchPlan* P, double* A, double* B, void* work, size_t work_bytes, bool use_graph,
                    hipStream_t s) {
    if (P == nullptr || (P->count > 0 && (A == nullptr || B == nullptr))) return kErrArg;
    if (work_bytes < batch_workspace_bytes(P) || (P->precision != 0 && work == nullptr)) return kErrWorkspace;
    if (!use_graph) return enqueue(P, A, B, work, s);
    if (P->graph_exec == nullptr || P->g_A != A || P->g_B != B || P->g_work != work) {
        if (P->graph_exec) {
            ELEMAMD_HIP_TRY(hipGraphExecDestroy(P->graph_exec));
            P->graph_exec = nullptr;
        }
        // capture on a private stream (the caller's may be the default
        // stream, which cannot be captured); the graph is launched on `s`
        if (P->capture_stream == nullptr)
            ELEMAMD_HIP_TRY(hipStreamCreateWithFlags(&P->capture_stream, hipStreamNonBlocking));
        hipStream_t cs = P->capture_stream;
        hipGraph_t g = nullptr;
        ELEMAMD_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(P, A, B, work, cs);
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc != kOk) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        ELEMAMD_HIP_TRY(e);
        const hipError_t ei = hipGraphInstantiate(&P->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            P->graph_exec = nullptr;
            return kErrHip;
        }
        P->g_A = A;
        P->g_B = B;
        P->g_work = work;
    }
    ELEMAMD_HIP_TRY(hipGraphLaunch(P->graph_exec, s));
    return kOk;
}

int batch_results(const BatchPlan* P, int* info, double* eta, int* iters, hipStream_t s) {
    if (P == nullptr) return kErrArg;
    if (P->count == 0) return kOk;
    std::vector<int> hi(3 * (size_t)P->count);
    std::vector<double> he(P->count);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(hi.data(), P->d_info, sizeof(int) * hi.size(), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipMemcpyAsync(he.data(), P->d_eta, sizeof(double) * he.size(), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    for (int p = 0; p < P->count; ++p) {
        const int i = P->order[p];
        if (info) info[i] = hi[p];
        if (iters) iters[i] = hi[2 * (size_t)P->count + p];
        if (eta) eta[i] = P->precision == 0 ? NAN : he[p];
    }
    return kOk;
}

int batch_backward_error(const BatchPlan* P, const double* A, const double* X, const double* B, double* R,
                         double* eta, hipStream_t s) {
    if (P == nullptr || eta == nullptr) return kErrArg;
    if (P->count == 0) return kOk;
    for (const BatchBucket& bk : P->buckets) {
        const BatchArgs a = bucket_args(P, bk, false);
        ELEMAMD_HIP_TRY(batch_residual(a, A, X, B, R, P->d_prefix + bk.res_prefix_off, bk.res_wgs,
                                       P->d_red + (i64)bk.first * (1 + 3 * P->kmax), P->d_eta + bk.first, nullptr,
                                       P->kmax, s));
    }
    std::vector<double> he(P->count);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(he.data(), P->d_eta, sizeof(double) * he.size(), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    for (int p = 0; p < P->count; ++p) eta[P->order[p]] = he[p];
    return kOk;
}

}  // namespace elemamd
