// extern "C" entry points declared in include/elemental_amd.h.
#include "../../include/elemental_amd.h"

#include <algorithm>

#include "batch.hpp"
#include "kernels.hpp"
#include "lapack.hpp"
#include "runtime.hpp"
#include "server.hpp"

using namespace elemamd;

struct elemamd_batch_plan {
    BatchPlan* p;
};
struct elemamd_server {
    Server* s;
};

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }
inline bool is_trans(char c) { c = up(c); return c == 'T' || c == 'C'; }
inline bool valid_orient(char c) { c = up(c); return c == 'N' || c == 'T' || c == 'C'; }

// One triangular solve with a lower L (forward, or backward with L^T) on
// the calling stream; the per-stream scratch holds the persistent-solve flags.
int solve_left_lower(bool trans, int n, int k, const double* L, i64 ldl, double* B, i64 ldb, hipStream_t s) {
    void* work = stream_scratch(solve_work_bytes<double>(n, k), s);
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    int* ctrl = (int*)((char*)work + ((solve_dinv_count(n) * sizeof(double) + 255) & ~(size_t)255));
    return solve_apply<double>(!trans, trans, n, k, L, ldl, B, ldb, nullptr, ctrl, nullptr, s);
}

}  // namespace

extern "C" {

const char* elemamd_version(void) { return "elemental_amd 0.1 (gfx950)"; }

int elemamd_info_read(const int* dinfo, void* stream) {
    if (dinfo == nullptr) return ELEMAMD_ERR_ARG;
    int v = 0;
    if (hipMemcpyAsync(&v, dinfo, sizeof(int), hipMemcpyDeviceToHost, S(stream)) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (hipStreamSynchronize(S(stream)) != hipSuccess) return ELEMAMD_ERR_HIP;
    return v;
}

int elemamd_potrf_unb_f64(int n, double* A, int64_t lda, int* dinfo, void* stream) {
    if (n < 0 || n > 128 || lda < std::max(1, n) || (n > 0 && A == nullptr)) return ELEMAMD_ERR_ARG;
    return launch_potrf_unb<double>(n, A, lda, dinfo, 0, S(stream)) == hipSuccess ? ELEMAMD_OK : ELEMAMD_ERR_HIP;
}

int elemamd_trsm_f64(char side, char uplo, char trans, char diag, int m, int n, double alpha, const double* A,
                     int64_t lda, double* B, int64_t ldb, void* stream) {
    side = up(side);
    uplo = up(uplo);
    diag = up(diag);
    if (uplo != 'L' || diag != 'N' || !valid_orient(trans)) return ELEMAMD_ERR_ARG;
    if (m < 0 || n < 0 || ldb < std::max(1, m)) return ELEMAMD_ERR_ARG;
    const int na = (side == 'L') ? m : n;
    if (lda < std::max(1, na)) return ELEMAMD_ERR_ARG;
    if (m == 0 || n == 0) return ELEMAMD_OK;
    if (alpha != 1.0 && launch_scale_f64(m, n, alpha, B, ldb, S(stream)) != hipSuccess) return ELEMAMD_ERR_HIP;
    if (side == 'R') {
        if (!is_trans(trans) || n > 128) return ELEMAMD_ERR_ARG;
        return launch_trsm_rlt<double>(m, n, A, lda, B, ldb, nullptr, kNoLimit, S(stream)) == hipSuccess ? ELEMAMD_OK
                                                                                               : ELEMAMD_ERR_HIP;
    }
    if (side != 'L') return ELEMAMD_ERR_ARG;
    return solve_left_lower(is_trans(trans), m, n, A, lda, B, ldb, S(stream));
}

int elemamd_herk_f64(char uplo, char trans, int n, int k, double alpha, const double* A, int64_t lda, double beta,
                     double* C, int64_t ldc, void* stream) {
    if (up(uplo) != 'L' || up(trans) != 'N') return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldc < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    return launch_syrk_ln_f64(n, n, k, alpha, A, lda, A, lda, beta, C, ldc, nullptr, 0, 1, S(stream)) == hipSuccess
               ? ELEMAMD_OK
                                                                                                  : ELEMAMD_ERR_HIP;
}

int elemamd_herk_f32(int mode, int n, int k, float alpha, const float* A, int64_t lda, float beta, float* C,
                     int64_t ldc, void* stream) {
    if (mode != ELEMAMD_TRAILING_F32 && mode != ELEMAMD_TRAILING_BF16X3 && mode != ELEMAMD_TRAILING_BF16X6)
        return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldc < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    hipStream_t s = S(stream);
    if (mode != ELEMAMD_TRAILING_F32 && k > 0) {
        // the factorization's form: split A into bf16 planes once, then the planes update
        const int np = mode == ELEMAMD_TRAILING_BF16X3 ? 2 : 3;
        const i64 ldk = ((i64)k + 31) / 32 * 32;
        const i64 ps = (i64)n * ldk;
        auto* planes = (unsigned short*)stream_scratch((size_t)np * ps * sizeof(unsigned short), s);
        if (planes == nullptr) return ELEMAMD_ERR_HIP;
        if (launch_split_bf16(np, n, k, A, lda, planes, ps, ldk, s) != hipSuccess) return ELEMAMD_ERR_HIP;
        return launch_syrk_ln_bf16_planes(np, n, n, k, alpha, planes, ps, ldk, beta, C, ldc, nullptr, 0, 1, s) ==
                       hipSuccess
                   ? ELEMAMD_OK
                   : ELEMAMD_ERR_HIP;
    }
    return launch_syrk_ln_f32(false, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, nullptr, 0, 1, s) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_trrk_f64(char uplo, char orientA, char orientB, int n, int k, double alpha, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream) {
    if (up(uplo) != 'L' || !is_trans(orientA) || up(orientB) != 'N') return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0 || lda < std::max(1, k) || ldb < std::max(1, k) || ldc < std::max(1, n))
        return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    return launch_gemm_tn_f64(n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, true, S(stream)) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_trsv_f64(char uplo, char trans, char diag, int n, const double* A, int64_t lda, double* x, int64_t incx,
                     void* stream) {
    if (up(uplo) != 'L' || up(diag) != 'N' || !valid_orient(trans) || incx != 1) return ELEMAMD_ERR_ARG;
    if (n < 0 || lda < std::max(1, n)) return ELEMAMD_ERR_ARG;
    if (n == 0) return ELEMAMD_OK;
    return solve_left_lower(is_trans(trans), n, 1, A, lda, x, std::max(1, n), S(stream));
}

int elemamd_cholesky_f64(char uplo, int n, double* A, int64_t lda, int nb, int* dinfo, void* stream) {
    if (up(uplo) != 'L') return ELEMAMD_ERR_ARG;
    return cholesky_lower<double>(n, A, lda, nb, kTrailingF64, dinfo, S(stream));
}

int elemamd_solve_after_f64(char uplo, char orientation, int n, int k, const double* A, int64_t lda, double* B,
                            int64_t ldb, void* stream) {
    if (up(uplo) != 'L' || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (n < 0 || k < 0) return ELEMAMD_ERR_ARG;
    void* work = stream_scratch(solve_work_bytes<double>(std::max(n, 1), std::max(k, 1)), S(stream));
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    return solve_after_lower<double>(n, k, A, lda, B, ldb, nullptr, work, S(stream));
}

int elemamd_hpd_solve_f64(char uplo, char orientation, int n, int k, double* A, int64_t lda, double* B, int64_t ldb,
                          int nb, int* dinfo, void* stream) {
    if (up(uplo) != 'L' || !valid_orient(orientation)) return ELEMAMD_ERR_ARG;
    if (k < 0 || ldb < std::max(1, n)) return ELEMAMD_ERR_ARG;
    void* work = stream_scratch(solve_work_bytes<double>(std::max(n, 1), std::max(k, 1)), S(stream));
    if (work == nullptr) return ELEMAMD_ERR_HIP;
    int rc = cholesky_lower<double>(n, A, lda, nb, kTrailingF64, dinfo, S(stream));
    if (rc) return rc;
    return solve_after_lower<double>(n, k, A, lda, B, ldb, dinfo, work, S(stream));
}

static MixedOpts to_opts(const elemamd_ir_opts* o) {
    MixedOpts m;
    if (o) {
        m.nb = o->nb > 0 ? o->nb : 128;
        m.trailing = o->trailing;
        m.max_iter = o->max_iter > 0 ? o->max_iter : 30;
        m.gmres_restart = o->gmres_restart >= 0 ? (o->gmres_restart == 0 ? 20 : o->gmres_restart) : 0;
        m.force_gmres = o->force_gmres;
        m.allow_shift = o->allow_shift;
        m.tol = o->tol > 0.0 ? o->tol : 1e-14;
        m.refine = o->refine;
    }
    return m;
}

size_t elemamd_hpd_solve_mixed_workspace(int n, int k, const elemamd_ir_opts* opts) {
    if (n < 0 || k < 0) return 0;
    return mixed_workspace_bytes(n, k, to_opts(opts));
}

int elemamd_hpd_solve_mixed_f64(int n, int k, const double* A, int64_t lda, double* B, int64_t ldb,
                                const elemamd_ir_opts* opts, void* work, size_t work_bytes, elemamd_ir_report* report,
                                void* stream) {
    MixedOpts o = to_opts(opts);
    if (o.trailing != kTrailingF32 && o.trailing != kTrailingBf16x3 && o.trailing != kTrailingBf16x6)
        return ELEMAMD_ERR_ARG;
    if (o.refine != ELEMAMD_REFINE_IR && o.refine != ELEMAMD_REFINE_CG) return ELEMAMD_ERR_ARG;
    MixedReport r{};
    int rc = hpd_solve_mixed(n, k, A, lda, B, ldb, o, work, work_bytes, &r, S(stream));
    if (report) {
        report->info = r.info;
        report->lowprec_info = r.lowprec_info;
        report->iterations = r.iterations;
        report->gmres_iterations = r.gmres_iterations;
        report->used_gmres = r.used_gmres;
        report->shifted = r.shifted;
        report->converged = r.converged;
        report->backward_error = r.backward_error;
        report->anorm_inf = r.anorm_inf;
    }
    return rc;
}

int elemamd_fill_normal_f64(double* out, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, int64_t offset,
                            void* stream) {
    if (rows < 0 || cols < 0 || ld < std::max<int64_t>(1, rows)) return ELEMAMD_ERR_ARG;
    return launch_fill_normal_f64(out, rows, cols, ld, seed, offset, S(stream)) == hipSuccess ? ELEMAMD_OK
                                                                                             : ELEMAMD_ERR_HIP;
}

int elemamd_sym_residual_f64(int n, int k, const double* A, int64_t lda, const double* X, int64_t ldx,
                             const double* B, int64_t ldb, double* R, int64_t ldr, void* stream) {
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldx < std::max(1, n) || ldb < std::max(1, n) ||
        ldr < std::max(1, n))
        return ELEMAMD_ERR_ARG;
    double* part = nullptr;
    const size_t pc = symm_part_count(n, k);
    if (pc > 0 && (part = (double*)stream_scratch(pc * sizeof(double), S(stream))) == nullptr) return ELEMAMD_ERR_HIP;
    return launch_symm_lower_residual(n, k, A, lda, X, ldx, B, ldb, R, ldr, part, S(stream)) == hipSuccess
               ? ELEMAMD_OK
               : ELEMAMD_ERR_HIP;
}

int elemamd_batch_plan_create(int count, const int* n, const int* k, int precision, int nbuckets,
                              const int* bucket_max_n, const int* bucket_nb, elemamd_batch_plan** plan) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    *plan = nullptr;
    BatchPlan* P = nullptr;
    const int rc = batch_plan_create(count, n, k, precision, nbuckets, bucket_max_n, bucket_nb, &P);
    if (rc != kOk) return rc;
    *plan = new elemamd_batch_plan{P};
    return ELEMAMD_OK;
}

int elemamd_batch_plan_destroy(elemamd_batch_plan* plan) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    delete plan->p;
    delete plan;
    return ELEMAMD_OK;
}

int elemamd_batch_plan_layout(const elemamd_batch_plan* plan, int64_t* a_off, int64_t* b_off, int64_t* a_total,
                              int64_t* b_total) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    const BatchPlan* P = plan->p;
    for (int i = 0; i < P->count; ++i) {
        if (a_off) a_off[i] = P->a_off[i];
        if (b_off) b_off[i] = P->b_off[i];
    }
    if (a_total) *a_total = P->a_total;
    if (b_total) *b_total = P->b_total;
    return ELEMAMD_OK;
}

int elemamd_batch_set_refinement(elemamd_batch_plan* plan, int sweeps, double tol) {
    if (plan == nullptr || sweeps < 0 || !(tol > 0.0)) return ELEMAMD_ERR_ARG;
    BatchPlan* P = plan->p;
    if (P->ir_sweeps != sweeps || P->tol != tol) {
        P->ir_sweeps = sweeps;
        P->tol = tol;
        P->g_A = nullptr;  // the captured sequence no longer matches: capture again
    }
    return ELEMAMD_OK;
}

size_t elemamd_batch_workspace(const elemamd_batch_plan* plan) {
    return plan == nullptr ? 0 : batch_workspace_bytes(plan->p);
}

int elemamd_batch_hpd_solve(elemamd_batch_plan* plan, double* A, double* B, void* work, size_t work_bytes,
                            int use_graph, void* stream) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    return batch_hpd_solve(plan->p, A, B, work, work_bytes, use_graph != 0, S(stream));
}

int elemamd_batch_results(const elemamd_batch_plan* plan, int* info, double* backward_error, int* iterations,
                          void* stream) {
    if (plan == nullptr) return ELEMAMD_ERR_ARG;
    return batch_results(plan->p, info, backward_error, iterations, S(stream));
}

int elemamd_batch_backward_error(const elemamd_batch_plan* plan, const double* A, const double* X, const double* B,
                                 double* R, double* eta, void* stream) {
    if (plan == nullptr || A == nullptr || X == nullptr || B == nullptr || R == nullptr) return ELEMAMD_ERR_ARG;
    return batch_backward_error(plan->p, A, X, B, R, eta, S(stream));
}

int elemamd_server_create(const elemamd_server_config* cfg, elemamd_server** server) {
    if (cfg == nullptr || server == nullptr) return ELEMAMD_ERR_ARG;
    *server = nullptr;
    ServerConfig c;
    c.slots = cfg->slots;
    c.precision = cfg->precision;
    c.nb = cfg->nb > 0 ? cfg->nb : 128;
    c.tol = cfg->tol > 0.0 ? cfg->tol : 1e-14;
    c.memory_budget = cfg->memory_budget;
    auto* S = new Server(c);
    const int rc = S->start();
    if (rc != kOk) {
        delete S;
        return rc;
    }
    *server = new elemamd_server{S};
    return ELEMAMD_OK;
}

int elemamd_server_submit(elemamd_server* server, int key, int reuse, int n, int k, const double* A, int64_t lda,
                          double* B, int64_t ldb, void* ready_stream, int64_t* ticket) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    ServeRequest q{key, reuse, n, k, A, lda, B, ldb, S(ready_stream)};
    return server->s->submit(q, ticket);
}

int elemamd_server_wait(elemamd_server* server, int64_t ticket, elemamd_serve_result* result) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    ServeResult r{};
    const int rc = server->s->wait(ticket, &r);
    if (rc != kOk) return rc;
    if (result) {
        result->ticket = r.ticket;
        result->status = r.status;
        result->info = r.info;
        result->factored = r.factored;
        result->iterations = r.iterations;
        result->gmres_iterations = r.gmres_iterations;
        result->slot = r.slot;
        result->backward_error = r.backward_error;
        result->t_submit = r.t_submit;
        result->t_start = r.t_start;
        result->t_done = r.t_done;
    }
    return ELEMAMD_OK;
}

int elemamd_server_evict(elemamd_server* server, int key) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    return server->s->evict(key);
}

int elemamd_server_stats(elemamd_server* server, elemamd_server_statistics* stats) {
    if (server == nullptr || stats == nullptr) return ELEMAMD_ERR_ARG;
    const ServerStats t = server->s->stats();
    stats->completed = t.completed;
    stats->reuse_hits = t.reuse_hits;
    stats->evictions = t.evictions;
    stats->peak_reserved = t.peak_used;
    stats->reserved = t.used;
    stats->budget = t.budget;
    stats->resident_factors = t.resident;
    return ELEMAMD_OK;
}

int elemamd_server_destroy(elemamd_server* server) {
    if (server == nullptr) return ELEMAMD_ERR_ARG;
    delete server->s;
    delete server;
    return ELEMAMD_OK;
}

}  // extern "C"
