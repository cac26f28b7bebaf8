// Host-side drivers: the blocked factorization loop, SolveAfter, and the
// mixed-precision iterative-refinement / GMRES-IR solver.  Everything here
// only enqueues kernels on the caller's stream; nothing is global.
#include "lapack.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

namespace elemamd {

// ---------------------------------------------------------------- LVar3
// cholesky::LVar3 (include/elemental/lapack-like/Cholesky/LVar3.hpp:83-123):
// while the bottom-right block is nonempty, take a diagonal block of
// min(nb, remaining) (partition_impl.hpp:403-424), factor it with LVar3Unb,
// solve the panel below it, and apply the symmetric rank-nb update to the
// trailing matrix.
//
// Scheduling on the GPU (same operations, other timing):
//  * Deferred trailing updates.  Block columns are taken in groups of G
//    (G*nb = the update depth).  Inside a group each block column is first
//    brought up to date with the group's earlier panels (a rank-(j*nb)
//    update of that block column only), then factored and its panel solved.
//    The trailing matrix right of the group receives the whole group's panels
//    in one rank-(G*nb) pass, so the dominant kernel reads and writes the
//    trailing matrix once per G panels instead of once per panel.
//  * Look-ahead.  The trailing pass is split: the caller's stream updates the
//    next group's block columns and goes straight on to factor them, while an
//    auxiliary stream applies the group to the rest of the trailing matrix.
//    The next group's update waits for the previous rest-update (they write
//    the same elements); everything joins the caller's stream at the end.
//  * Float storage (the low-precision factor): the panel solve multiplies by
//    the inverses of the diagonal block's 32 x 32 pieces (potrf_f32.hip)
//    instead of the substitution the FP64 path keeps.
//  * Failure.  Every launch carries the device status word: a potrf or panel
//    solve after a failed pivot does nothing, and a trailing pass applies
//    only the panels of blocks before the failing one -- the reference's
//    state when it throws, up to the rounding order of the updates.
template <typename T>
static hipError_t syrk_launch(int trailing, int m, int n, int k, const T* A, i64 lda, T* C, i64 ldc, const int* info,
                              int kcol0, int nbp, hipStream_t s, bool beside_chain = false) {
    if constexpr (sizeof(T) == 8)
        return launch_syrk_ln_f64(m, n, k, -1.0, (const double*)A, lda, (const double*)A, lda, 1.0, (double*)C, ldc,
                                  info, kcol0, nbp, s, beside_chain);
    else
        return launch_syrk_ln_f32(trailing == kTrailingBf16x3, m, n, k, -1.0f, (const float*)A, lda, (const float*)A,
                                  lda, 1.0f, (float*)C, ldc, info, kcol0, nbp, s);
}

// update depth (columns per trailing pass)
template <typename T>
static int update_depth(int trailing) {
    return sizeof(T) == 8 ? kUpdateDepthF64 : (trailing == kTrailingF32 ? kUpdateDepthF32 : kUpdateDepthBf16x3);
}

// bf16 planes of the float factor's panels (two alternating sets) and the
// inverses of the diagonal block's 32 x 32 pieces
static int plane_count(int nb, int trailing) {
    return nb % 32 == 0 ? (trailing == kTrailingBf16x3 ? 2 : trailing == kTrailingBf16x6 ? 3 : 0) : 0;
}
static size_t plane_region_bytes(int n, int nb, int trailing) {
    const int np = plane_count(nb, trailing);
    if (np == 0 || n <= 0) return 0;
    const int G = std::max(1, update_depth<float>(trailing) / nb);
    const i64 gw = std::min<i64>((i64)G * nb, n);
    const i64 ldk = (gw + 31) / 32 * 32;
    return (2 * (size_t)np * (size_t)n * (size_t)ldk * sizeof(unsigned short) + 255) & ~(size_t)255;
}
size_t lowprec_scratch_bytes(int n, int nb, int trailing) {
    if (nb < 1 || nb > 128) return 0;
    // the planes, then the diagonal pieces' inverses (4 x 32 x 32)
    return plane_region_bytes(n, nb, trailing) + 4 * 1024 * sizeof(float);
}

template <typename T>
int cholesky_lower(int n, T* A, i64 lda, int nb, int trailing, int* dinfo, hipStream_t s, hipEvent_t trailing_ready,
                   const std::function<int(int)>* after_group, void* scratch_in) {
    if (n < 0 || lda < std::max(1, n) || nb < 1 || nb > 128 || dinfo == nullptr) return kErrArg;
    ELEMAMD_HIP_TRY(launch_fill_int(dinfo, 0, 1, s));
    if (n == 0) return kOk;
    StreamPair* sp = stream_pair(s);
    if (sp == nullptr) return kErrHip;
    const int G = std::max(1, update_depth<T>(trailing) / nb);
    const i64 gw64 = (i64)G * nb;
    const int gw = (int)std::min<i64>(gw64, n);
    // split-BF16x3 / x6: the panel solve also writes its result as 2 / 3
    // bf16 planes (row-major, rows from the group's first column down), which
    // every update of the group reads: the in-group block-column updates, the
    // next group's block columns and the rest.  Two sets, alternating by
    // group: the rest-update of a group still reads its planes while the next
    // group's are written.
    // (plane chunks are 32 columns: block columns must start on chunk
    // boundaries; split-BF16x3 has an in-kernel-split form for other nb)
    const int np = sizeof(T) == 4 ? plane_count(nb, trailing) : 0;
    if (sizeof(T) == 4 && trailing == kTrailingBf16x6 && np == 0) return kErrArg;
    const bool planes = np > 0;
    const i64 ldk = ((i64)gw + 31) / 32 * 32;
    const size_t plane_elems = (size_t)n * ldk;
    // float storage: the planes, then the diagonal pieces' inverses (4 x 32 x 32)
    const size_t plane_bytes = planes ? plane_region_bytes(n, nb, trailing) : 0;
    unsigned short* pbuf = nullptr;
    float* Xd = nullptr;
    if (sizeof(T) == 4) {
        char* scratch = scratch_in != nullptr ? (char*)scratch_in
                                              : (char*)stream_scratch(lowprec_scratch_bytes(n, nb, trailing), s);
        if (scratch == nullptr) return kErrHip;
        pbuf = (unsigned short*)scratch;
        Xd = (float*)(scratch + plane_bytes);
    }
    int group = 0;
    bool rest_pending = false;
    for (int c0 = 0; c0 < n; c0 += gw, ++group) {
        const int c1 = (int)std::min<i64>(n, (i64)c0 + gw);
        unsigned short* Pl = nullptr;  // this group's planes (stride plane_elems), row 0 = matrix row c0
        if (planes) Pl = pbuf + (size_t)(group & 1) * np * plane_elems;
        // the group's block columns, in order
        for (int k = c0; k < c1; k += nb) {
            const int b = std::min(nb, n - k);
            T* Akk = A + k + (i64)k * lda;
            // the block column's pending in-group update (panels c0 .. k-1):
            // with planes, the diagonal block here and the rows below it
            // inside the panel kernel (each slab before its solve)
            const unsigned short* upd = nullptr;
            if (k > c0) {  // panels c0 .. k-1 of this group into block column k
                if (planes) {
                    const unsigned short* Pk = Pl + (i64)(k - c0) * ldk;
                    upd = Pk;
                    ELEMAMD_HIP_TRY(launch_syrk_diag_bf16_planes(np, b, k - c0, Pk, (i64)plane_elems, ldk,
                                                                 (float*)Akk, lda, dinfo, c0, nb, s));
                } else {
                    ELEMAMD_HIP_TRY(syrk_launch<T>(trailing, n - k, b, k - c0, A + k + (i64)c0 * lda, lda, Akk, lda,
                                                   dinfo, c0, nb, s));
                }
            }
            if constexpr (sizeof(T) == 8) {
                ELEMAMD_HIP_TRY(launch_potrf_unb<T>(b, Akk, lda, dinfo, k, s));
                if (n - k - b > 0)
                    ELEMAMD_HIP_TRY(launch_panel_f64(n - k - b, b, (const double*)Akk, lda, (double*)Akk + b, lda,
                                                     dinfo, k + b, s));
            } else {  // the low-precision factor: diagonal inverse, panel as a product with it
                ELEMAMD_HIP_TRY(launch_potrf_f32(b, (float*)Akk, lda, Xd, dinfo, k, s));
                if (n - k - b > 0)
                    ELEMAMD_HIP_TRY(launch_panel_f32(
                        n - k - b, b, (const float*)Akk, lda, Xd, (float*)Akk + b, lda, dinfo, k + b,
                        planes ? Pl + (i64)(k + b - c0) * ldk + (k - c0) : nullptr, np, (i64)plane_elems, ldk, s,
                        upd, k - c0));
            }
        }
        if (after_group != nullptr) {
            const int rc = (*after_group)(c1);
            if (rc) return rc;
        }
        const int m2 = n - c1;
        if (m2 <= 0) break;
        const int kg = c1 - c0;
        const T* P = A + c1 + (i64)c0 * lda;  // the group's panels below it, m2 x kg
        T* C = A + c1 + (i64)c1 * lda;        // trailing matrix, m2 x m2
        const int bn = std::min(gw, m2);      // next group's block columns
        if (rest_pending) ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, sp->ev_b, 0));
        if (group == 0 && trailing_ready != nullptr) ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, trailing_ready, 0));
        if (planes) Pl += (i64)(c1 - c0) * ldk;  // rows from c1 on
        const bool rest = m2 > bn;
        if (rest) {  // fork: the rest of the trailing matrix on aux
            ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_a, s));
            ELEMAMD_HIP_TRY(hipStreamWaitEvent(sp->aux, sp->ev_a, 0));
            if (planes)
                ELEMAMD_HIP_TRY(launch_syrk_ln_bf16_planes(np, m2 - bn, m2 - bn, kg, -1.0f, Pl + (i64)bn * ldk,
                                                           (i64)plane_elems, ldk, 1.0f,
                                                           (float*)(C + bn + (i64)bn * lda), lda, dinfo, c0, nb,
                                                           sp->aux));
            else
                ELEMAMD_HIP_TRY(syrk_launch<T>(trailing, m2 - bn, m2 - bn, kg, P + bn, lda, C + bn + (i64)bn * lda,
                                               lda, dinfo, c0, nb, sp->aux, true));
            ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_b, sp->aux));
        }
        rest_pending = rest;
        // critical chain on the caller's stream: the next group's block columns
        if (planes)
            ELEMAMD_HIP_TRY(launch_syrk_ln_bf16_planes(np, m2, bn, kg, -1.0f, Pl, (i64)plane_elems, ldk, 1.0f,
                                                       (float*)C, lda, dinfo, c0, nb, s, 1));
        else
            ELEMAMD_HIP_TRY(syrk_launch<T>(trailing, m2, bn, kg, P, lda, C, lda, dinfo, c0, nb, s));
    }
    if (rest_pending) ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, sp->ev_b, 0));
    return kOk;
}

template int cholesky_lower<double>(int, double*, i64, int, int, int*, hipStream_t, hipEvent_t,
                                    const std::function<int(int)>*, void*);
template int cholesky_lower<float>(int, float*, i64, int, int, int*, hipStream_t, hipEvent_t,
                                   const std::function<int(int)>*, void*);

// ----------------------------------------------------------- SolveAfter
// cholesky::SolveAfter(LOWER, NORMAL) (SolveAfter.hpp:20-75): forward then
// backward triangular solve with the lower factor.  One column is the Trsv
// pair in the reference; the persistent kernels handle k = 1 the same way.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T>
size_t solve_work_bytes(int n, int k) {
    return align256(solve_dinv_count(n) * sizeof(T)) + align256(solve_ctrl_count(n, k) * sizeof(int));
}

template <typename T>
int solve_prepare(int n, const T* L, i64 ldl, T* Dinv, hipStream_t s) {
    ELEMAMD_HIP_TRY(launch_trtri_diag64<T>(n, L, ldl, Dinv, s));
    return kOk;
}

template <typename T>
int solve_apply(bool forward, bool backward, int n, int k, const T* L, i64 ldl, T* B, i64 ldb, const T* Dinv,
                int* ctrl, const int* dinfo, hipStream_t s, bool upper, bool unit) {
    // ctrl: solve_ctrl_count(n, k) ints = flags, then the helpers' partial sums;
    // ctrl[1] is the error word of the whole solve (both directions)
    T* part = (T*)((char*)ctrl + ((chain_ctrl_count(n, k) * sizeof(int) + 255) & ~(size_t)255));
    if (n > 0 && k > 0) ELEMAMD_HIP_TRY(launch_fill_int(ctrl + 1, 0, 1, s));
    if (forward)
        ELEMAMD_HIP_TRY(launch_trsm_chain<T>(false, n, k, L, ldl, B, ldb, Dinv, ctrl, part, dinfo, s, upper, unit));
    if (backward)
        ELEMAMD_HIP_TRY(launch_trsm_chain<T>(true, n, k, L, ldl, B, ldb, Dinv, ctrl, part, dinfo, s, upper, unit));
    return kOk;
}

int solve_read_error(const int* ctrl, hipStream_t s) {
    int lost = 0;
    ELEMAMD_HIP_TRY(hipMemcpyAsync(&lost, ctrl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    return lost != 0 ? kErrHip : kOk;
}

template <typename T>
int* solve_ctrl(int n, void* work) {
    return (int*)((char*)work + align256(solve_dinv_count(n) * sizeof(T)));
}

template <typename T>
int solve_after_lower(int n, int k, const T* L, i64 ldl, T* B, i64 ldb, const int* dinfo, void* work, hipStream_t s,
                      bool upper, bool inverse_blocks) {
    if (n < 0 || k < 0 || ldl < std::max(1, n) || ldb < std::max(1, n)) return kErrArg;
    if (n == 0 || k == 0) return kOk;
    if (work == nullptr) return kErrWorkspace;
    const bool inv = inverse_blocks && !upper;
    T* Dinv = inv ? (T*)work : nullptr;
    int* ctrl = solve_ctrl<T>(n, work);
    if (inv) {
        int rc = solve_prepare<T>(n, L, ldl, Dinv, s);
        if (rc) return rc;
    }
    return solve_apply<T>(true, true, n, k, L, ldl, B, ldb, Dinv, ctrl, dinfo, s, upper, false);
}

template size_t solve_work_bytes<double>(int, int);
template size_t solve_work_bytes<float>(int, int);
template int solve_prepare<double>(int, const double*, i64, double*, hipStream_t);
template int solve_prepare<float>(int, const float*, i64, float*, hipStream_t);
template int solve_apply<double>(bool, bool, int, int, const double*, i64, double*, i64, const double*, int*,
                                 const int*, hipStream_t, bool, bool);
template int solve_apply<float>(bool, bool, int, int, const float*, i64, float*, i64, const float*, int*, const int*,
                                hipStream_t, bool, bool);
template int* solve_ctrl<double>(int, void*);
template int* solve_ctrl<float>(int, void*);
template int solve_after_lower<double>(int, int, const double*, i64, double*, i64, const int*, void*, hipStream_t,
                                       bool, bool);
template int solve_after_lower<float>(int, int, const float*, i64, float*, i64, const int*, void*, hipStream_t, bool,
                                      bool);

// ------------------------------------------------------ mixed precision
namespace {

struct Carve {
    char* base;
    size_t off = 0, cap;
    Carve(void* p, size_t c) : base((char*)p), cap(c) {}
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
};

struct MixedWork {
    double* X;     // n x k iterate
    double* R;     // n x k residual
    float* Rf;     // n x k float correction
    double* D;     // n x k correction (GMRES)
    double* W;     // n x k GMRES work vector
    double* V;     // (m+1) blocks of n x k: Krylov basis
    double* nrm;   // 8k + n + 8 doubles: norms, row sums, scratch
    double* h;     // (m+1) x k Gram-Schmidt coefficients
    void* sweep;   // block vectors and flags of the correction sweeps
    double* part;  // partial sums of the split symmetric product
    double* P;     // n x k CG search directions
    double* Q;     // n x k A P
    double* Z;     // n x k preconditioned residual
    int* active;   // k CG column flags (0: converged, frozen)
    size_t bytes;
};

MixedWork carve_refine(void* work, size_t cap, int n, int k, int m, bool cg) {
    Carve c(work, cap);
    MixedWork w;
    const size_t nk = (size_t)n * k;
    w.X = c.take<double>(nk);
    w.R = c.take<double>(nk);
    w.Rf = c.take<float>(nk);
    w.D = c.take<double>(m > 0 ? nk : 0);
    w.W = c.take<double>(m > 0 ? nk : 0);
    w.V = c.take<double>(m > 0 ? nk * (size_t)(m + 1) : 0);
    w.nrm = c.take<double>((size_t)8 * k + n + 8);
    w.h = c.take<double>((size_t)(m + 2) * k);
    w.sweep = c.take<char>(sweep_work_bytes(n));
    w.part = c.take<double>(symm_part_count(n, k));
    w.P = c.take<double>(cg ? nk : 0);
    w.Q = c.take<double>(cg ? nk : 0);
    w.Z = c.take<double>(cg ? nk : 0);
    w.active = c.take<int>(cg ? (size_t)k : 0);
    w.bytes = c.off + 256;
    return w;
}

struct Ctx {
    int n, k;
    const double* A;
    i64 lda;
    const double* Bin;
    i64 ldb;
    MixedWork w;
    const MixedFactor* f;
    hipStream_t s;
    double anorm;
};

// z := M^{-1} y with M = S S^T in float, applied to each column scaled by
// colscale (device, k entries) to stay inside the float range.
int precond_apply(Ctx& c, const double* y, double* z, const double* colscale) {
    const int n = c.n, k = c.k;
    ELEMAMD_HIP_TRY(launch_copy_f64_to_f32_scaled(n, k, y, n, colscale, c.w.Rf, n, c.s));
    ELEMAMD_HIP_TRY(launch_sweep_solve(n, k, c.f->S, n, c.f->tiles, c.w.Rf, n, c.w.sweep, c.s));
    ELEMAMD_HIP_TRY(launch_copy_f32_to_f64_scaled(n, k, c.w.Rf, n, colscale, z, n, c.s));
    return kOk;
}

// Column norms: out[c] = ||X(:,c)||_inf, copied to host.
int host_colmax(Ctx& c, const double* X, double* dev_out, std::vector<double>& host) {
    ELEMAMD_HIP_TRY(launch_col_absmax(c.n, c.k, X, c.n, dev_out, c.s));
    host.resize(c.k);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(host.data(), dev_out, sizeof(double) * c.k, hipMemcpyDeviceToHost, c.s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(c.s));
    return kOk;
}

// One restarted GMRES cycle for A d = r (per column, left preconditioned by
// M = S S^T), started from d = 0.  rscale[c] = ||r(:,c)||_inf on device.
int gmres_cycle(Ctx& c, int m, double inner_tol, int* iters_out) {
    const int n = c.n, k = c.k;
    const size_t nk = (size_t)n * k;
    double* dscale = c.w.nrm + 4 * k;  // device: per-column scale scratch
    double* dnorm = c.w.nrm + 5 * k;   // device: per-column 2-norms / dots scratch
    double* h = c.w.h;
    // z0 = M^{-1} r  -> V_0 after normalization
    std::vector<double> hs;
    int rc = host_colmax(c, c.w.R, dscale, hs);
    if (rc) return rc;
    rc = precond_apply(c, c.w.R, c.w.V, dscale);
    if (rc) return rc;
    ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, c.w.V, n, (i64)nk, c.w.V, n, dnorm, c.s));
    std::vector<double> beta(k);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(beta.data(), dnorm, sizeof(double) * k, hipMemcpyDeviceToHost, c.s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(c.s));
    for (int j = 0; j < k; ++j) beta[j] = std::sqrt(beta[j]);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(dnorm, beta.data(), sizeof(double) * k, hipMemcpyHostToDevice, c.s));
    ELEMAMD_HIP_TRY(launch_col_scale(n, k, c.w.V, n, dnorm, 1, c.w.V, n, c.s));

    // Hessenberg (column-major (m+1) x m per RHS), Givens rotations, rhs g.
    std::vector<double> H((size_t)k * (m + 1) * m, 0.0), cs((size_t)k * m), sn((size_t)k * m),
        g((size_t)k * (m + 1), 0.0);
    std::vector<char> done(k, 0);
    for (int col = 0; col < k; ++col) {
        g[(size_t)col * (m + 1)] = beta[col];
        if (!(beta[col] > 0.0)) done[col] = 1;
    }
    std::vector<double> hcol((size_t)(m + 1) * k), h2((size_t)(m + 1) * k), wn(k);
    int jdone = 0;
    for (int j = 0; j < m; ++j) {
        double* Vj = c.w.V + (size_t)j * nk;
        double* Vj1 = c.w.V + (size_t)(j + 1) * nk;
        // W = M^{-1} A V_j
        ELEMAMD_HIP_TRY(launch_symm_lower(n, k, c.A, c.lda, Vj, n, c.w.D, n, c.w.part, c.s));
        rc = host_colmax(c, c.w.D, dscale, hs);
        if (rc) return rc;
        rc = precond_apply(c, c.w.D, c.w.W, dscale);
        if (rc) return rc;
        // CGS2 against V_0..V_j
        const int nv = j + 1;
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, nv, c.w.V, n, (i64)nk, c.w.W, n, h, c.s));
        ELEMAMD_HIP_TRY(launch_batched_gemv(n, k, nv, -1.0, c.w.V, n, (i64)nk, h, 1.0, c.w.W, n, c.s));
        ELEMAMD_HIP_TRY(hipMemcpyAsync(hcol.data(), h, sizeof(double) * nv * k, hipMemcpyDeviceToHost, c.s));
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, nv, c.w.V, n, (i64)nk, c.w.W, n, h, c.s));
        ELEMAMD_HIP_TRY(launch_batched_gemv(n, k, nv, -1.0, c.w.V, n, (i64)nk, h, 1.0, c.w.W, n, c.s));
        ELEMAMD_HIP_TRY(hipMemcpyAsync(h2.data(), h, sizeof(double) * nv * k, hipMemcpyDeviceToHost, c.s));
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, c.w.W, n, (i64)nk, c.w.W, n, dnorm, c.s));
        ELEMAMD_HIP_TRY(hipMemcpyAsync(wn.data(), dnorm, sizeof(double) * k, hipMemcpyDeviceToHost, c.s));
        ELEMAMD_HIP_TRY(hipStreamSynchronize(c.s));
        for (int col = 0; col < k; ++col) wn[col] = std::sqrt(std::max(wn[col], 0.0));
        ELEMAMD_HIP_TRY(hipMemcpyAsync(dnorm, wn.data(), sizeof(double) * k, hipMemcpyHostToDevice, c.s));
        ELEMAMD_HIP_TRY(launch_col_scale(n, k, c.w.W, n, dnorm, 1, Vj1, n, c.s));
        bool all = true;
        for (int col = 0; col < k; ++col) {
            double* Hc = &H[(size_t)col * (m + 1) * m];
            double* hc = &Hc[(size_t)j * (m + 1)];
            for (int i = 0; i < nv; ++i) hc[i] = hcol[i + (size_t)col * nv] + h2[i + (size_t)col * nv];
            hc[j + 1] = wn[col];
            for (int i = 0; i < j; ++i) {  // previous rotations
                const double a = hc[i], b = hc[i + 1];
                const double cc = cs[(size_t)col * m + i], ss = sn[(size_t)col * m + i];
                hc[i] = cc * a + ss * b;
                hc[i + 1] = -ss * a + cc * b;
            }
            const double a = hc[j], b = hc[j + 1];
            const double r = std::hypot(a, b);
            const double cc = (r > 0.0) ? a / r : 1.0, ss = (r > 0.0) ? b / r : 0.0;
            cs[(size_t)col * m + j] = cc;
            sn[(size_t)col * m + j] = ss;
            hc[j] = r;
            hc[j + 1] = 0.0;
            double* gc = &g[(size_t)col * (m + 1)];
            gc[j + 1] = -ss * gc[j];
            gc[j] = cc * gc[j];
            if (!done[col] && std::fabs(gc[j + 1]) <= inner_tol * beta[col]) done[col] = 1;
            if (!done[col]) all = false;
        }
        jdone = j + 1;
        if (all) break;
    }
    *iters_out += jdone;
    // y = H^{-1} g (upper triangular, per column), then d = V y
    std::vector<double> y((size_t)jdone * k, 0.0);
    for (int col = 0; col < k; ++col) {
        const double* Hc = &H[(size_t)col * (m + 1) * m];
        const double* gc = &g[(size_t)col * (m + 1)];
        double* yc = &y[(size_t)col * jdone];
        for (int i = jdone - 1; i >= 0; --i) {
            double t = gc[i];
            for (int l = i + 1; l < jdone; ++l) t -= Hc[i + (size_t)l * (m + 1)] * yc[l];
            const double d = Hc[i + (size_t)i * (m + 1)];
            yc[i] = (d != 0.0) ? t / d : 0.0;
        }
    }
    if (jdone > 0) {
        ELEMAMD_HIP_TRY(hipMemcpyAsync(h, y.data(), sizeof(double) * jdone * k, hipMemcpyHostToDevice, c.s));
        ELEMAMD_HIP_TRY(launch_batched_gemv(n, k, jdone, 1.0, c.w.V, n, (i64)nk, h, 0.0, c.w.D, n, c.s));
    } else {
        ELEMAMD_HIP_TRY(launch_scale_f64(n, k, 0.0, c.w.D, n, c.s));
    }
    ELEMAMD_HIP_TRY(hipStreamSynchronize(c.s));
    return kOk;
}

// ELEMAMD_CG_TRACE=1 prints the CG refinement's backward errors to stderr.
bool cg_trace() {
    static const bool on = [] {
        const char* v = std::getenv("ELEMAMD_CG_TRACE");
        return v != nullptr && v[0] == '1';
    }();
    return on;
}

// Normwise backward error per column from the device column maxima of R
// (dev[0..k)) and X (dev[k..2k)), copied to host; returns the maximum.
int host_eta(Ctx& c, const std::vector<double>& bn, std::vector<double>& eta_c, double& eta) {
    const int n = c.n, k = c.k;
    std::vector<double> nx(2 * (size_t)k);
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, c.w.R, n, c.w.nrm + k, c.s));
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, c.w.X, n, c.w.nrm + 2 * k, c.s));
    ELEMAMD_HIP_TRY(hipMemcpyAsync(nx.data(), c.w.nrm + k, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, c.s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(c.s));
    eta_c.resize(k);
    eta = 0.0;
    for (int col = 0; col < k; ++col) {
        eta_c[col] = column_backward_error(nx[col], c.anorm, nx[k + col], bn[col]);
        eta = std::max(eta, eta_c[col]);
    }
    return kOk;
}

// Preconditioned conjugate gradients on A X = B, one independent recurrence
// per column, from X = 0, R = B: the preconditioner is the low-precision
// factor (the correction sweeps, in float), everything else FP64.  The
// preconditioner's rounding differs from call to call, so the direction
// update uses the flexible coefficient beta = -z_new.q / p.q (Q = A P),
// which equals the Fletcher-Reeves one for a fixed preconditioner.  When the
// recursive residual meets the target the true residual B - A X is formed;
// columns still short of it restart from it.  Columns that reach the target
// are frozen.  Returns with *conv = 1 when every column's true backward
// error is <= tol, otherwise (stagnation, step limit, breakdown) with X
// holding the best iterate so far for classic refinement to continue from.
int cg_refine(Ctx& c, const MixedOpts& o, const std::vector<double>& bn, MixedReport& r, double& eta, int* conv) {
    const int n = c.n, k = c.k;
    const size_t nk = (size_t)n * k;
    MixedWork& w = c.w;
    hipStream_t s = c.s;
    double* rz = w.nrm + 3 * k;
    double* pq = w.nrm + 4 * k;
    double* zq = w.nrm + 5 * k;
    *conv = 0;
    std::vector<int> act(k, 1), act_dev(k, -1);
    std::vector<double> eta_c;
    StreamPair* sp = stream_pair(s);
    if (sp == nullptr) return kErrHip;
    // column maxima of R and X land in pinned memory behind an event, so the
    // next step's preconditioner can be enqueued before the host reads them
    // (one grow-only buffer per host thread, kept until the thread ends)
    struct Pinned {
        double* p = nullptr;
        size_t cap = 0;
        ~Pinned() {
            if (p) (void)hipHostFree(p);
        }
    };
    static thread_local Pinned nx;
    if (nx.cap < 2 * (size_t)k) {
        if (nx.p) ELEMAMD_HIP_TRY(hipHostFree(nx.p));
        nx.p = nullptr;
        nx.cap = 0;
        ELEMAMD_HIP_TRY(hipHostMalloc((void**)&nx.p, sizeof(double) * 2 * k, hipHostMallocDefault));
        nx.cap = 2 * (size_t)k;
    }
    auto push_active = [&]() -> hipError_t {
        if (act == act_dev) return hipSuccess;
        act_dev = act;
        return hipMemcpyAsync(w.active, act_dev.data(), sizeof(int) * k, hipMemcpyHostToDevice, s);
    };
    auto eta_enqueue = [&]() -> hipError_t {
        hipError_t e = launch_col_absmax(n, k, w.R, n, w.nrm + k, s);
        if (e == hipSuccess) e = launch_col_absmax(n, k, w.X, n, w.nrm + 2 * k, s);
        if (e == hipSuccess) e = hipMemcpyAsync(nx.p, w.nrm + k, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(sp->ev_d, s);
        return e;
    };
    auto eta_finish = [&]() -> hipError_t {
        hipError_t e = hipEventSynchronize(sp->ev_d);
        if (e != hipSuccess) return e;
        eta_c.resize(k);
        eta = 0.0;
        for (int col = 0; col < k; ++col) {
            eta_c[col] = column_backward_error(nx.p[col], c.anorm, nx.p[k + col], bn[col]);
            eta = std::max(eta, eta_c[col]);
        }
        return hipSuccess;
    };
    // Z = M^{-1} R, then P = Z - (z.q / p.q) P (flexible beta); touches
    // neither X nor R, so it may be enqueued before this step's convergence
    // is known (it is wasted only on the last step)
    auto direction = [&]() -> int {
        int rc = precond_apply(c, w.R, w.Z, w.nrm + k);
        if (rc) return rc;
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, w.Z, n, (i64)nk, w.Q, n, zq, s));
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, w.R, n, (i64)nk, w.Z, n, rz, s));
        ELEMAMD_HIP_TRY(launch_cg_p(n, k, zq, pq, w.Z, w.P, s));
        return kOk;
    };
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, w.R, n, w.nrm + k, s));  // preconditioner scale of R = B
    bool restart = true;
    double best = INFINITY, prev = INFINITY;
    int since_best = 0;
    for (;;) {
        ELEMAMD_HIP_TRY(push_active());
        if (restart) {  // P = Z = M^{-1} R
            int rc = precond_apply(c, w.R, w.Z, w.nrm + k);
            if (rc) return rc;
            ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, w.R, n, (i64)nk, w.Z, n, rz, s));
            ELEMAMD_HIP_TRY(launch_copy_f64(n, k, w.Z, n, w.P, n, s));
            restart = false;
        }
        ELEMAMD_HIP_TRY(launch_symm_lower(n, k, c.A, c.lda, w.P, n, w.Q, n, w.part, s));
        ELEMAMD_HIP_TRY(launch_batched_dots(n, k, 1, w.P, n, (i64)nk, w.Q, n, pq, s));
        ELEMAMD_HIP_TRY(launch_cg_xr(n, k, rz, pq, w.active, w.P, w.Q, w.X, w.R, s));
        r.iterations += 1;
        ELEMAMD_HIP_TRY(eta_enqueue());
        // the backward error drops about tenfold per step: unless the last
        // one was within 20x of the target, this step will not converge, so
        // the next direction is enqueued now
        const bool ahead = prev > 20.0 * o.tol && r.iterations < o.max_iter;
        if (ahead) {
            int rc = direction();
            if (rc) return rc;
        }
        ELEMAMD_HIP_TRY(eta_finish());
        prev = eta;
        if (cg_trace()) std::fprintf(stderr, "cg step %d: recursive eta %.3e\n", r.iterations, eta);
        if (!(eta < INFINITY)) return kOk;  // breakdown: classic refinement takes over
        if (eta <= o.tol) {  // confirm with the true residual
            ELEMAMD_HIP_TRY(launch_symm_lower_residual(n, k, c.A, c.lda, w.X, n, c.Bin, c.ldb, w.R, n, w.part, s));
            int rc = host_eta(c, bn, eta_c, eta);
            if (rc) return rc;
            if (cg_trace()) {
                int nbad = 0;
                for (int col = 0; col < k; ++col) nbad += eta_c[col] > o.tol;
                std::fprintf(stderr, "cg true residual: eta %.3e, %d columns above the target\n", eta, nbad);
            }
            if (eta <= o.tol) {
                *conv = 1;
                return kOk;
            }
            for (int col = 0; col < k; ++col) act[col] = eta_c[col] > o.tol;
            restart = true;
            if (r.iterations >= o.max_iter) return kOk;
            continue;
        }
        if (r.iterations >= o.max_iter) return kOk;
        if (eta < 0.5 * best) {
            best = eta;
            since_best = 0;
        } else if (++since_best >= 3) {
            return kOk;  // stagnation
        }
        for (int col = 0; col < k; ++col)
            if (eta_c[col] <= o.tol) act[col] = 0;
        if (eta <= 8.0 * o.tol) {
            // within 8x of the target: one classic step x += M^{-1} r (the
            // preconditioned residual is at hand, no product with A; a
            // classic step contracts ~10x here) closes the gap; the true
            // residual confirms it, or CG restarts from it
            if (!ahead) {
                int rc = precond_apply(c, w.R, w.Z, w.nrm + k);
                if (rc) return rc;
            }
            ELEMAMD_HIP_TRY(launch_axpy_f64(n, k, 1.0, w.Z, n, w.X, n, s));
            r.iterations += 1;
            ELEMAMD_HIP_TRY(launch_symm_lower_residual(n, k, c.A, c.lda, w.X, n, c.Bin, c.ldb, w.R, n, w.part, s));
            int rc = host_eta(c, bn, eta_c, eta);
            if (rc) return rc;
            if (cg_trace()) std::fprintf(stderr, "cg closing step %d: true eta %.3e\n", r.iterations, eta);
            if (eta <= o.tol) {
                *conv = 1;
                return kOk;
            }
            if (!(eta < INFINITY)) return kOk;
            for (int col = 0; col < k; ++col) act[col] = eta_c[col] > o.tol;
            restart = true;
            if (r.iterations >= o.max_iter) return kOk;
            continue;
        }
        if (!ahead) {
            int rc = direction();
            if (rc) return rc;
        }
    }
}

}  // namespace

MixedFactor carve_factor(void* mem, int n) {
    Carve c(mem, 0);
    MixedFactor f{};
    f.S = c.take<float>((size_t)n * n);
    f.Dinv = c.take<float>(solve_dinv_count(n));
    f.Xp = c.take<float>(sweep_xinv_count(n));
    f.tiles = c.take<float>(sweep_tiles_count(n));
    f.dinfo = c.take<int>(4);
    f.scratch = c.take<double>((size_t)n + 8);  // row sums of |A|, their maximum
    return f;
}

size_t mixed_factor_bytes(int n) {
    Carve c(nullptr, 0);
    c.take<float>((size_t)n * n);
    c.take<float>(solve_dinv_count(n));
    c.take<float>(sweep_xinv_count(n));
    c.take<float>(sweep_tiles_count(n));
    c.take<int>(4);
    c.take<double>((size_t)n + 8);
    return c.off + 256;
}

size_t mixed_refine_bytes(int n, int k, const MixedOpts& o) {
    return carve_refine(nullptr, 0, n, k, o.gmres_restart, o.refine == kRefineCG).bytes;
}

// The plain FP64 finish of a failed low-precision solve: an n x n copy of A
// to factor, the original B, the residual, the solve's control words and the
// residual's partial sums.
struct FallbackWork {
    double* Af;    // n x n, ld n
    double* B0;    // n x k, ld n: the right-hand sides as given
    double* R;     // n x k residual
    double* nrm;   // 3k column maxima
    double* part;  // symm_part_count(n, k)
    void* solve;   // solve_work_bytes<double>(n, k)
    size_t bytes;
};
static FallbackWork carve_fallback(void* mem, int n, int k) {
    Carve c(mem, 0);
    FallbackWork w;
    const size_t nk = (size_t)n * k;
    w.Af = c.take<double>((size_t)n * n);
    w.B0 = c.take<double>(nk);
    w.R = c.take<double>(nk);
    w.nrm = c.take<double>((size_t)3 * k);
    w.part = c.take<double>(symm_part_count(n, k));
    w.solve = c.take<char>(solve_work_bytes<double>(n, k));
    w.bytes = c.off + 256;
    return w;
}

size_t mixed_workspace_bytes(int n, int k, const MixedOpts& o) {
    return align256(mixed_factor_bytes(n)) + align256(lowprec_scratch_bytes(n, o.nb, o.trailing)) +
           align256(mixed_refine_bytes(n, k, o)) + (o.fallback ? carve_fallback(nullptr, n, k).bytes : 0);
}

// Low-precision factor of A: S = float(A) (lower), FP64 diagonal potrf,
// float panels and trailing updates; then the operands of the block-inverse
// correction sweeps (diagonal-block inverses, look-ahead products, the
// mirror of L in the upper triangle of S); ||A||_inf.  On breakdown the factor of A + sigma*diag(A) is retried.
int mixed_factor(int n, const double* A, i64 lda, const MixedOpts& o, MixedFactor& f, hipStream_t s, void* scratch) {
    if (n < 0 || lda < std::max(1, n) || o.nb < 1 || o.nb > 128) return kErrArg;
    f.lowprec_info = 0;
    f.shifted = 0;
    f.anorm = 0.0;
    if (n == 0) return kOk;
    StreamPair* sp = stream_pair(s);
    if (sp == nullptr) return kErrHip;
    // ||A||_inf (only the refinement's backward errors need it) on aux3
    // beside the factorization's first steps, into pinned memory behind
    // ev_c; aux3 starts after the caller's work before this call.  Unsplit
    // (one workgroup per 64-row block, about one per CU): it is needed only
    // at the end of the factorization, and this way every CU keeps room for
    // the chain's first diagonal factor and panel, which otherwise waited
    // for the split form's three workgroups per CU to leave
    double* rowsum = f.scratch;
    double* dtmp = f.scratch + n;
    ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_c, s));
    ELEMAMD_HIP_TRY(hipStreamWaitEvent(sp->aux3, sp->ev_c, 0));
    ELEMAMD_HIP_TRY(launch_sym_lower_abs_rowsum(n, A, lda, rowsum, nullptr, sp->aux3));
    ELEMAMD_HIP_TRY(launch_vec_max(n, rowsum, dtmp, sp->aux3));
    ELEMAMD_HIP_TRY(hipMemcpyAsync(sp->pinned, dtmp, sizeof(double), hipMemcpyDeviceToHost, sp->aux3));
    ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_c, sp->aux3));
    double sigma = 0.0;
    int info = 0;
    // the float copy: the first group of block columns on the caller's
    // stream, the columns right of it on aux under that group's factor steps
    // (the factorization waits for them before its first trailing update)
    const int gw = std::min(n, std::max(1, update_depth<float>(o.trailing) / std::max(1, o.nb)) * o.nb);
    // the solve preparation (diagonal-block inverses, sweep operands) of each
    // 512-column block on aux3 once the factorization has made that block's
    // columns final; it reads only the lower triangle
    const int SB = sweep_block_cols();
    const int nT = (n + SB - 1) / SB;
    int prepared = 0;
    const std::function<int(int)> prep = [&](int c1) -> int {
        const int ready_blocks = c1 >= n ? nT : c1 / SB;
        if (ready_blocks <= prepared) return kOk;
        // not before 3/4 of the columns: earlier, the bulk update keeps the
        // CUs busy and the preparation's narrow, latency-bound launches beside
        // it slow the factor chain by more than they take off the end
        // (measured at n = 16384: from the start 28.0 ms per solve, from 3/4
        // 27.5, all after the factorization 27.9)
        if (c1 < n && 4 * (i64)c1 < 3 * (i64)n) return kOk;
        ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_g, s));
        ELEMAMD_HIP_TRY(hipStreamWaitEvent(sp->aux3, sp->ev_g, 0));
        ELEMAMD_HIP_TRY(launch_trtri_diag64_range<float>(n, prepared * (SB / 64), ready_blocks * (SB / 64), f.S, n,
                                                         f.Dinv, sp->aux3));
        ELEMAMD_HIP_TRY(
            launch_sweep_prepare_blocks(n, prepared, ready_blocks, f.S, n, f.Dinv, f.Xp, f.tiles, sp->aux3));
        prepared = ready_blocks;
        return kOk;
    };
    // the caller's stream waits for everything enqueued on aux3
    auto join_prep = [&]() -> int {
        ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_h, sp->aux3));
        ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, sp->ev_h, 0));
        return kOk;
    };
    for (int attempt = 0; attempt < 4; ++attempt) {
        hipEvent_t ready = nullptr;
        if (attempt > 0) {
            // the failed attempt's mirror (aux) and preparation (aux3) still
            // read S: the next attempt's copy into S waits for both
            ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, sp->ev_b, 0));
            if (int rc = join_prep()) return rc;
        }
        prepared = 0;
        if (sigma == 0.0 && gw < n && o.nb > 0) {
            ELEMAMD_HIP_TRY(launch_copy_lower_f64_to_f32_cols(n, 0, gw, A, lda, f.S, n, s));
            ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_a, s));
            ELEMAMD_HIP_TRY(hipStreamWaitEvent(sp->aux, sp->ev_a, 0));
            ELEMAMD_HIP_TRY(launch_copy_lower_f64_to_f32_cols(n, gw, n, A, lda, f.S, n, sp->aux));
            ELEMAMD_HIP_TRY(hipEventRecord(sp->ev_f, sp->aux));
            ready = sp->ev_f;
        } else {
            ELEMAMD_HIP_TRY(launch_copy_lower_f64_to_f32(n, A, lda, f.S, n, s));
            if (sigma > 0.0) ELEMAMD_HIP_TRY(launch_shift_diag(n, A, lda, f.S, n, sigma, s));
        }
        int rc = cholesky_lower<float>(n, f.S, n, o.nb, o.trailing, f.dinfo, s, ready, &prep, scratch);
        if (rc == kOk && prepared < nT) rc = prep(n);
        if (rc == kOk) {
            // the mirror (upper triangle) on aux before the host reads the
            // breakdown flag; it reads only the lower triangle, and after a
            // breakdown the next attempt's mirror follows it on aux
            rc = hipEventRecord(sp->ev_a, s) == hipSuccess && hipStreamWaitEvent(sp->aux, sp->ev_a, 0) == hipSuccess &&
                         launch_sweep_mirror(n, f.S, n, sp->aux) == hipSuccess &&
                         hipEventRecord(sp->ev_b, sp->aux) == hipSuccess
                     ? kOk
                     : kErrHip;
        }
        if (rc) {  // the work on aux and aux3 ends before the caller's stream moves on
            if (ready != nullptr) (void)hipStreamWaitEvent(s, ready, 0);
            (void)join_prep();
            return rc;
        }
        ELEMAMD_HIP_TRY(hipMemcpyAsync(&info, f.dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
        ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
        f.lowprec_info = info;
        if (info == 0) break;
        if (!o.allow_shift) break;
        sigma = (sigma == 0.0) ? 1e-5 : sigma * 10.0;
        f.shifted = 1;
    }
    // the caller's stream waits for the mirror (aux) and the preparation (aux3)
    ELEMAMD_HIP_TRY(hipStreamWaitEvent(s, sp->ev_b, 0));
    if (int rc = join_prep()) return rc;
    ELEMAMD_HIP_TRY(hipEventSynchronize(sp->ev_c));  // the norm, long done by now
    f.anorm = sp->pinned[0];
    if (info != 0) return kErrNotConverged;  // broke down even after shifting
    return kOk;
}

// The dispatch rule's plain FP64 route: Cholesky + SolveAfter on a copy of
// A (A itself stays read-only), from the right-hand sides as given, and the
// backward error of the result.  kErrNotConverged if A is not numerically
// HPD in FP64 either (r.info = the failing column) or the result misses the
// target (NaN or Inf in the input).
static int fallback_fp64(int n, int k, const double* A, i64 lda, double* B, i64 ldb, const MixedOpts& o,
                         const MixedFactor& f, const FallbackWork& fw, MixedReport& r, hipStream_t s) {
    r.fallback = 1;
    r.converged = 0;
    ELEMAMD_HIP_TRY(launch_copy_f64(n, n, A, lda, fw.Af, n, s));
    int rc = cholesky_lower<double>(n, fw.Af, n, o.nb, kTrailingF64, f.dinfo, s);
    if (rc) return rc;
    int info = 0;
    ELEMAMD_HIP_TRY(hipMemcpyAsync(&info, f.dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    r.info = info;
    if (info != 0) return kErrNotConverged;
    ELEMAMD_HIP_TRY(launch_copy_f64(n, k, fw.B0, n, B, ldb, s));
    rc = solve_after_lower<double>(n, k, fw.Af, n, B, ldb, nullptr, fw.solve, s, false, true);
    if (rc) return rc;
    rc = solve_read_error(solve_ctrl<double>(n, fw.solve), s);
    if (rc) return rc;
    // eta = max_c ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)
    ELEMAMD_HIP_TRY(launch_symm_lower_residual(n, k, A, lda, B, ldb, fw.B0, n, fw.R, n, fw.part, s));
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, fw.R, n, fw.nrm, s));
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, B, ldb, fw.nrm + k, s));
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, fw.B0, n, fw.nrm + 2 * k, s));
    std::vector<double> h((size_t)3 * k);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(h.data(), fw.nrm, sizeof(double) * 3 * k, hipMemcpyDeviceToHost, s));
    ELEMAMD_HIP_TRY(hipStreamSynchronize(s));
    double eta = 0.0;
    for (int c = 0; c < k; ++c) {
        eta = std::max(eta, column_backward_error(h[c], f.anorm, h[k + c], h[2 * k + c]));
    }
    r.backward_error = eta;
    r.converged = eta <= o.tol ? 1 : 0;
    return r.converged ? kOk : kErrNotConverged;
}

int hpd_solve_mixed(int n, int k, const double* A, i64 lda, double* B, i64 ldb, const MixedOpts& o, void* work,
                    size_t work_bytes, MixedReport* rep, hipStream_t s) {
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) return kErrArg;
    if (o.nb < 1 || o.nb > 128) return kErrArg;
    if (n == 0 || k == 0) {
        if (rep) {
            *rep = MixedReport{};
            rep->converged = 1;
        }
        return kOk;
    }
    // the caller's workspace: the factor, the factorization's scratch, the
    // refinement's vectors, the FP64 fallback's buffers
    const size_t fb = align256(mixed_factor_bytes(n));
    const size_t sb = align256(lowprec_scratch_bytes(n, o.nb, o.trailing));
    const size_t rb = align256(mixed_refine_bytes(n, k, o));
    if (work == nullptr || work_bytes < mixed_workspace_bytes(n, k, o)) return kErrWorkspace;
    MixedFactor f = carve_factor(work, n);
    FallbackWork fw{};
    if (o.fallback) {
        fw = carve_fallback((char*)work + fb + sb + rb, n, k);
        ELEMAMD_HIP_TRY(launch_copy_f64(n, k, B, ldb, fw.B0, n, s));
    }
    MixedReport r{};
    int rc = mixed_factor(n, A, lda, o, f, s, (char*)work + fb);
    if (rc == kErrNotConverged) {
        r.lowprec_info = f.lowprec_info;
        r.shifted = f.shifted;
        r.anorm_inf = f.anorm;
    } else if (rc) {
        return rc;
    } else {
        rc = mixed_refine(n, k, A, lda, B, ldb, o, f, (char*)work + fb + sb, rb, &r, s);
        if (rc != kOk && rc != kErrNotConverged) return rc;
    }
    if (rc == kErrNotConverged && o.fallback) rc = fallback_fp64(n, k, A, lda, B, ldb, o, f, fw, r, s);
    if (rep) *rep = r;
    return rc;
}

// FP64-residual refinement (classic, then GMRES-IR on stagnation) of A X = B
// with a low-precision factor; B is overwritten with X.  Synchronous.
int mixed_refine(int n, int k, const double* A, i64 lda, double* B, i64 ldb, const MixedOpts& o,
                 const MixedFactor& f, void* work, size_t work_bytes, MixedReport* rep, hipStream_t s) {
    MixedReport r{};
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) return kErrArg;
    r.lowprec_info = f.lowprec_info;
    r.shifted = f.shifted;
    r.anorm_inf = f.anorm;
    if (n == 0 || k == 0) {
        r.converged = 1;
        if (rep) *rep = r;
        return kOk;
    }
    const int m = o.gmres_restart;
    MixedWork w = carve_refine(work, work_bytes, n, k, m, o.refine == kRefineCG);
    if (work == nullptr || work_bytes < w.bytes) return kErrWorkspace;
    Ctx c{n, k, A, lda, B, ldb, w, &f, s, f.anorm};
    std::vector<double> bn, xn, rn;
    ELEMAMD_HIP_TRY(launch_col_absmax(n, k, B, ldb, w.nrm, s));
    bn.resize(k);
    ELEMAMD_HIP_TRY(hipMemcpyAsync(bn.data(), w.nrm, sizeof(double) * k, hipMemcpyDeviceToHost, s));

    ELEMAMD_HIP_TRY(launch_sweep_clear_error(n, w.sweep, s));
    // X = 0, R = B
    ELEMAMD_HIP_TRY(launch_copy_f64(n, k, B, ldb, w.R, n, s));
    ELEMAMD_HIP_TRY(launch_scale_f64(n, k, 0.0, w.X, n, s));
    bool gmres = o.force_gmres && m > 0;
    double prev = INFINITY;
    int stall = 0;
    double eta = INFINITY;
    double gmres_prev = INFINITY;  // backward error before the last GMRES-IR cycle
    int it0 = 0, cg_conv = 0;
    if (o.refine == kRefineCG && !gmres) {
        int rc = cg_refine(c, o, bn, r, eta, &cg_conv);
        if (rc) return rc;
        if (eta < INFINITY) {
            it0 = 1;  // classic refinement continues from the CG iterate
        } else {      // breakdown: classic refinement from X = 0
            ELEMAMD_HIP_TRY(launch_copy_f64(n, k, B, ldb, w.R, n, s));
            ELEMAMD_HIP_TRY(launch_scale_f64(n, k, 0.0, w.X, n, s));
        }
    }
    const int base = r.iterations;
    for (int it = it0; !cg_conv && it <= o.max_iter; ++it) {
        if (it > 0) {
            ELEMAMD_HIP_TRY(launch_symm_lower_residual(n, k, A, lda, w.X, n, B, ldb, w.R, n, w.part, s));
            r.iterations = base + it;
        }
        // norms
        int rc = host_colmax(c, w.R, w.nrm + k, rn);
        if (rc) return rc;
        rc = host_colmax(c, w.X, w.nrm + 2 * k, xn);
        if (rc) return rc;
        eta = 0.0;
        for (int col = 0; col < k; ++col) {
            eta = std::max(eta, column_backward_error(rn[col], c.anorm, xn[col], bn[col]));
        }
        if (it > 0 && eta <= o.tol) break;
        if (it == o.max_iter) break;
        // with a fallback, a GMRES-IR cycle that does not halve the backward
        // error ends the refinement (the plain FP64 solve finishes the job)
        if (o.fallback && gmres_prev < INFINITY && !(eta <= 0.5 * gmres_prev)) break;
        gmres_prev = INFINITY;
        if (it > 1) {
            if (eta > 0.5 * prev) ++stall;
            else stall = 0;
            if (!gmres && m > 0 && (stall >= 2 || !(eta < prev * 4.0))) gmres = true;
        }
        prev = std::min(prev, eta);
        if (!gmres) {
            // d = M^{-1} r in float, x += d
            ELEMAMD_HIP_TRY(launch_copy_f64_to_f32_scaled(n, k, w.R, n, w.nrm + k, w.Rf, n, s));
            ELEMAMD_HIP_TRY(launch_sweep_solve(n, k, f.S, n, f.tiles, w.Rf, n, w.sweep, s));
            ELEMAMD_HIP_TRY(launch_axpy_f32_to_f64_scaled(n, k, w.Rf, n, w.nrm + k, w.X, n, s));
        } else {
            r.used_gmres = 1;
            gmres_prev = eta;
            int gi = 0;
            rc = gmres_cycle(c, m, 1e-8, &gi);
            if (rc) return rc;
            r.gmres_iterations += gi;
            // X += D
            ELEMAMD_HIP_TRY(launch_axpy_f64(n, k, 1.0, w.D, n, w.X, n, s));
        }
    }
    r.backward_error = eta;
    r.converged = (eta <= o.tol) ? 1 : 0;
    ELEMAMD_HIP_TRY(launch_copy_f64(n, k, w.X, n, B, ldb, s));
    int lost = 0;  // a correction sweep whose bounded wait expired
    ELEMAMD_HIP_TRY(sweep_read_error(n, w.sweep, &lost, s));
    if (lost) r.converged = 0;
    if (rep) *rep = r;
    if (lost) return kErrHip;
    return r.converged ? kOk : kErrNotConverged;
}

}  // namespace elemamd
