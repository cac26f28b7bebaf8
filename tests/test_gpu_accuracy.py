"""Accuracy of the benchmark configurations and of the pieces their numbers
rest on, through the C ABI.

* The FP64 residual R = B - A X (elemamd_sym_residual_f64) that the refinement
  and the benchmark's backward error use: bit for bit against the oracle's
  residual and an exact integer product, on both the one-pass and the split
  (fixed-order combine) forms of the kernel.
* The bounded waits of the persistent triangular solves: a wait that gives
  up (forced from the host with a tiny spin limit; every kernel still exits
  normally) is reported as an error, never as success.
* The mixed solve at the benchmark's shape (n = 16384, k = 32, kappa = 1e6,
  split-BF16x3 + CG): FP64 normwise backward error <= 1e-14 from a host
  residual, and forward error against the oracle's FP64 HPDSolve.
* The precision sweep at n = 32768, k = 8, kappa in {1e2, 1e4, 1e6, 1e8},
  FP32 and split-BF16x3 trailing updates, plus plain FP64.
* The low-precision factor's breakdown paths: the shifted refactor with
  GMRES-IR, and a non-HPD matrix with shifting off.
* NaN / Inf inputs: the mixed path reports non-convergence instead of a
  converged answer, and nothing hangs.
* The uneven batch at full size (8192 matrices, n log-uniform in
  [16, 1024]): every matrix reaches FP64 backward error <= 1e-14, cross-checked
  against the oracle on a sample.
* Herk / Trrk with alpha and beta other than -1 and 1 (the epilogues).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import elemental_amd as E  # noqa: E402
from elemental_amd import Matrix, inputs, lib  # noqa: E402
from elemental_amd import batch as EB  # noqa: E402

EPS = np.finfo(np.float64).eps


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def colmajor(t):
    out = torch.empty(t.shape[1], t.shape[0], dtype=t.dtype, device=t.device).t()
    out.copy_(t)
    return out


def dev(a, ld=None):
    return Matrix.from_numpy(np.asfortranarray(a), ldim=ld)


def host(M):
    torch.cuda.synchronize()
    return M.to_numpy()


def eta_inf_host(A_full, X, B):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf),
    with numpy on the host (A given as the full symmetric matrix)."""
    R = B - A_full @ X
    an = np.abs(A_full).sum(axis=1).max()
    return float(max(np.abs(R[:, c]).max() / (an * np.abs(X[:, c]).max() + np.abs(B[:, c]).max())
                     for c in range(B.shape[1])))


# ------------------------------------------------------- FP64 residual kernel
def _int_problem(n, k, seed):
    rng = np.random.default_rng(seed)
    Lw = np.tril(rng.integers(-9, 10, (n, n)).astype(np.float64))
    Asym = Lw + np.tril(Lw, -1).T
    # the strictly upper triangle holds values the kernel must not read
    Astored = np.asfortranarray(Lw + np.triu(rng.integers(500, 900, (n, n)).astype(np.float64), 1))
    X = np.asfortranarray(rng.integers(-7, 8, (n, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-60, 61, (n, k)).astype(np.float64))
    return Asym, Astored, X, B


@pytest.mark.parametrize("n,k", [(1, 1), (50, 3), (64, 32), (777, 33), (1536, 8), (2048, 32)])
def test_sym_residual_bitexact(oracle, n, k):
    """n <= 64 runs the one-pass kernel, larger n the split sweep with the
    fixed-order combine; on integers every form is exact."""
    Asym, Astored, X, B = _int_problem(n, k, seed=n * 7 + k)
    ref = oracle.sym_lower_residual(Astored, X, B)
    assert np.array_equal(ref, B - Asym @ X)
    Ad, Xd, Bd = dev(Astored, ld=n + 3), dev(X), dev(B, ld=n + 1)
    Rd = Matrix(n, k)
    assert lib().elemamd_sym_residual_f64(n, k, Ad.ptr(), Ad.LDim(), Xd.ptr(), Xd.LDim(), Bd.ptr(), Bd.LDim(),
                                          Rd.ptr(), Rd.LDim(), S()) == 0
    assert np.array_equal(host(Rd), ref)


@pytest.mark.slow
@pytest.mark.parametrize("n,k", [(4096, 32), (16384, 32)])
def test_sym_residual_integer_exact_full_size(n, k):
    g = torch.Generator(device="cuda")
    g.manual_seed(n + k)
    L = torch.tril(torch.randint(-9, 10, (n, n), generator=g, device="cuda", dtype=torch.int32).to(torch.float64))
    Asym = L + torch.tril(L, -1).t()
    Astored = colmajor(L + torch.triu(torch.full((n, n), 777.0, dtype=torch.float64, device="cuda"), 1))
    del L
    X = colmajor(torch.randint(-7, 8, (n, k), generator=g, device="cuda", dtype=torch.int32).to(torch.float64))
    B = colmajor(torch.randint(-60, 61, (n, k), generator=g, device="cuda", dtype=torch.int32).to(torch.float64))
    ref = B - Asym @ X  # integers below 2^53: exact in any order
    del Asym
    R = colmajor(torch.zeros(n, k, dtype=torch.float64, device="cuda"))
    assert lib().elemamd_sym_residual_f64(n, k, P(Astored), n, P(X), n, P(B), n, P(R), n, S()) == 0
    torch.cuda.synchronize()
    assert torch.equal(R, ref)


# ---------------------------------------------------- bounded-wait reporting
def test_solve_reports_expired_wait(monkeypatch):
    """With a spin limit of one poll the persistent solve's helpers give up
    their waits (and still exit: every access is in bounds whatever the
    flags say); SolveAfter, HPDSolve, Trsm and Trsv must return
    ELEMAMD_ERR_HIP.  With the default limit the same calls succeed."""
    n, k = 4096, 16
    A = inputs.gram_host(n, seed=5)
    L = np.asfortranarray(np.linalg.cholesky(A))
    B = inputs.normal_host(n, k, seed=6)
    Ld = dev(L)
    monkeypatch.setenv("ELEMAMD_SPIN_LIMIT", "1")
    Bd = dev(B)
    assert lib().elemamd_solve_after_f64(b"L", b"N", n, k, Ld.ptr(), n, Bd.ptr(), n, S()) == -2
    Bd = dev(B)
    assert lib().elemamd_trsm_f64(b"L", b"L", b"N", b"N", n, k, 1.0, Ld.ptr(), n, Bd.ptr(), n, S()) == -2
    xd = dev(B[:, :1])
    assert lib().elemamd_trsv_f64(b"L", b"T", b"N", n, Ld.ptr(), n, xd.ptr(), 1, S()) == -2
    Ad, Bd = dev(A), dev(B)
    d = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib().elemamd_hpd_solve_f64(b"L", b"N", n, k, Ad.ptr(), n, Bd.ptr(), n, 128, P(d), S()) == -2
    monkeypatch.delenv("ELEMAMD_SPIN_LIMIT")
    Bd = dev(B)
    assert lib().elemamd_solve_after_f64(b"L", b"N", n, k, Ld.ptr(), n, Bd.ptr(), n, S()) == 0
    X = host(Bd)
    assert eta_inf_host(A, X, B) <= 1e-15


# ------------------------------------------------- the benchmark's workload
@pytest.mark.slow
def test_mixed_c2_full_size(oracle):
    """n = 16384, k = 32, A = Q diag(s) Q^T with s log-spaced 1 .. 1e-6:
    split-BF16x3 factor + CG refinement reaches FP64 backward error <= 1e-14
    (host residual), and agrees with the oracle's FP64 HPDSolve to the
    accuracy the condition number allows."""
    n, k, kappa = 16384, 32, 1e6
    Ad = inputs.spectrum_device(n, kappa, seed=2024)
    Bd = inputs.normal_device(n, k, seed=2025)
    torch.cuda.synchronize()
    A, B = Ad.to_numpy(), Bd.to_numpy()
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", refine="cg"))
    X = host(Bd)
    assert rep.converged and rep.status == 0, rep
    assert rep.iterations <= 12, rep
    eta = eta_inf_host(A, X, B)
    assert eta <= 1e-14, (eta, rep)
    oracle.bind_host_blas(16)
    Lr, Xr = A.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve(Lr, Xr, 128, host_blas=True) == 0
    fwd = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    assert fwd <= 100 * kappa * EPS, fwd
    # plain FP64 on the same matrix: the factor against the oracle's
    Pd, Xd = dev(A), dev(B)
    E.HPDSolve(E.LOWER, E.NORMAL, Pd, Xd)
    Lg, Xg = host(Pd), host(Xd)
    assert np.linalg.norm(np.tril(Lg - Lr)) / np.linalg.norm(np.tril(Lr)) <= 1e-12
    assert eta_inf_host(A, Xg, B) <= 1e-15


@pytest.mark.slow
def test_precision_sweep_n32768():
    """n = 32768, k = 8: FP32 and split-BF16x3 trailing updates at kappa =
    1e2 .. 1e8 all reach FP64 backward error <= 1e-14 (host residual) with
    bounded refinement work; plain FP64 at kappa = 1e6 likewise."""
    n, k = 32768, 8
    G = inputs.normal_device(n, n, seed=3301).tensor().clone()
    Q, R = torch.linalg.qr(G)
    Q.mul_(torch.sign(torch.diagonal(R)).unsqueeze(0))
    del R, G
    Bd = inputs.normal_device(n, k, seed=3302)
    torch.cuda.synchronize()
    B = Bd.to_numpy()
    for kappa in (1e2, 1e4, 1e6, 1e8):
        s = torch.logspace(0.0, -float(np.log10(kappa)), n, dtype=torch.float64, device="cuda")
        A = Matrix(n, n)
        At = A.tensor()
        At.copy_((Q * s.unsqueeze(0)) @ Q.t())
        At.copy_(0.5 * (At + At.t()))
        torch.cuda.synchronize()
        Ah = A.to_numpy()
        for trailing in ("f32", "bf16x3"):
            X = Matrix.from_numpy(B)
            refine = "cg" if kappa <= 1e6 else "ir"
            rep = E.HPDSolveMixed(A, X, E.IROptions(trailing=trailing, refine=refine))
            assert rep.converged, (kappa, trailing, rep)
            assert rep.iterations <= 30 and rep.gmres_iterations <= 400, (kappa, trailing, rep)
            eta = eta_inf_host(Ah, host(X), B)
            print(f"kappa={kappa:.0e} {trailing}: eta={eta:.2e} iters={rep.iterations} "
                  f"gmres={rep.gmres_iterations} shifted={rep.shifted}")
            assert eta <= 1e-14, (kappa, trailing, eta, rep)
        if kappa == 1e6:
            X = Matrix.from_numpy(B)
            E.HPDSolve(E.LOWER, E.NORMAL, A, X)
            assert eta_inf_host(Ah, host(X), B) <= 1e-15
        del A, At, Ah


# ------------------------------------------------- low-precision breakdown
def test_mixed_breakdown_falls_back_to_fp64(oracle):
    """kappa = 1e9 breaks the split-BF16x3 factor of A (a non-positive float
    pivot).  The dispatch rule finishes in plain FP64 (Cholesky + SolveAfter
    on a copy of A): FP64 backward error, X as the oracle's FP64 HPDSolve, A
    untouched.  n = 2048 so the sweep preparation runs beside the
    factorization on the third stream."""
    n, k = 2048, 4
    A = inputs.spectrum_host(n, 1e9, seed=71)
    B = inputs.normal_host(n, k, seed=72)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", refine="ir"))
    assert rep.lowprec_info > 0 and not rep.shifted, rep
    assert rep.fallback and rep.converged and rep.status == 0 and rep.info == 0, rep
    X = host(Bd)
    eta = eta_inf_host(A, X, B)
    assert eta <= 1e-14 and rep.backward_error <= 1e-14, (eta, rep)  # both at the rounding level
    Lr, Xr = A.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve(Lr, Xr, 128) == 0
    # both are FP64 backward-stable solves of a kappa = 1e9 system
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e9 * 100 * EPS
    assert np.array_equal(host(Ad), A)


def test_mixed_shifted_refactor_runs_then_falls_back():
    """allow_shift: after the breakdown the factor of A + sigma diag(A)
    succeeds and GMRES-IR preconditioned by it makes progress; with the
    fallback switched off the call stops short of the target (the float
    preconditioner's rounding keeps GMRES-IR near 1e-8 at this kappa), with
    it on the FP64 route reaches it."""
    n, k = 2048, 4
    A = inputs.spectrum_host(n, 1e9, seed=71)
    B = inputs.normal_host(n, k, seed=72)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", refine="ir", max_iter=8, allow_shift=True,
                                              fallback=False))
    assert rep.shifted and rep.lowprec_info == 0 and rep.used_gmres and not rep.fallback, rep
    assert rep.status in (0, -3) and rep.backward_error <= 1e-6, rep
    Bd = dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", refine="ir", allow_shift=True))
    assert rep.shifted and rep.converged, rep
    assert eta_inf_host(A, host(Bd), B) <= 1e-14


def test_mixed_non_hpd_without_shift_reports_column():
    n, k, col = 2048, 3, 1500
    A = inputs.gram_host(n, seed=81)
    A[col, col] = -1.0
    B = inputs.normal_host(n, k, seed=82)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", allow_shift=False))
    assert rep.status == -3 and not rep.converged, rep
    assert rep.lowprec_info == col + 1, rep
    assert not rep.shifted
    # the FP64 fallback meets the same column: A is not numerically HPD
    assert rep.fallback and rep.info == col + 1, rep
    rep = E.HPDSolveMixed(dev(A), dev(B), E.IROptions(trailing="bf16x3", fallback=False))
    assert rep.status == -3 and rep.lowprec_info == col + 1 and not rep.fallback and rep.info == 0, rep


@pytest.mark.parametrize("where", ["b_nan", "a_inf"])
def test_mixed_nonfinite_input_not_converged(where):
    """A NaN in B or an Inf on A's diagonal: the mixed path reports
    non-convergence (the reference would return NaNs without a word); the
    call returns, and the next solve on the same stream is correct."""
    n, k = 700, 2
    A = inputs.spectrum_host(n, 1e4, seed=91)
    B = inputs.normal_host(n, k, seed=92)
    Ab, Bb = A.copy(order="F"), B.copy(order="F")
    if where == "b_nan":
        Bb[17, 1] = np.nan
    else:
        Ab[300, 300] = np.inf
    rep = E.HPDSolveMixed(dev(Ab), dev(Bb), E.IROptions(trailing="f32", refine="cg", max_iter=10))
    assert not rep.converged and rep.status == -3, rep
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="f32", refine="cg"))
    assert rep.converged and eta_inf_host(A, host(Bd), B) <= 1e-14


# ----------------------------------------------------- uneven batch at size
@pytest.mark.slow
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_8192_uneven(oracle, precision):
    n, k = EB.uneven_sizes(8192, 16, 1024, 8, seed=4242)
    plan = EB.BatchPlan(n, k, precision)
    A, B = plan.empty_a(), plan.empty_b()
    EB.fill_gram(plan, A, B, seed=50000)
    A0, B0 = A.clone(), B.clone()
    plan.solve(A, B, use_graph=True)
    info, eta_dev, sweeps = plan.results()
    assert np.all(info == 0)
    eta = plan.backward_error(A0, B, B0)
    assert np.all(eta <= 1e-14), (eta.max(), int(np.argmax(eta)))
    if precision != "f64":
        assert np.all(eta_dev <= 1e-14)
    # cross-check the device backward error and the solution on a sample
    oracle.bind_host_blas(8)
    rng = np.random.default_rng(1)
    for i in rng.choice(len(n), 12, replace=False):
        Ai = plan.a_view(A0, int(i)).cpu().numpy().copy(order="F")
        Bi = plan.b_view(B0, int(i)).cpu().numpy().copy(order="F")
        Xi = plan.b_view(B, int(i)).cpu().numpy()
        e = eta_inf_host(Ai, Xi, Bi)
        # the two residuals round differently: compare at the rounding level
        assert e <= 1e-14 and abs(e - eta[i]) <= 0.5 * e + 1e-16, (i, e, eta[i])
        Lr, Xr = Ai.copy(order="F"), Bi.copy(order="F")
        assert oracle.hpd_solve(Lr, Xr, 128, host_blas=True) == 0
        assert np.linalg.norm(Xi - Xr) / np.linalg.norm(Xr) <= 1e-12


# ------------------------------------------------------ epilogue scalars
@pytest.mark.parametrize("alpha,beta", [(0.5, -2.0), (-1.25, 0.0), (3.0, 1.0)])
def test_herk_f64_alpha_beta(oracle, alpha, beta):
    n, k = 300, 70
    A = oracle.normal((n, k), seed=11)
    C = oracle.normal((n, n), seed=12)
    ref = C.copy(order="F")
    oracle.dsyrk("N", alpha, A, beta, ref)
    Ad, Cd = dev(A), dev(C, ld=n + 5)
    assert lib().elemamd_herk_f64(b"L", b"N", n, k, alpha, Ad.ptr(), Ad.LDim(), beta, Cd.ptr(), Cd.LDim(), S()) == 0
    got = host(Cd)
    scale = abs(alpha) * np.abs(A).max() ** 2 * k + abs(beta) * np.abs(C).max()
    assert np.max(np.abs(np.tril(got - ref))) <= 8 * k * EPS * scale
    assert np.array_equal(np.triu(got, 1), np.triu(C, 1))


@pytest.mark.parametrize("alpha,beta", [(2.0, 0.25), (-0.5, 0.0), (1.0, -3.0)])
def test_trrk_alpha_beta(oracle, alpha, beta):
    n, k = 260, 40
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((k, n)))
    Bm = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    ref = C.copy(order="F")
    oracle.trrk_tn(alpha, A, Bm, beta, ref, 64)
    Cd = dev(C)
    E.Trrk(E.LOWER, E.TRANSPOSE, E.NORMAL, alpha, dev(A), dev(Bm), beta, Cd)
    got = host(Cd)
    assert np.max(np.abs(np.tril(got - ref))) <= 16 * k * EPS * (abs(alpha) * np.abs(A).max() * np.abs(Bm).max() * k
                                                               + abs(beta) * np.abs(C).max())
    assert np.array_equal(np.triu(got, 1), np.triu(C, 1))
