"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: bit-exact where the GPU kernel performs the reference's own operation
sequence (LVar3Unb diagonal block, the R/L/T panel solve) and for the
integer-exact known answers at any size; otherwise the north-star
tolerances: ||L - L_ref||/||L_ref|| <= 1e-12, relative residual <= 1e-13,
FP64 normwise backward error <= 1e-14 for the refined mixed-precision solve.
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

EPS = np.finfo(np.float64).eps


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dev(a, ld=None):
    return Matrix.from_numpy(np.asfortranarray(a), ldim=ld)


def host(M):
    torch.cuda.synchronize()
    return M.to_numpy()


def full_host(M):
    """The whole ld x width buffer (to check padding and the upper triangle)."""
    torch.cuda.synchronize()
    return M._st.cpu().numpy()


# ------------------------------------------------------------ diagonal block
@pytest.mark.parametrize("n", [1, 2, 7, 64, 127, 128])
@pytest.mark.parametrize("kind", ["gram", "integer"])
def test_potrf_unb_bitexact(oracle, n, kind):
    A = inputs.gram_host(n, seed=n) if kind == "gram" else inputs.integer_factor(n, seed=n)[0]
    ref = A.copy(order="F")
    assert oracle.lvar3unb(ref) == 0
    M = dev(A, ld=n + 3)
    d = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib().elemamd_potrf_unb_f64(n, M.ptr(), M.LDim(), ctypes.c_void_p(d.data_ptr()), S()) == 0
    assert lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), S()) == 0
    assert np.array_equal(host(M), ref)  # lower = L bit for bit, upper untouched


@pytest.mark.parametrize("col", [0, 9, 100])
def test_potrf_unb_failing_column(oracle, col):
    n = 128
    A = inputs.gram_host(n, seed=3)
    A[col, col] = -1.0
    ref = A.copy(order="F")
    info = oracle.lvar3unb(ref)
    # only a_cc changed: pivots before it are those of the HPD matrix, and
    # pivot c is -1 minus a sum of squares
    assert info == col + 1
    M = dev(A)
    d = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib().elemamd_potrf_unb_f64(n, M.ptr(), n, ctypes.c_void_p(d.data_ptr()), S()) == 0
    assert lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), S()) == info
    assert np.array_equal(host(M), ref)  # same partially factored state as the reference


# --------------------------------------------------------------- panel trsm
@pytest.mark.parametrize("m", [1, 63, 64, 65, 1000])
@pytest.mark.parametrize("nb", [1, 7, 64, 127, 128])
def test_trsm_panel_bitexact(oracle, m, nb):
    L = np.linalg.cholesky(inputs.gram_host(nb, seed=nb))
    L = np.asfortranarray(L)
    B = oracle.normal((m, nb), seed=m * 131 + nb)
    ref = B.copy(order="F")
    oracle.dtrsm("R", "T", 1.0, L, ref)
    Ld, Bd = dev(L), dev(B, ld=m + 1)
    assert lib().elemamd_trsm_f64(b"R", b"L", b"T", b"N", m, nb, 1.0, Ld.ptr(), Ld.LDim(), Bd.ptr(), Bd.LDim(),
                                  S()) == 0
    assert np.array_equal(host(Bd), ref)


# -------------------------------------------------------------- trailing herk
@pytest.mark.parametrize("n,k", [(1, 1), (100, 16), (128, 128), (129, 130), (515, 128), (1000, 64)])
def test_herk_f64(oracle, n, k):
    A = oracle.normal((n, k), seed=n + k)
    C = oracle.normal((n, n), seed=7 * n)
    ref = C.copy(order="F")
    oracle.dsyrk("N", -1.0, A, 1.0, ref)
    Ad, Cd = dev(A), dev(C, ld=n + 2)
    assert lib().elemamd_herk_f64(b"L", b"N", n, k, -1.0, Ad.ptr(), Ad.LDim(), 1.0, Cd.ptr(), Cd.LDim(), S()) == 0
    got = host(Cd)
    scale = np.abs(A).max() ** 2 * k + np.abs(C).max()
    assert np.max(np.abs(np.tril(got - ref))) <= 8 * k * EPS * scale
    assert np.array_equal(np.triu(got, 1), np.triu(C, 1))


def test_herk_f64_integer_exact():
    rng = np.random.default_rng(1)
    n, k = 777, 200
    A = np.asfortranarray(rng.integers(-7, 8, (n, k)).astype(float))
    C = np.asfortranarray(rng.integers(-9, 10, (n, n)).astype(float))
    Ad, Cd = dev(A), dev(C)
    assert lib().elemamd_herk_f64(b"L", b"N", n, k, -1.0, Ad.ptr(), n, 1.0, Cd.ptr(), n, S()) == 0
    got = host(Cd)
    assert np.array_equal(np.tril(got), np.tril(C - A @ A.T))
    assert np.array_equal(np.triu(got, 1), np.triu(C, 1))


@pytest.mark.parametrize("mode,tol", [(1, 2e-6), (2, 2e-4), (3, 2e-6)])
def test_herk_lowprec(oracle, mode, tol):
    # FP32 MFMA (mode 1), split-BF16x3 (mode 2) and split-BF16x6 (mode 3)
    # against the FP64 product; tolerance relative to k*max|a|^2: ~ float
    # eps, ~2^-16 and ~ float eps respectively
    n, k = 700, 128
    A = oracle.normal((n, k), seed=5)
    C = oracle.normal((n, n), seed=6)
    ref = np.tril(C - A @ A.T)
    Af = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()  # column-major n x k
    Cf = torch.from_numpy(np.ascontiguousarray(C.T)).float().cuda()
    assert lib().elemamd_herk_f32(mode, n, k, -1.0, ctypes.c_void_p(Af.data_ptr()), n, 1.0,
                                  ctypes.c_void_p(Cf.data_ptr()), n, S()) == 0
    torch.cuda.synchronize()
    got = Cf.double().cpu().numpy().T
    err = np.max(np.abs(np.tril(got) - ref)) / (k * np.abs(A).max() ** 2)
    assert err < tol
    assert np.array_equal(np.triu(got, 1), np.triu(C.astype(np.float32).astype(np.float64), 1))


# -------------------------------------------------------------- factorization
@pytest.mark.parametrize("n,nb", [(1, 128), (2, 128), (127, 128), (129, 128), (300, 64), (1000, 128), (2049, 128),
                                  (600, 1), (512, 37)])
def test_cholesky_vs_oracle(oracle, n, nb):
    A = inputs.gram_host(n, seed=n + nb)
    ref = A.copy(order="F")
    assert oracle.lvar3(ref, nb) == 0
    M = dev(A, ld=n + 5)
    E.Cholesky(E.LOWER, M, nb=nb)
    got = host(M)
    Lr = np.tril(ref)
    assert np.linalg.norm(np.tril(got) - Lr) / np.linalg.norm(Lr) <= 1e-12
    assert np.array_equal(np.triu(got, 1), np.triu(A, 1))


@pytest.mark.parametrize("n,nb", [(129, 128), (1000, 64), (3000, 128)])
def test_cholesky_integer_exact(n, nb):
    A, L0 = inputs.integer_factor(n, seed=n)
    M = dev(A)
    E.Cholesky(E.LOWER, M, nb=nb)
    got = host(M)
    assert np.array_equal(np.tril(got), L0)
    assert np.array_equal(np.triu(got, 1), np.triu(A, 1))


@pytest.mark.parametrize("col", [0, 127, 128, 300])
def test_cholesky_non_hpd_raises_like_reference(oracle, col):
    n = 400
    A = inputs.gram_host(n, seed=9)
    L = np.linalg.cholesky(A)
    A[col, col] -= 1.5 * L[col, col] ** 2
    ref = A.copy(order="F")
    info = oracle.lvar3(ref, 128)
    assert info == col + 1
    M = dev(A)
    with pytest.raises(E.LogicError, match="A was not numerically HPD") as ei:
        E.Cholesky(E.LOWER, M)
    assert ei.value.column == col
    got = host(M)
    # blocks before the failing one are factored like the reference's
    b0 = (col // 128) * 128
    if b0 > 0:
        assert np.linalg.norm(np.tril(got[:, :b0] - ref[:, :b0])) <= 1e-12 * np.linalg.norm(np.tril(ref[:, :b0]))


# -------------------------------------------------------------------- solves
@pytest.mark.parametrize("n,k", [(1, 1), (7, 1), (128, 3), (300, 16), (1000, 33), (2048, 16)])
def test_solve_after_vs_oracle(oracle, n, k):
    A = inputs.gram_host(n, seed=n)
    L = A.copy(order="F")
    assert oracle.lvar3(L, 128) == 0
    B = oracle.normal((n, k), seed=k)
    ref = B.copy(order="F")
    oracle.solve_after(L, ref)
    Ld, Bd = dev(L), dev(B, ld=n + 1)
    E.SolveAfter(E.LOWER, E.NORMAL, Ld, Bd)
    X = host(Bd)
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) <= 1e-13
    R = oracle.sym_lower_residual(A, X, B)
    assert np.linalg.norm(R) / np.linalg.norm(B) <= 1e-13


@pytest.mark.parametrize("n,k", [(129, 1), (1000, 5), (2500, 32)])
def test_solve_after_integer_exact(n, k):
    A, L0 = inputs.integer_factor(n, seed=n + 1)
    rng = np.random.default_rng(n)
    X0 = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(float))
    Bd = dev(A @ X0)
    E.SolveAfter(E.LOWER, E.NORMAL, dev(L0), Bd)
    assert np.array_equal(host(Bd), X0)


@pytest.mark.parametrize("trans", [b"N", b"T"])
def test_trsv_vs_oracle(oracle, trans):
    n = 700
    L = np.asfortranarray(np.linalg.cholesky(inputs.gram_host(n, seed=2)))
    x = oracle.normal((n, 1), seed=4)
    ref = x.copy(order="F")
    oracle.dtrsv(trans.decode(), L, ref)
    Ld, xd = dev(L), dev(x)
    assert lib().elemamd_trsv_f64(b"L", trans, b"N", n, Ld.ptr(), n, xd.ptr(), 1, S()) == 0
    assert np.linalg.norm(host(xd) - ref) / np.linalg.norm(ref) <= 1e-13


def test_trsm_left_vs_oracle(oracle):
    n, k = 900, 12
    L = np.asfortranarray(np.linalg.cholesky(inputs.gram_host(n, seed=8)))
    B = oracle.normal((n, k), seed=5)
    for t in ("N", "T"):
        ref = B.copy(order="F")
        oracle.dtrsm("L", t, 2.0, L, ref)
        Bd = dev(B)
        E.Trsm(E.LEFT, E.LOWER, E.NORMAL if t == "N" else E.TRANSPOSE, E.NON_UNIT, 2.0, dev(L), Bd)
        assert np.linalg.norm(host(Bd) - ref) / np.linalg.norm(ref) <= 1e-13


@pytest.mark.parametrize("n,k", [(10, 4), (130, 16), (300, 128), (700, 33)])
def test_trrk_vs_oracle(oracle, n, k):
    A = oracle.normal((k, n), seed=n)
    B = oracle.normal((k, n), seed=n + 1)
    C = oracle.normal((n, n), seed=n + 2)
    ref = C.copy(order="F")
    oracle.trrk_tn(-1.0, A, B, 1.0, ref, 64)
    Cd = dev(C)
    E.Trrk(E.LOWER, E.TRANSPOSE, E.NORMAL, -1.0, dev(A), dev(B), 1.0, Cd)
    got = host(Cd)
    assert np.max(np.abs(np.tril(got - ref))) <= 16 * k * EPS * (np.abs(A).max() * np.abs(B).max() * k + 1)
    assert np.array_equal(np.triu(got, 1), np.triu(C, 1))


# --------------------------------------------------- n = 2048 factor + solve
def test_hpd_solve_n2048_k16(oracle):
    """A = G G^T/n + I, B 2048 x 16: relative residual <= 1e-13 and
    ||L - L_ref||/||L_ref|| <= 1e-12 against the oracle's FP64 LVar3."""
    n, k = 2048, 16
    A = inputs.gram_host(n, seed=12345)
    B = inputs.normal_host(n, k, seed=777)
    Lr, Xr = A.copy(order="F"), B.copy(order="F")
    oracle.bind_host_blas(8)
    assert oracle.hpd_solve(Lr, Xr, 128, host_blas=True) == 0
    Ad, Bd = dev(A), dev(B)
    E.HPDSolve(E.LOWER, E.NORMAL, Ad, Bd)
    L, X = host(Ad), host(Bd)
    assert np.linalg.norm(np.tril(L - Lr)) / np.linalg.norm(np.tril(Lr)) <= 1e-12
    R = oracle.sym_lower_residual(A, X, B)
    assert np.linalg.norm(R) / np.linalg.norm(B) <= 1e-13
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-12


# ---------------------------------------------------------- mixed precision
def _eta_inf(A, X, B, oracle):
    R = oracle.sym_lower_residual(A, X, B)
    anorm = np.abs(A).sum(axis=1).max()
    return max(np.abs(R[:, c]).max() / (anorm * np.abs(X[:, c]).max() + np.abs(B[:, c]).max())
               for c in range(B.shape[1]))


@pytest.mark.parametrize("trailing,kappa,gmres,refine", [("f32", 1e6, False, "ir"), ("bf16x3", 1e4, False, "ir"),
                                                         ("f32", 1e6, True, "ir"), ("bf16x3", 1e6, False, "ir"),
                                                         ("bf16x6", 1e6, False, "ir"), ("f32", 1e6, False, "cg"),
                                                         ("bf16x3", 1e6, False, "cg"), ("bf16x3", 1e8, False, "cg")])
def test_mixed_ir_reaches_fp64_backward_error(oracle, trailing, kappa, gmres, refine):
    n, k = 1536, 8
    A = inputs.spectrum_host(n, kappa, seed=21)
    B = inputs.normal_host(n, k, seed=22)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing=trailing, force_gmres=gmres, refine=refine))
    X = host(Bd)
    assert rep.converged, rep
    eta = _eta_inf(A, X, B, oracle)
    assert eta <= 1e-14, (eta, rep)
    assert np.array_equal(host(Ad), A)  # A is not modified
    # forward error against the oracle's FP64 solution: both are backward
    # stable, so they agree to ~kappa * eps
    Lr, Xr = A.copy(order="F"), B.copy(order="F")
    oracle.bind_host_blas(8)
    assert oracle.hpd_solve(Lr, Xr, 128, host_blas=True) == 0
    fwd = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    assert fwd <= 100 * kappa * EPS


@pytest.mark.parametrize("trailing,nb", [("bf16x3", 64), ("bf16x3", 96), ("bf16x3", 37), ("bf16x6", 64),
                                         ("bf16x3", 128)])
def test_mixed_block_sizes(oracle, trailing, nb):
    """Plane-path block columns of 64 and 96 (look-ahead update of the rows
    below the diagonal block when nb is a multiple of 64), the in-kernel split
    for nb = 37, on a matrix with a partial last block."""
    n, k = 1000, 3
    A = inputs.spectrum_host(n, 1e4, seed=61 + nb)
    B = inputs.normal_host(n, k, seed=62 + nb)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing=trailing, nb=nb, gmres_restart=-1))
    assert rep.converged and not rep.used_gmres and not rep.shifted, rep
    assert rep.iterations <= 8, rep
    assert _eta_inf(A, host(Bd), B, oracle) <= 1e-14


@pytest.mark.parametrize("n,k", [(2048, 4), (777, 33)])
def test_mixed_cg_fewer_products_than_ir(oracle, n, k):
    """Preconditioned CG with the split-BF16x3 factor: fewer FP64 products
    with A than classic refinement to the same backward error, every column
    (33 = two correction chunks) confirmed by the true residual."""
    A = inputs.spectrum_host(n, 1e6, seed=51 + n)
    B = inputs.normal_host(n, k, seed=52 + n)
    its = {}
    for refine in ("ir", "cg"):
        Ad, Bd = dev(A), dev(B)
        rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="bf16x3", refine=refine, gmres_restart=-1))
        assert rep.converged and not rep.used_gmres, (refine, rep)
        assert _eta_inf(A, host(Bd), B, oracle) <= 1e-14
        its[refine] = rep.iterations
    assert its["cg"] < its["ir"], its


def test_mixed_bf16x6_between_bf16x3_and_fp32():
    """Split-BF16x6 keeps every product term down to 2^-24, split-BF16x3 the
    terms down to 2^-16: the x6 factor is the better preconditioner, so
    refinement takes fewer steps than with x3, and no fewer than with the
    FP32 MFMA (the bf16 MFMA's truncated product sum keeps x6 short of FP32)."""
    n, k = 2048, 4
    A = inputs.spectrum_host(n, 1e6, seed=41)
    B = inputs.normal_host(n, k, seed=42)
    its = {}
    for trailing in ("f32", "bf16x3", "bf16x6"):
        Ad, Bd = dev(A), dev(B)
        rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing=trailing, gmres_restart=-1))
        assert rep.converged and not rep.used_gmres, (trailing, rep)
        its[trailing] = rep.iterations
    assert its["f32"] <= its["bf16x6"] < its["bf16x3"], its


@pytest.mark.parametrize("n,k", [(64, 1), (200, 3), (256, 17), (777, 33), (1300, 40), (2048, 32)])
def test_mixed_ir_correction_sweeps_shapes(oracle, n, k):
    """The block-inverse correction sweeps over partial blocks, one or two
    16-column halves and several 32-column chunks: with kappa = 1e3 and an
    FP32 factor each refinement step gains ~kappa * 2^-24, so three or four
    classic steps reach 1e-14 unless a sweep is wrong."""
    A = inputs.spectrum_host(n, 1e3, seed=31 + n)
    B = inputs.normal_host(n, k, seed=32 + n)
    Ad, Bd = dev(A), dev(B)
    rep = E.HPDSolveMixed(Ad, Bd, E.IROptions(trailing="f32"))
    X = host(Bd)
    assert rep.converged and not rep.used_gmres, rep
    assert rep.iterations <= 5, rep
    assert _eta_inf(A, X, B, oracle) <= 1e-14


# ------------------------------------------------------------ full size
@pytest.mark.slow
def test_full_size_integer_factor_and_solve_exact():
    """n = 16384: the factor of A = L0 L0^T must be L0 exactly and the
    solution of A X = A X0 must be X0 exactly (all intermediates are
    integers below 2^53).  Cholesky + SolveAfter (substitution inside the
    diagonal blocks); HPDSolve's solves multiply by the blocks' inverses,
    whose entries grow far past 2^53 for these integer factors."""
    n, k = 16384, 32
    A, L0 = inputs.integer_device(n, seed=5)
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    X0 = torch.randint(-3, 4, (n, k), generator=g, device="cuda", dtype=torch.int32).to(torch.float64)
    B = Matrix(n, k)
    B.tensor().copy_(A.tensor() @ X0)
    upper_before = torch.triu(A.tensor(), 1).clone()
    E.Cholesky(E.LOWER, A)
    E.SolveAfter(E.LOWER, E.NORMAL, A, B)
    torch.cuda.synchronize()
    assert torch.equal(torch.tril(A.tensor()), L0.tensor())
    assert torch.equal(torch.triu(A.tensor(), 1), upper_before)
    assert torch.equal(B.tensor(), X0)


def test_device_normal_generator_matches_host():
    a = inputs.normal_device(301, 7, seed=42, offset=3)
    b = inputs.normal_host(301, 7, seed=42, offset=3)
    assert np.max(np.abs(host(a) - b)) <= 8 * EPS * np.abs(b).max()


def test_cpp_dropin_api():
    """The C++ drop-in header (elem::Matrix, elem::HPDSolve, elem::Cholesky on a
    view, the non-HPD logic_error, elem::Trsm) driven by a compiled program."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "elem_api_test")
    assert os.path.exists(exe), "build() compiles tests/cpp/elem_api_test"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
