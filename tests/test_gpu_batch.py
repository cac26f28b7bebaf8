"""Batched uneven-size factor-and-solve through the C ABI.

Bars: every matrix in a batch gives bit-for-bit the result it gives in a
batch of its own (the batch shares launches, never arithmetic); the plain
FP64 batch agrees with the single-matrix path (factor within 1e-13, solution
within 1e-12 relative: the single path groups its trailing updates and
blocks its panel solves differently) and with the oracle's HPDSolve
(1e-12 relative); every matrix of the refinement modes reaches FP64 normwise
backward error <= 1e-14; hipGraph replays equal the eager launch sequence;
a non-HPD matrix reports its failing column and leaves the others intact.
The persistent launch form (one launch per bucket, workgroups taking whole
matrices) gives the stepwise form's results bit for bit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from elemental_amd import batch as EB  # noqa: E402
from elemental_amd import lib  # noqa: E402

SEED = 7100


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def make(n, k, precision="f64", buckets=None, seed=SEED):
    plan = EB.BatchPlan(n, k, precision, buckets)
    A, B = plan.empty_a(), plan.empty_b()
    EB.fill_gram(plan, A, B, seed)
    torch.cuda.synchronize()
    return plan, A, B


def colmajor_copy(t):
    """A new column-major (unit row stride) copy of a 2-D tensor."""
    out = torch.empty(t.shape[1], t.shape[0], dtype=t.dtype, device=t.device).t()
    out.copy_(t)
    return out


def rel(a, b):
    d = (a - b).abs().max().item()
    s = b.abs().max().item()
    return d / s if s > 0 else d


SIZES = dict(count=48, lo=16, hi=300, kmax=8, seed=11)


def sizes(**kw):
    a = dict(SIZES)
    a.update(kw)
    return EB.uneven_sizes(a["count"], a["lo"], a["hi"], a["kmax"], a["seed"])


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16x3"])
def test_alone_equals_in_batch(precision):
    n, k = sizes()
    plan, A, B = make(n, k, precision)
    A0 = A.clone()
    plan.set_refinement(6)
    plan.solve(A, B, use_graph=False)
    info, eta, _ = plan.results()
    assert (info == 0).all()
    for i in [0, 5, 17, len(n) - 1]:
        one = EB.BatchPlan(n[i:i + 1], k[i:i + 1], precision, plan.buckets)
        one.set_refinement(6)
        A1, B1 = one.empty_a(), one.empty_b()
        EB.fill_gram(one, A1, B1, SEED + i)  # matrix i of the batch used stream SEED + i
        assert torch.equal(one.a_view(A1, 0), plan.a_view(A0, i))
        one.solve(A1, B1, use_graph=False)
        torch.cuda.synchronize()
        assert torch.equal(one.b_view(B1, 0), plan.b_view(B, i)), f"matrix {i} (n={n[i]}) differs alone"
        assert torch.equal(one.a_view(A1, 0), plan.a_view(A, i))


def test_f64_batch_matches_single_path_and_oracle(oracle):
    n, k = sizes()
    nb = 64
    plan, A, B = make(n, k, "f64", buckets=[(1 << 30, nb)])
    A0, B0 = A.clone(), B.clone()
    plan.solve(A, B, use_graph=False)
    info, _, _ = plan.results()
    assert (info == 0).all()
    eta = plan.backward_error(A0, B, B0)
    assert eta.max() <= 1e-14, eta.max()
    for i in range(0, len(n), 7):
        ni, ki = int(n[i]), int(k[i])
        Ai = colmajor_copy(plan.a_view(A0, i))  # ld = n
        Bi = colmajor_copy(plan.b_view(B0, i))
        d = torch.zeros(1, dtype=torch.int32, device="cuda")
        rc = lib().elemamd_hpd_solve_f64(b"L", b"N", ni, ki, ctypes.c_void_p(Ai.data_ptr()), ni,
                                         ctypes.c_void_p(Bi.data_ptr()), ni, nb, ctypes.c_void_p(d.data_ptr()), S())
        assert rc == 0 and lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), S()) == 0
        Lb = torch.tril(plan.a_view(A, i))
        # same potrf code; the single path defers trailing updates in larger
        # groups and solves its panels in sub-blocks, so the rounding differs
        assert rel(torch.tril(Ai), Lb) <= 1e-13, f"factor of matrix {i} vs the single path"
        assert rel(plan.b_view(B, i), Bi) <= 1e-12
        # oracle: the reference's HPDSolve on the host
        Ah = np.array(plan.a_view(A0, i).cpu().numpy(), order="F", copy=True)
        Bh = np.array(plan.b_view(B0, i).cpu().numpy(), order="F", copy=True)
        assert np.abs(Ah - Ah.T).max() <= 1e-14 * np.abs(Ah).max(), f"matrix {i} not symmetric on the host"
        assert np.linalg.eigvalsh(Ah).min() > 0, f"matrix {i} not HPD on the host"
        info_h = oracle.hpd_solve(Ah, Bh, nb=nb)
        assert info_h == 0, (i, ni, nb, info_h)
        assert np.abs(plan.b_view(B, i).cpu().numpy() - Bh).max() <= 1e-12 * np.abs(Bh).max()


@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_refinement_reaches_fp64_backward_error(precision):
    n, k = sizes(count=96, hi=1024)
    plan, A, B = make(n, k, precision)
    A0, B0 = A.clone(), B.clone()
    plan.set_refinement(6, 1e-14)
    plan.solve(A, B, use_graph=True)
    info, eta, it = plan.results()
    assert (info == 0).all()
    assert eta.max() <= 1e-14, (eta.max(), int(np.argmax(eta)))
    assert torch.equal(A, A0)  # the refinement modes read A only
    eta2 = plan.backward_error(A0, B, B0)
    assert eta2.max() <= 1e-14
    assert it.min() >= 1


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_graph_replays_equal_eager(precision):
    n, k = sizes(count=32)
    plan, A, B = make(n, k, precision)
    plan.set_refinement(5)
    A0, B0 = A.clone(), B.clone()
    plan.solve(A, B, use_graph=False)
    torch.cuda.synchronize()
    A_e, X_e = A.clone(), B.clone()
    for rep in range(3):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        plan.solve(A, B, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(B, X_e), f"replay {rep}"
        assert torch.equal(A, A_e), f"replay {rep}"


def test_non_hpd_member_reports_its_column():
    n = np.array([40, 200, 130, 17], dtype=np.int32)
    k = np.array([2, 3, 1, 4], dtype=np.int32)
    plan, A, B = make(n, k, "f64", buckets=[(1 << 30, 32)])
    bad, col = 1, 77
    plan.a_view(A, bad)[col, col] = -5.0
    ok_plan, A2, B2 = make(n, k, "f64", buckets=[(1 << 30, 32)])
    plan.solve(A, B, use_graph=False)
    ok_plan.solve(A2, B2, use_graph=False)
    info, _, _ = plan.results()
    assert info[bad] == col + 1
    for i in range(len(n)):
        if i != bad:
            assert info[i] == 0
            assert torch.equal(plan.b_view(B, i), ok_plan.b_view(B2, i))


def test_plan_rejects_bad_arguments():
    with pytest.raises(ValueError):
        EB.BatchPlan([4, 0], [1, 1])
    with pytest.raises(ValueError):
        EB.BatchPlan([4], [1], buckets=[(64, 0)])
    with pytest.raises(ValueError):
        EB.BatchPlan([4], [1], buckets=[(64, 32), (32, 32)])


# ------------------------------------------------------- persistent launch
PERSIST_BUCKETS = [(96, 32), (400, 64), (1 << 30, 128)]


def _grid(buckets, g):
    return [(m, nb, g) for m, nb in buckets]


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("grid", [7, 256])
def test_persistent_equals_stepwise(precision, grid):
    """Same matrices, same block sizes: the persistent launch (fewer
    workgroups than matrices, or more) against the stepwise launches."""
    n, k = sizes(count=64, hi=700)
    step, A1, B1 = make(n, k, precision, buckets=PERSIST_BUCKETS)
    pers, A2, B2 = make(n, k, precision, buckets=_grid(PERSIST_BUCKETS, grid))
    A0, B0 = A1.clone(), B1.clone()
    for p in (step, pers):
        p.set_refinement(6, 1e-14)
    step.solve(A1, B1, use_graph=False)
    pers.solve(A2, B2, use_graph=True)
    torch.cuda.synchronize()
    i1, e1, it1 = step.results()
    i2, e2, it2 = pers.results()
    assert (i1 == 0).all() and (i2 == 0).all()
    assert torch.equal(A1, A2)
    assert torch.equal(B1, B2)
    assert np.array_equal(e1, e2, equal_nan=True) and np.array_equal(it1, it2)
    # plain FP64 overwrote A with L: the backward error is taken against the input
    assert pers.backward_error(A0, B2, B0).max() <= 1e-14


def test_persistent_graph_replays_and_non_hpd_member():
    n = np.array([40, 200, 130, 17, 600], dtype=np.int32)
    k = np.array([2, 3, 1, 4, 8], dtype=np.int32)
    b = [(1 << 30, 32, 3)]
    plan, A, B = make(n, k, "f64", buckets=b)
    bad, col = 1, 77
    plan.a_view(A, bad)[col, col] = -5.0
    A0, B0 = A.clone(), B.clone()
    ok_plan, A2, B2 = make(n, k, "f64", buckets=b)
    ok_plan.solve(A2, B2, use_graph=False)
    plan.solve(A, B, use_graph=False)
    torch.cuda.synchronize()
    A_e, X_e = A.clone(), B.clone()
    info, _, _ = plan.results()
    assert info[bad] == col + 1
    for i in range(len(n)):
        if i != bad:
            assert info[i] == 0
            assert torch.equal(plan.b_view(B, i), ok_plan.b_view(B2, i))
    for rep in range(3):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        plan.solve(A, B, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(B, X_e), f"replay {rep}"
        assert torch.equal(A, A_e), f"replay {rep}"
        assert plan.results()[0][bad] == col + 1


def test_persistent_rejects_bad_grid():
    with pytest.raises(ValueError):
        EB.BatchPlan([100, 300], [1, 2], buckets=[(1 << 30, 128, -1)])
