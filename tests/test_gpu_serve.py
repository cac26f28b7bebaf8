"""Serving loop through the C ABI.

Bars: a request's result is bit-for-bit the same alone (one slot, nothing
else running) and in a concurrent mix with reuse of resident factors (each
request runs on one stream with deterministic kernels); every result meets
FP64 normwise backward error <= 1e-14; the reserved HBM never exceeds the
budget; reuse requests skip the factorization; a non-HPD request reports its
column without disturbing the others.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from elemental_amd import inputs  # noqa: E402
from elemental_amd.serve import Server, result_dict  # noqa: E402

SIZES = [64, 300, 1000, 129, 700, 1500, 256, 90]


def problem(n, seed, k=3):
    A = inputs.spectrum_device(n, 1e4, seed).tensor()
    B = inputs.normal_device(n, k, seed + 1000).tensor()
    return A, B


def eta(A, X, B):
    R = B - A @ X
    an = A.abs().sum(dim=1).max()
    return float((R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max())


def workload():
    probs = [problem(n, 40 + i) for i, n in enumerate(SIZES)]
    # (key, reuse, problem index, rhs seed): every other matrix is asked again with new right-hand sides
    reqs = [(i, False, i, None) for i in range(len(SIZES))]
    reqs += [(i, True, i, 900 + i) for i in range(0, len(SIZES), 2)]
    return probs, reqs


def rhs(probs, i, seed):
    n, k = probs[i][1].shape
    return probs[i][1].clone() if seed is None else inputs.normal_device(n, k, seed).tensor().clone()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_alone_equals_in_mix(precision):
    probs, reqs = workload()
    torch.cuda.synchronize()
    # alone: one slot, one request at a time
    alone = []
    with Server(slots=1, precision=precision) as S:
        for key, reuse, i, seed in reqs:
            X = rhs(probs, i, seed)
            r = S.wait(S.submit(key, probs[i][0], X, reuse=reuse))
            assert r.status == 0, (key, r.status)
            alone.append(X)
    # mix: four slots, everything submitted at once
    with Server(slots=4, precision=precision) as S:
        outs, tickets = [], []
        for key, reuse, i, seed in reqs:
            X = rhs(probs, i, seed)
            outs.append(X)
            tickets.append(S.submit(key, probs[i][0], X, reuse=reuse))
        res = [S.wait(t) for t in tickets]
        st = S.stats()
    assert all(r.status == 0 for r in res)
    assert st.completed == len(reqs)
    assert len({r.slot for r in res}) > 1  # the requests really were spread over slots
    for (key, reuse, i, seed), a, b in zip(reqs, alone, outs):
        assert torch.equal(a, b), f"request key={key} reuse={reuse} differs in the mix"
        Bi = rhs(probs, i, seed)
        assert eta(probs[i][0], b, Bi) <= 1e-14


def test_reuse_skips_factorization():
    A, B = problem(800, 7)
    with Server(slots=2, precision="f64") as S:
        X1 = B.clone()
        r1 = S.wait(S.submit(5, A, X1))
        B2 = inputs.normal_device(800, 3, 77).tensor()
        X2 = B2.clone()
        r2 = S.wait(S.submit(5, A, X2, reuse=True))
        st = S.stats()
    assert r1.status == 0 and r2.status == 0, (result_dict(r1), result_dict(r2))
    assert r1.factored == 1 and r2.factored == 0, (result_dict(r1), result_dict(r2))
    assert st.reuse_hits == 1 and st.resident_factors == 1
    assert eta(A, X2, B2) <= 1e-14
    assert r2.t_done - r2.t_start < r1.t_done - r1.t_start


def test_admission_respects_memory_budget():
    probs = [problem(n, 60 + i, k=2) for i, n in enumerate([1200, 1100, 1000, 900, 1300, 1250])]
    # room for about two factors at a time: the rest must wait, and old factors get evicted
    budget = 2 * 1300 * 1300 * 8 + (8 << 20)
    with Server(slots=4, precision="f64", memory_budget=budget) as S:
        outs, tickets = [], []
        for i, (A, B) in enumerate(probs):
            X = B.clone()
            outs.append(X)
            tickets.append(S.submit(i, A, X))
        res = [S.wait(t) for t in tickets]
        st = S.stats()
    assert all(r.status == 0 for r in res)
    assert st.peak_reserved <= budget
    assert st.evictions >= 1
    for (A, B), X in zip(probs, outs):
        assert eta(A, X, B) <= 1e-14


def test_non_hpd_request_reports_column():
    A, B = problem(300, 11)
    Abad = A.clone()
    Abad[150, 150] = -1.0
    with Server(slots=2, precision="f64") as S:
        Xbad, Xok = B.clone(), B.clone()
        tb = S.submit(1, Abad, Xbad)
        to = S.submit(2, A, Xok)
        rb, ro = S.wait(tb), S.wait(to)
    assert rb.status == -5 and 1 <= rb.info <= 151
    assert ro.status == 0 and eta(A, Xok, B) <= 1e-14


def test_submit_rejects_bad_requests():
    A, B = problem(64, 3)
    with Server(slots=1, precision="f64", memory_budget=1 << 16) as S:
        with pytest.raises(RuntimeError):
            S.submit(0, A, B.clone())  # can never fit in the budget
    with Server(slots=1) as S:
        with pytest.raises(ValueError):
            S.submit(0, A, B.t())  # not column-major
