"""GPU parity of the ladder from the resident virtual block (ab|cd)
(ladder_vv.hip) against the oracle, and against the Cholesky-direct ladder of
the same plan inputs.  TCI_LADDER_VV=1 forces the resident ladder whenever it
fits (with or without autotuning); TCI_LADDER_VV=0 switches it off."""
import contextlib
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu
TOL = 1e-10
SPLIT_TOL = {"bf16x3": 5e-5, "fp16x3": 2e-6}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _sigma(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


@contextlib.contextmanager
def _env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    os.environ.update({k: str(v) for k, v in kv.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = [
    dict(n_int=2, n_v=5, n_chol=7, seed=1),
    dict(n_int=4, n_v=37, n_chol=45, seed=2),
    dict(n_int=6, n_v=70, n_chol=50, seed=3, domain_range=(9, 70), weak_below=12),
    dict(n_int=5, n_v=64, n_chol=80, seed=6, domain_range=(20, 64), h_offdiag=0.2),
    dict(n_int=4, n_v=150, n_chol=300, seed=5),
]


@pytest.mark.parametrize("autotune", ["0", "1"])
@pytest.mark.parametrize("args", CASES)
def test_vv_sigma_fp64(args, autotune):
    _say("case: resident ladder fp64", args, "autotune", autotune)
    prob = make_problem(**args)
    ints = Integrals(prob)
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=autotune):
        plan = tci.Plan.from_problem(prob)
        assert "fp64 ladder (ab|cd) resident" in plan.tuning() and "-> resident" in plan.tuning()
        for k in range(2):
            x = random_ci_vector(prob, 80 + k)
            assert _rel(_sigma(plan, x), oracle_sigma(prob, x, ints)) < TOL


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("args", CASES)
def test_vv_sigma_split(args, precision):
    _say("case: resident ladder", precision, args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=0):
        plan = tci.Plan.from_problem(prob)
        for k in range(2):
            x = random_ci_vector(prob, 90 + k)
            got = _sigma(plan, x, precision=precision)
            assert _rel(got, oracle_sigma(prob, x, ints)) < SPLIT_TOL[precision]
        assert precision + " ladder (ab|cd) resident" in plan.tuning()


def test_vv_matches_cholesky_direct_and_graph_replays():
    """Same inputs, resident and Cholesky-direct ladders: FP64 sigma to 1e-12;
    the resident FP64 and bf16x3 builds replayed from HIP graphs match their
    eager builds over several replays with eager work between them."""
    args = dict(n_int=5, n_v=48, n_chol=70, seed=7, domain_range=(12, 48), weak_below=14)
    _say("case: resident vs Cholesky-direct", args)
    prob = make_problem(**args)
    with _env(TCI_LADDER_VV=0, TCI_AUTOTUNE=0):
        p0 = tci.Plan.from_problem(prob)
        assert "resident" not in p0.tuning()
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=0):
        p1 = tci.Plan.from_problem(prob)
        xs = [random_ci_vector(prob, 100 + k) for k in range(3)]
        for x in xs:
            assert _rel(_sigma(p1, x), _sigma(p0, x)) < 1e-12
        for prec in ("fp64", "bf16x3"):
            eager = [_sigma(p1, x, precision=prec) for x in xs]
            for rep in range(3):
                for x, e in zip(xs, eager):
                    g = _sigma(p1, x, precision=prec, graph=True)
                    assert np.array_equal(g, e), (prec, rep)
                _sigma(p1, xs[0], precision=prec)  # eager work between replays


def test_vv_davidson_roots():
    """Davidson roots with the resident ladder (FP64 and mixed) equal those of
    the Cholesky-direct ladder."""
    args = dict(n_int=5, n_v=30, n_chol=40, seed=3)
    _say("case: resident ladder Davidson", args)
    prob = make_problem(**args)
    with _env(TCI_LADDER_VV=0, TCI_AUTOTUNE=0):
        e0, _ = tci.Plan.from_problem(prob).davidson(roots=2, energy_tol=1e-11, norm_tol=1e-8)
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=0):
        plan = tci.Plan.from_problem(prob)
        e1, _ = plan.davidson(roots=2, energy_tol=1e-11, norm_tol=1e-8)
        e2, r2 = plan.davidson(roots=2, energy_tol=1e-11, norm_tol=1e-8, precision="bf16x3")
    _say("fp64", e0, "resident", e1, "resident mixed", e2, r2)
    assert r2["converged"]
    assert np.max(np.abs(np.asarray(e1) - np.asarray(e0))) < 1e-9
    assert np.max(np.abs(np.asarray(e2) - np.asarray(e0))) < 1e-6


@pytest.mark.parametrize("tile", ["128x192T", "128x192T2", "128x192d2", "128x192d3", "128x128d2"])
@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("args", CASES)
def test_vv_transposed_split_product(args, precision, tile):
    """S^T = V X^T (walks along the tiles' N side), one and two K-splits,
    transposed back before the gather, and the untransposed product with two
    or three direct-load stages: against the oracle at the split bound, and
    against the default schedule to FP32 summation order."""
    _say("case: resident ladder transposed", precision, tile, args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    xs = [random_ci_vector(prob, 110 + k) for k in range(2)]
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=0):
        ref = [_sigma(tci.Plan.from_problem(prob), x, precision=precision) for x in xs[:1]]
    with _env(TCI_LADDER_VV=1, TCI_AUTOTUNE=0, TCI_VV_TILE=tile):
        plan = tci.Plan.from_problem(prob)
        got = [_sigma(plan, x, precision=precision) for x in xs]
        assert _rel(got[0], ref[0]) < 1e-6
        for x, g in zip(xs, got):
            assert _rel(g, oracle_sigma(prob, x, ints)) < SPLIT_TOL[precision]
        assert np.array_equal(_sigma(plan, xs[0], precision=precision, graph=True), got[0])
