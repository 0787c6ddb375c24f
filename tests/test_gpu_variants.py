"""GPU parity of product paths that the other sigma tests do not reach:
FP32 storage of L^J for the ladder, pair domains of 301-400 virtuals, and the
fused single-kernel FP64 ladder (TCI_LADDER=fused).  FP64 results must match
the oracle to 1e-10 relative 2-norm."""
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import Problem, make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu
TOL = 1e-10
SPLIT_TOL = {"bf16x3": 5e-5, "fp16x3": 2e-6}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _gpu_sigma(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


def _fp32_records(prob):
    """The same problem with its Cholesky records rounded to FP32, so FP32
    storage loses nothing and FP64 parity stays exact."""
    L32 = prob.L.astype(np.float32)
    p64 = Problem(prob.n_int, prob.n_v, prob.n_chol, prob.eps, L32.astype(np.float64), prob.pairs,
                  prob.dom_off, prob.dom_idx)
    p32 = Problem(prob.n_int, prob.n_v, prob.n_chol, prob.eps, L32, prob.pairs, prob.dom_off, prob.dom_idx)
    return p64, p32


@pytest.mark.parametrize("args", [
    dict(n_int=4, n_v=37, n_chol=45, seed=11),
    dict(n_int=6, n_v=70, n_chol=50, seed=12, domain_range=(9, 70), weak_below=12),
    dict(n_int=3, n_v=200, n_chol=24, seed=13, domain_range=(130, 200)),
])
def test_fp32_ladder_storage(args):
    """FP32 records (chol dtype) and FP32 device storage of L^J_ab for the
    ladder (ladder_dtype f32), as at 40/1000/3000: FP64 sigma equals the
    oracle on the FP32-rounded records, split precisions within their bounds."""
    _say("case: fp32 L^J storage", args)
    p64, p32 = _fp32_records(make_problem(**args))
    ints = Integrals(p64)
    plan = tci.Plan.from_problem(p32, ladder_dtype="f32")
    for k in range(2):
        x = random_ci_vector(p64, 20 + k)
        ref = oracle_sigma(p64, x, ints)
        got = _gpu_sigma(plan, x)
        _say(f"  fp64 rel err {_rel(got, ref):.2e}")
        assert _rel(got, ref) < TOL
        for prec in ("bf16x3", "fp16x3"):
            err = _rel(_gpu_sigma(plan, x, precision=prec), ref)
            _say(f"  {prec} rel err {err:.2e}")
            assert err < SPLIT_TOL[prec]


def test_domains_301_to_400():
    """Pair domains of 301-400 virtuals (the upper range of the 40/1000/3000
    workload), with weak pairs, in FP64 and both split precisions."""
    args = dict(n_int=3, n_v=420, n_chol=16, seed=14, domain_range=(290, 400), weak_below=301)
    _say("case: domains 301-400", args)
    prob = make_problem(**args)
    d = np.diff(prob.dom_off)
    _say(f"  walks {prob.n_pairs}, domain sizes {sorted(d.tolist())}")
    assert prob.n_pairs >= 3 and d.min() >= 301 and d.max() <= 400
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 30)
    ref = oracle_sigma(prob, x, ints)
    got = _gpu_sigma(plan, x)
    _say(f"  fp64 rel err {_rel(got, ref):.2e}")
    assert _rel(got, ref) < TOL
    for prec in ("bf16x3", "fp16x3"):
        err = _rel(_gpu_sigma(plan, x, precision=prec), ref)
        _say(f"  {prec} rel err {err:.2e}")
        assert err < SPLIT_TOL[prec]


@pytest.mark.parametrize("args", [
    dict(n_int=4, n_v=37, n_chol=45, seed=15),
    dict(n_int=5, n_v=90, n_chol=30, seed=16, domain_range=(10, 90), weak_below=14),
])
def test_fused_fp64_ladder(args):
    """The fused single-kernel FP64 ladder (TCI_LADDER=fused) reproduces the
    oracle and the default two-phase ladder."""
    _say("case: fused FP64 ladder", args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 40)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_LADDER"] = "fused"
    try:
        fused = tci.Plan.from_problem(prob)
    finally:
        del os.environ["TCI_LADDER"]
    a = _gpu_sigma(fused, x)
    b = _gpu_sigma(tci.Plan.from_problem(prob), x)
    _say(f"  fused rel err {_rel(a, ref):.2e}  two-phase {_rel(b, ref):.2e}")
    assert _rel(a, ref) < TOL and _rel(b, ref) < TOL
