"""Split-precision (bf16x3) sigma and mixed Davidson on the GPU.

The bf16x3 products keep ~17 significant bits per operand (hi + lo bf16) and
accumulate in FP32, so sigma agrees with the oracle to a few 1e-6 relative
(bound used: 5e-5 in the 2-norm); scaled fp16x3 keeps ~22 bits (bound 2e-6).  Split precision covers the ladder and,
where the plan runs them as dense products, the ring, 3-external and
4-internal terms; the remaining terms stay FP64.  The mixed
Davidson refines in FP64 after the switch, so its roots must match the FP64
solver's to 1 microhartree (they typically agree to 1e-9).
"""
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu
TOL_MIXED = 5e-5


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


SPLIT_TOL = {"bf16x3": TOL_MIXED, "fp16x3": 2e-6}


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("args", [
    dict(n_int=2, n_v=5, n_chol=7, seed=1),
    dict(n_int=4, n_v=37, n_chol=45, seed=2),
    dict(n_int=6, n_v=70, n_chol=50, seed=3, domain_range=(9, 70), weak_below=12),
    dict(n_int=3, n_v=300, n_chol=30, seed=4, domain_range=(130, 300)),
    dict(n_int=4, n_v=150, n_chol=300, seed=5),
])
def test_sigma_split(args, precision):
    _say("case:", precision, "sigma", args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    for k in range(2):
        x = random_ci_vector(prob, 70 + k)
        ref = oracle_sigma(prob, x, ints)
        mixed = _sigma(plan, x, precision=precision)
        fp64 = _sigma(plan, x)
        err = _rel(mixed, ref)
        _say(f"  {precision} rel err {err:.3e}  fp64 rel err {_rel(fp64, ref):.3e}")
        assert err < SPLIT_TOL[precision]
        assert _rel(fp64, ref) < 1e-10
    # a vector without doubles (reference + singles) exercises the split
    # 3-external products alone
    lay = prob.layout()
    x = np.zeros(lay.n_csf)
    x[: 1 + prob.n_int * prob.n_v] = random_ci_vector(prob, 90)[: 1 + prob.n_int * prob.n_v]
    assert _rel(_sigma(plan, x, precision=precision), oracle_sigma(prob, x, ints)) < SPLIT_TOL[precision]


def test_sigma_bf16x3_flags_and_repeat():
    _say("case: bf16x3 flags / repeat")
    prob = make_problem(3, 24, 20, 6, domain_range=(5, 24), weak_below=6)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 3)
    a = _sigma(plan, x, precision="bf16x3")
    b = _sigma(plan, x, precision="bf16x3")
    assert np.array_equal(a, b)  # deterministic: no atomics in the mixed ladder
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray.from_numpy(np.ones(plan.n_csf))
    plan.sigma(c, out, precision="bf16x3", accumulate=True, add_eref=True)
    assert np.allclose(out.numpy(), 1.0 + a + plan.info.e_ref * x, rtol=0, atol=1e-12 * max(1, np.abs(a).max()))


def test_switching_split_precisions():
    """One plan serves both split precisions (the split copies are rebuilt on a switch)."""
    _say("case: bf16x3 -> fp16x3 -> bf16x3 on one plan")
    prob = make_problem(3, 30, 25, 8, domain_range=(10, 30), weak_below=12)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 4)
    ref = oracle_sigma(prob, x, ints)
    a = _sigma(plan, x, precision="bf16x3")
    b = _sigma(plan, x, precision="fp16x3")
    c = _sigma(plan, x, precision="bf16x3")
    assert np.array_equal(a, c)
    assert _rel(b, ref) < SPLIT_TOL["fp16x3"] and _rel(a, ref) < SPLIT_TOL["bf16x3"]
    assert _rel(b, ref) < _rel(a, ref)


@pytest.mark.parametrize("args,roots", [
    (dict(n_int=5, n_v=60, n_chol=80, seed=11, l_scale=0.03), 1),
    (dict(n_int=6, n_v=90, n_chol=120, seed=12, domain_range=(30, 90), weak_below=35, l_scale=0.03), 3),
])
def test_mixed_davidson_matches_fp64(args, roots):
    _say("case: mixed davidson", args, roots)
    prob = make_problem(**args)
    plan = tci.Plan.from_problem(prob)
    e64, r64 = plan.davidson(roots=roots, energy_tol=1e-9, norm_tol=1e-6)
    emx, rmx = plan.davidson(roots=roots, energy_tol=1e-9, norm_tol=1e-6, precision="bf16x3", switch_tol=1e-6)
    _say(f"  fp64 {e64} ({r64['iterations']} it)  mixed {emx} ({rmx['iterations']} it, "
         f"{rmx['sigma_mixed']} mixed + {rmx['sigma_fp64']} fp64 sigma)")
    assert r64["converged"] and rmx["converged"]
    assert rmx["sigma_mixed"] > 0
    assert np.abs(emx - e64).max() < 1e-6


def test_per_pair_launch_matches_grouped():
    """One launch per walk and phase (TCI_LADDER_PER_PAIR=1) computes the same
    tiles in the same order as the grouped launches: identical sigma."""
    _say("case: per-walk ladder launches vs grouped")
    prob = make_problem(5, 60, 30, 9, domain_range=(10, 60), weak_below=12)
    x = random_ci_vector(prob, 5)
    # the split ladder is laid out on a plan's first split sigma, which is
    # when the launch mode is read; both plans keep the untuned K-splits
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision="bf16x3")
        os.environ["TCI_LADDER_PER_PAIR"] = "1"
        b = _sigma(tci.Plan.from_problem(prob), x, precision="bf16x3")
    finally:
        os.environ.pop("TCI_LADDER_PER_PAIR", None)
        del os.environ["TCI_AUTOTUNE"]
    assert np.array_equal(a, b)
