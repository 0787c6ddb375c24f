"""Batched sigma builds (tci_sigma_apply_batch) and the batched Davidson steps.

In split precision with per-walk ladder programs the ladder of all vectors runs
as one pass; its K-splits may differ from the single-vector schedule, so the
batch agrees with single builds to FP32-accumulation rounding (and with the
oracle to the split-precision bound).  Everything else is per vector, and FP64
or full-domain (merged) plans build the vectors one by one: identical results.
"""
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu
SPLIT_TOL = {"bf16x3": 5e-5, "fp16x3": 2e-6}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _single(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


def _batch(plan, xs, **kw):
    c = tci.DeviceArray.from_numpy(np.concatenate(xs))
    out = tci.DeviceArray(len(xs) * plan.n_csf)
    plan.sigma_batch(c, out, len(xs), **kw)
    return out.numpy().reshape(len(xs), plan.n_csf)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("args", [
    dict(n_int=5, n_v=60, n_chol=40, seed=81, domain_range=(10, 60), weak_below=12),
    dict(n_int=3, n_v=200, n_chol=24, seed=53, domain_range=(100, 192), weak_below=0),
])
def test_batch_matches_single(args, precision):
    _say("case: batched sigma", precision, args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    xs = [random_ci_vector(prob, 90 + k) for k in range(3)]
    got = _batch(plan, xs, precision=precision)
    for k, x in enumerate(xs):
        one = _single(plan, x, precision=precision)
        ref = oracle_sigma(prob, x, ints)
        _say(f"  vector {k}: batch vs single {_rel(got[k], one):.2e}  vs oracle {_rel(got[k], ref):.2e}")
        assert _rel(got[k], ref) < SPLIT_TOL[precision]
        assert _rel(got[k], one) < SPLIT_TOL[precision] / 10
    again = _batch(plan, xs, precision=precision)
    assert np.array_equal(again, got)
    # accumulate and E_ref flags
    base = np.ones((3, plan.n_csf))
    c = tci.DeviceArray.from_numpy(np.concatenate(xs))
    out = tci.DeviceArray.from_numpy(base.ravel())
    plan.sigma_batch(c, out, 3, precision=precision, accumulate=True, add_eref=True)
    acc = out.numpy().reshape(3, plan.n_csf)
    for k in range(3):
        assert np.allclose(acc[k], 1.0 + got[k] + plan.info.e_ref * xs[k], rtol=0,
                           atol=1e-12 * max(1.0, np.abs(got[k]).max(), abs(plan.info.e_ref)))


def test_batch_falls_back_to_single_builds():
    _say("case: batched sigma fallbacks (fp64, full domains)")
    prob = make_problem(4, 30, 30, 82)  # full domains: merged ladder
    plan = tci.Plan.from_problem(prob)
    xs = [random_ci_vector(prob, 95 + k) for k in range(2)]
    for precision in ["fp64", "bf16x3"]:
        got = _batch(plan, xs, precision=precision)
        for k, x in enumerate(xs):
            assert np.array_equal(got[k], _single(plan, x, precision=precision))


def test_batched_davidson_matches_unbatched():
    _say("case: batched davidson steps")
    prob = make_problem(6, 90, 120, 12, domain_range=(30, 90), weak_below=35, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    kw = dict(roots=3, energy_tol=1e-9, norm_tol=1e-6, precision="bf16x3", switch_tol=1e-6)
    e0, r0 = plan.davidson(**kw)
    os.environ["TCI_DAVIDSON_BATCH"] = "1"
    try:
        e1, r1 = plan.davidson(**kw)
    finally:
        del os.environ["TCI_DAVIDSON_BATCH"]
    e64, _ = plan.davidson(roots=3, energy_tol=1e-9, norm_tol=1e-6)
    _say(f"  batched {e1} ({r1['iterations']} it)  unbatched {e0} ({r0['iterations']} it)  fp64 {e64}")
    assert r1["converged"] and r0["converged"]
    assert np.abs(e1 - e64).max() < 1e-6 and np.abs(e0 - e64).max() < 1e-6
