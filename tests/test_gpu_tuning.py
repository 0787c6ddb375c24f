"""Autotuned launch schedules.

A plan times K-split fill targets and workgroup orders of its ladder programs
on the device (FP64 at creation, split precisions on first use) and keeps the
fastest.  The schedule changes the order of FP64 partial sums only, so tuned
and untuned plans agree to rounding, and both reproduce the oracle.
"""
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sigma(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("args", [
    dict(n_int=5, n_v=60, n_chol=40, seed=51, domain_range=(10, 60), weak_below=12),
    dict(n_int=4, n_v=40, n_chol=30, seed=52),
])
def test_tuned_schedule_matches_default(args):
    _say("case: tuned vs default schedule", args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 9)
    ref = oracle_sigma(prob, x, ints)
    tuned = tci.Plan.from_problem(prob)
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        fixed = tci.Plan.from_problem(prob)
        f64, fmx = _sigma(fixed, x), _sigma(fixed, x, precision="bf16x3")
    finally:
        del os.environ["TCI_AUTOTUNE"]
    t64, tmx = _sigma(tuned, x), _sigma(tuned, x, precision="bf16x3")
    log = tuned.tuning()
    _say(log)
    assert "fp64 ladder" in log and "split ladder" in log
    assert fixed.tuning() == ""
    assert _rel(t64, ref) < 1e-10 and _rel(f64, ref) < 1e-10
    assert _rel(t64, f64) < 1e-12
    assert _rel(tmx, ref) < 5e-5 and _rel(fmx, ref) < 5e-5
    # the schedule is kept: repeated and re-prepared sigmas are identical
    assert np.array_equal(_sigma(tuned, x, precision="bf16x3"), tmx)
    _sigma(tuned, x, precision="fp16x3")
    assert np.array_equal(_sigma(tuned, x, precision="bf16x3"), tmx)
    assert tuned.tuning().count("split ladder") == 1


def test_phase2_192_tiles():
    """Phase 2 of the split ladder on single 192 x 192 tiles for domains of
    129-192 virtuals (the shape the tuner may pick) reproduces the oracle and
    the 128-row tiles to rounding."""
    _say("case: split ladder phase 2 on 192 x 192 tiles")
    prob = make_problem(3, 200, 24, 53, domain_range=(100, 192), weak_below=0)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 10)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision="bf16x3")
        os.environ["TCI_SPLIT_P2_192"] = "1"
        big = tci.Plan.from_problem(prob)
        b = _sigma(big, x, precision="bf16x3")
        c = _sigma(big, x, precision="fp16x3")
    finally:
        os.environ.pop("TCI_SPLIT_P2_192", None)
        del os.environ["TCI_AUTOTUNE"]
    d = np.diff(np.asarray(prob.dom_off))
    assert ((d > 128) & (d <= 192)).sum() >= 2
    _say(f"  rel err 128-tiles {_rel(a, ref):.2e}  192-tiles {_rel(b, ref):.2e}  fp16x3 {_rel(c, ref):.2e}")
    assert _rel(a, ref) < 5e-5 and _rel(b, ref) < 5e-5 and _rel(c, ref) < 2e-6
    assert _rel(b, a) < 1e-5


@pytest.mark.parametrize("env", ["TCI_FP64_P192", "TCI_FP64_P1_192"])
def test_fp64_192_tiles(env):
    """Both FP64 ladder phases (TCI_FP64_P192), or phase 1 alone
    (TCI_FP64_P1_192), of domains of 129-192 on 192 x 192 tiles (eight waves
    per workgroup) reproduce the oracle and the default tiles."""
    _say("case: FP64 ladder on 192 x 192 tiles", env)
    prob = make_problem(3, 200, 24, 53, domain_range=(100, 192), weak_below=0)
    ints = Integrals(prob)
    xs = [random_ci_vector(prob, 11 + k) for k in range(2)]
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        base = tci.Plan.from_problem(prob)
        os.environ[env] = "1"
        big = tci.Plan.from_problem(prob)
    finally:
        os.environ.pop(env, None)
        del os.environ["TCI_AUTOTUNE"]
    for x in xs:
        ref = oracle_sigma(prob, x, ints)
        a, b = _sigma(base, x), _sigma(big, x)
        _say(f"  rel err default {_rel(a, ref):.2e}  192-tiles {_rel(b, ref):.2e}")
        assert _rel(a, ref) < 1e-10 and _rel(b, ref) < 1e-10
        assert _rel(b, a) < 1e-12


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_fused_split_ladder(precision):
    """The fused split ladder (T' kept on chip) for domains of 129-192
    reproduces the oracle and the two-phase ladder to the split precision."""
    _say("case: fused split ladder", precision)
    prob = make_problem(3, 200, 24, 53, domain_range=(100, 192), weak_below=0)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 12)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
        os.environ["TCI_SPLIT_FUSED"] = "1"
        fused = tci.Plan.from_problem(prob)
        b = _sigma(fused, x, precision=precision)
        b2 = _sigma(fused, x, precision=precision)
    finally:
        os.environ.pop("TCI_SPLIT_FUSED", None)
        del os.environ["TCI_AUTOTUNE"]
    tol = 5e-5 if precision == "bf16x3" else 2e-6
    _say(f"  rel err two-phase {_rel(a, ref):.2e}  fused {_rel(b, ref):.2e}  fused vs two-phase {_rel(b, a):.2e}")
    assert _rel(a, ref) < tol and _rel(b, ref) < tol
    assert _rel(b, a) < tol / 5
    assert np.array_equal(b, b2)


@pytest.mark.parametrize("precision,dr,n_chol,waves", [
    ("bf16x3", (40, 140), 24, "1"),
    ("fp16x3", (40, 140), 24, "1"),
    ("bf16x3", (65, 65), 7, "1"),
    ("bf16x3", (128, 128), 40, "1"),
    ("bf16x3", (40, 140), 24, "8"),
    ("fp16x3", (65, 128), 9, "8"),
])
def test_fused128_split_ladder(precision, dr, n_chol, waves):
    """The whole-walk fused split ladder for domains of 65-128 (T' kept in
    registers) reproduces the oracle and the two-phase ladder to the split
    precision, next to two-phase walks of other sizes; J ranges of one and of
    several Cholesky vectors."""
    _say("case: fused128 split ladder", precision, dr, n_chol, "eight waves" if waves == "8" else "four waves")
    prob = make_problem(5, 200, n_chol, 61, domain_range=dr, weak_below=0)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 13)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
        os.environ["TCI_SPLIT_FUSED128"] = waves
        fused = tci.Plan.from_problem(prob)
        b = _sigma(fused, x, precision=precision)
        b2 = _sigma(fused, x, precision=precision)
        os.environ["TCI_SPLIT_FUSED"] = "1"  # both fused kernels and the two-phase walks below 65
        c = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
    finally:
        os.environ.pop("TCI_SPLIT_FUSED", None)
        os.environ.pop("TCI_SPLIT_FUSED128", None)
        del os.environ["TCI_AUTOTUNE"]
    tol = 5e-5 if precision == "bf16x3" else 2e-6
    _say(f"  rel err two-phase {_rel(a, ref):.2e}  fused128 {_rel(b, ref):.2e}  both fused {_rel(c, ref):.2e}"
         f"  fused128 vs two-phase {_rel(b, a):.2e}")
    assert _rel(a, ref) < tol and _rel(b, ref) < tol and _rel(c, ref) < tol
    assert _rel(b, a) < tol / 5
    assert np.array_equal(b, b2)


@pytest.mark.parametrize("precision,dr,n_chol", [
    ("bf16x3", (100, 192), 24),
    ("fp16x3", (100, 192), 24),
    ("bf16x3", (129, 129), 7),
    ("bf16x3", (192, 192), 40),
])
def test_fused192_split_ladder(precision, dr, n_chol):
    """The whole-walk fused split ladder for domains of 129-192 (C in LDS, L^J
    from global memory, T' in registers) reproduces the oracle and the
    two-phase ladder to the split precision, alone and next to the fused
    kernel for 65-128; J ranges of one and of several Cholesky vectors."""
    _say("case: fused192 split ladder", precision, dr, n_chol)
    prob = make_problem(5, 200, n_chol, 62, domain_range=dr, weak_below=0)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 14)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_AUTOTUNE"] = "0"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
        os.environ["TCI_SPLIT_FUSED192"] = "1"
        fused = tci.Plan.from_problem(prob)
        b = _sigma(fused, x, precision=precision)
        b2 = _sigma(fused, x, precision=precision)
        os.environ["TCI_SPLIT_FUSED128"] = "1"
        c = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
    finally:
        os.environ.pop("TCI_SPLIT_FUSED192", None)
        os.environ.pop("TCI_SPLIT_FUSED128", None)
        del os.environ["TCI_AUTOTUNE"]
    tol = 5e-5 if precision == "bf16x3" else 2e-6
    _say(f"  rel err two-phase {_rel(a, ref):.2e}  fused192 {_rel(b, ref):.2e}  both fused {_rel(c, ref):.2e}"
         f"  fused192 vs two-phase {_rel(b, a):.2e}")
    assert _rel(a, ref) < tol and _rel(b, ref) < tol and _rel(c, ref) < tol
    assert _rel(b, a) < tol / 5
    assert np.array_equal(b, b2)


@pytest.mark.parametrize("precision,env", [
    ("bf16x3", {"TCI_SPLIT_DMA": "1"}),
    ("fp16x3", {"TCI_SPLIT_DMA": "1"}),
    ("bf16x3", {"TCI_SPLIT_DMA1": "2"}),
    ("bf16x3", {"TCI_SPLIT_DMA1": "3"}),
    ("fp16x3", {"TCI_SPLIT_DMA": "1", "TCI_SPLIT_DMA1": "2"}),
])
def test_split_direct_load_stages(precision, env):
    """The domains of 129-192 with their phase-2 192 x 192 tiles
    (TCI_SPLIT_DMA) or phase-1 128 x 192 tiles (TCI_SPLIT_DMA1: 2 or 3 stages)
    fed by direct global-to-LDS loads give the register-staged result up to
    FP32 summation order and the oracle to the split precision; domains below
    and at 192."""
    _say("case: split tiles with direct-load stages", precision, env)
    prob = make_problem(4, 200, 30, 63, domain_range=(129, 192), weak_below=0)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 15)
    ref = oracle_sigma(prob, x, ints)
    os.environ["TCI_AUTOTUNE"] = "0"
    os.environ["TCI_SPLIT_P2_192"] = "1"
    try:
        a = _sigma(tci.Plan.from_problem(prob), x, precision=precision)
        os.environ.update(env)
        plan = tci.Plan.from_problem(prob)
        b = _sigma(plan, x, precision=precision)
        b2 = _sigma(plan, x, precision=precision)
    finally:
        for k in env:
            os.environ.pop(k, None)
        os.environ.pop("TCI_SPLIT_P2_192", None)
        del os.environ["TCI_AUTOTUNE"]
    tol = 5e-5 if precision == "bf16x3" else 2e-6
    _say(f"  rel err register-staged {_rel(a, ref):.2e}  direct loads {_rel(b, ref):.2e}  vs each other {_rel(b, a):.2e}")
    assert _rel(a, ref) < tol and _rel(b, ref) < tol
    assert _rel(b, a) < tol / 5
    assert np.array_equal(b, b2)
