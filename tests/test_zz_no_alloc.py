"""tci_htimesc_noalloc (host vectors staged through the library's fixed device
array) against tci_htimesc (per-call device buffers), called through the C ABI:
two calls in a row on different plans, the second with a shorter vector, each
checked against tci_htimesc and against the oracle; and the size limit is
rejected with a clear error before anything is copied."""
import ctypes
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma

pytestmark = pytest.mark.gpu
MAX_CSF = 1 << 22


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _noalloc(lib, plan, x, base):
    out = np.ascontiguousarray(base.copy())
    x = np.ascontiguousarray(x)
    rc = lib.tci_htimesc_noalloc(plan._h, x.ctypes.data, out.ctypes.data, plan.n_csf)
    assert rc == 0, lib.tci_last_error().decode()
    return out


def _alloc(lib, plan, x, base):
    out = np.ascontiguousarray(base.copy())
    x = np.ascontiguousarray(x)
    rc = lib.tci_htimesc(plan._h, x.ctypes.data, out.ctypes.data, plan.n_csf)
    assert rc == 0, lib.tci_last_error().decode()
    return out


def test_noalloc_matches_htimesc_twice():
    lib = tci.load_library()
    cases = [dict(n_int=6, n_v=70, n_chol=50, seed=3, domain_range=(10, 70), weak_below=16),
             dict(n_int=3, n_v=20, n_chol=12, seed=4)]           # second call: shorter vector
    sizes = []
    for k, args in enumerate(cases):
        _say("case: htimesc_noalloc call", k + 1, args)
        prob = make_problem(**args)
        plan = tci.Plan.from_problem(prob)
        sizes.append(plan.n_csf)
        x = random_ci_vector(prob, 11 + k)
        base = np.random.default_rng(k).standard_normal(plan.n_csf)
        got = _noalloc(lib, plan, x, base)
        want = _alloc(lib, plan, x, base)
        ref = base + oracle_sigma(prob, x, Integrals(prob)) + plan.info.e_ref * x
        _say(f"  n={plan.n_csf} noalloc vs htimesc {_rel(got, want):.2e}, vs oracle {_rel(got, ref):.2e}")
        assert _rel(got, want) < 1e-13
        assert _rel(got, ref) < 1e-10
    assert sizes[1] < sizes[0]


def test_noalloc_rejects_oversized_vector():
    lib = tci.load_library()
    prob = make_problem(2, 5, 7, 1)
    plan = tci.Plan.from_problem(prob)
    x = np.zeros(4)
    rc = lib.tci_htimesc_noalloc(plan._h, x.ctypes.data, x.ctypes.data, MAX_CSF + 1)
    msg = lib.tci_last_error().decode()
    _say("  oversized:", rc, msg)
    assert rc == 1 and "4194304" in msg
