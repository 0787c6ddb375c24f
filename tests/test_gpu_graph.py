"""Sigma builds replayed from captured HIP graphs (TCI_SIGMA_GRAPH).

A graph replays the same kernels with the same arguments as the eager build;
only the CSF vectors are read from a device slot set before each replay.  No
kernel of the build uses atomics for sums, so every replay must equal the eager
build bit for bit.  Replays are interleaved with eager builds and host
synchronisations, on several vectors, and a precision switch (which rebuilds
the split operand copies) must drop and recapture the graphs.
"""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import make_problem, random_ci_vector

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sigma(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


PROBLEMS = [
    dict(n_int=5, n_v=60, n_chol=30, seed=21, domain_range=(10, 60), weak_below=12),  # local domains
    dict(n_int=4, n_v=48, n_chol=40, seed=22),                                         # full domains
]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("args", PROBLEMS)
def test_graph_replays_match_eager(args, precision):
    _say("case: graph replay", precision, args)
    prob = make_problem(**args)
    plan = tci.Plan.from_problem(prob)
    xs = [random_ci_vector(prob, 40 + k) for k in range(3)]
    eager = [_sigma(plan, x, precision=precision) for x in xs]
    for rep in range(3):
        for k, x in enumerate(xs):
            got = _sigma(plan, x, precision=precision, graph=True)
            assert np.array_equal(got, eager[k]), (rep, k)
            # eager work on the same plan between replays
            other = _sigma(plan, xs[(k + 1) % 3], precision=precision)
            assert np.array_equal(other, eager[(k + 1) % 3])
    # a second flag set gets its own graph
    c = tci.DeviceArray.from_numpy(xs[0])
    out = tci.DeviceArray.from_numpy(np.ones(plan.n_csf))
    plan.sigma(c, out, precision=precision, accumulate=True, add_eref=True, graph=True)
    ref = tci.DeviceArray.from_numpy(np.ones(plan.n_csf))
    plan.sigma(c, ref, precision=precision, accumulate=True, add_eref=True)
    assert np.array_equal(out.numpy(), ref.numpy())


def test_graphs_follow_precision_switches():
    _say("case: graph replay across precision switches")
    prob = make_problem(**PROBLEMS[0])
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 5)
    seq = ["bf16x3", "fp16x3", "fp64", "bf16x3", "fp16x3"]
    for p in seq:
        g = _sigma(plan, x, precision=p, graph=True)
        e = _sigma(plan, x, precision=p)
        assert np.array_equal(g, e), p


def test_graph_sigma_profiled():
    _say("case: graph replay with profiling")
    prob = make_problem(3, 20, 16, 23)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 6)
    plan.profile(True)
    e = _sigma(plan, x)
    g = _sigma(plan, x, graph=True)
    g2 = _sigma(plan, x, graph=True)
    assert np.array_equal(g, e) and np.array_equal(g2, e)
    pr = plan.profile_read("fp64")
    assert pr["sigma_calls"] == 3 and pr["ladder_launches"] == 1 and pr["sigma_ms"] > 0


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_davidson_with_graph_sigma(precision):
    _say("case: davidson with graph sigma", precision)
    prob = make_problem(5, 60, 80, 24, domain_range=(20, 60), weak_below=25, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    kw = dict(roots=2, energy_tol=1e-9, norm_tol=1e-6, precision=precision, switch_tol=1e-6)
    e0, r0 = plan.davidson(**kw)
    e1, r1 = plan.davidson(graph=True, **kw)
    _say(f"  eager {e0} ({r0['iterations']} it)  graph {e1} ({r1['iterations']} it)")
    assert r0["converged"] and r1["converged"]
    assert r0["iterations"] == r1["iterations"]
    assert np.array_equal(e0, e1)
