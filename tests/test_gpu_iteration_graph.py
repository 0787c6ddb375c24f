"""Davidson iterations replayed from captured HIP graphs (TCI_DAVIDSON_GRAPH).

One replay runs an iteration's device work: the residuals and their norms,
the norm test, the preconditioner, Gram-Schmidt against the basis (the second
sweep switched on the device), orthonormalisation within the new block, the
sigma builds of the new vectors and the new columns of the subspace matrix.
It must follow the eager solver, whose dots, projections and norms sum in
the same order: the same iterations, sigma counts and roots (to 1e-10 Eh;
the second-sweep test runs on the device there and on the host here, so the
last bits may differ), against the dense spectrum and the
oracle Davidson, across groups of four roots, budget collapses, the
mixed -> FP64 switch, roots that stop on their residual norm inside a replay,
repeated solves that reuse the captured graphs, and the generalized (ACPF)
problem.
"""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.davidson import davidson as oracle_davidson
from oracle.problem import make_problem
from oracle.sigma import Integrals, dense_hamiltonian, diagonal as oracle_diagonal, sigma as oracle_sigma

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solve(plan, roots, graph, **kw):
    vec = tci.DeviceArray(roots * plan.n_csf)
    e, res = plan.davidson(roots=roots, vectors=vec, graph=graph, **kw)
    return e, res, vec.numpy().reshape(roots, plan.n_csf)


def _same_vectors(a, b, tol):
    for x, y in zip(a, b):
        s = 1.0 if np.dot(x, y) >= 0 else -1.0
        assert np.abs(x - s * y).max() < tol


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("roots", [2, 3])
def test_iteration_graph_matches_eager(precision, roots):
    _say("case: iteration graph vs eager", precision, roots)
    prob = make_problem(5, 60, 80, 24, domain_range=(20, 60), weak_below=25, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    kw = dict(energy_tol=1e-9, norm_tol=1e-6, precision=precision, switch_tol=1e-6)
    e0, r0, v0 = _solve(plan, roots, False, **kw)
    e1, r1, v1 = _solve(plan, roots, "iteration", **kw)
    _say(f"  eager {e0} ({r0['iterations']} it, {r0['sigma_mixed']}+{r0['sigma_fp64']} sigma)  "
         f"graph {e1} ({r1['iterations']} it, {r1['sigma_mixed']}+{r1['sigma_fp64']} sigma, "
         f"{r1['seconds_sigma']:.4f} of {r1['seconds']:.4f} s in sigma)")
    assert r0["converged"] and r1["converged"]
    assert r0["iterations"] == r1["iterations"]
    assert (r0["sigma_mixed"], r0["sigma_fp64"]) == (r1["sigma_mixed"], r1["sigma_fp64"])
    assert np.abs(e0 - e1).max() < 1e-10
    _same_vectors(v0, v1, 1e-6)
    assert 0.0 < r1["seconds_sigma"] < r1["seconds"]
    # a second solve replays the graphs captured by the first
    e2, r2, v2 = _solve(plan, roots, "iteration", **kw)
    assert r2["iterations"] == r1["iterations"]
    assert np.array_equal(e2, e1) and np.array_equal(v2, v1)


def test_iteration_graph_matches_oracle_trajectory():
    _say("case: iteration graph vs oracle")
    prob = make_problem(5, 30, 40, 3)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    E, X, hist = oracle_davidson(lambda v: oracle_sigma(prob, v, ints), oracle_diagonal(prob, ints),
                                 roots=2, energy_tol=1e-6, norm_tol=1e-6)
    e, res = plan.davidson(roots=2, energy_tol=1e-6, norm_tol=1e-6, graph="iteration")
    assert res["converged"]
    assert np.abs(e - plan.info.e_ref - E).max() < 1e-8
    assert res["iterations"] == hist[-1][0]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_iteration_graph_six_roots_vs_dense(precision):
    """Six roots: residuals, norms, sweeps and columns in groups of four and two."""
    _say("case: iteration graph six roots", precision)
    prob = make_problem(3, 10, 12, 7)
    ref = np.linalg.eigvalsh(dense_hamiltonian(prob, Integrals(prob)))[:6]
    plan = tci.Plan.from_problem(prob)
    kw = dict(energy_tol=1e-11, norm_tol=1e-8, precision=precision)
    e0, r0 = plan.davidson(roots=6, **kw)
    e, res = plan.davidson(roots=6, graph="iteration", **kw)
    assert res["converged"]
    assert res["iterations"] == r0["iterations"]
    assert np.abs(e - plan.info.e_ref - ref).max() < 1e-8


@pytest.mark.parametrize("roots,max_vectors,precision", [(2, 10, "fp64"), (3, 15, "fp64"), (2, 12, "bf16x3")])
def test_iteration_graph_budget_collapse(roots, max_vectors, precision):
    _say("case: iteration graph collapse", roots, max_vectors, precision)
    prob = make_problem(3, 10, 12, 7)
    ref = np.linalg.eigvalsh(dense_hamiltonian(prob, Integrals(prob)))[:roots]
    plan = tci.Plan.from_problem(prob)
    kw = dict(roots=roots, energy_tol=1e-10, norm_tol=1e-7, max_vectors=max_vectors, precision=precision)
    e0, r0 = plan.davidson(**kw)
    e, res = plan.davidson(graph="iteration", **kw)
    _say(f"  {res['iterations']} iterations, {res['collapses']} collapses")
    assert res["converged"] and res["collapses"] >= 1
    assert (res["iterations"], res["collapses"]) == (r0["iterations"], r0["collapses"])
    assert np.abs(e - plan.info.e_ref - ref).max() < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("roots", [3, 4])
def test_iteration_graph_residual_norm_stops(precision, roots):
    """An energy tolerance no root can meet: every root stops on its residual
    norm, which the replay tests on the device (its new vector zeroed, its
    sigma built and discarded, later vectors moved down by the host).  The
    roots and iteration counts equal the eager solver's; the replays build at
    least one discarded sigma more."""
    _say("case: iteration graph residual-norm stops", precision, roots)
    prob = make_problem(4, 40, 50, 11, domain_range=(12, 40), weak_below=14, l_scale=0.05)
    plan = tci.Plan.from_problem(prob)
    kw = dict(energy_tol=1e-15, norm_tol=3e-4, precision=precision, switch_tol=1e-15, maxit=200)
    e0, r0, v0 = _solve(plan, roots, False, **kw)
    e1, r1, v1 = _solve(plan, roots, "iteration", **kw)
    _say(f"  eager {r0['iterations']} it {r0['sigma_mixed']}+{r0['sigma_fp64']}  "
         f"graph {r1['iterations']} it {r1['sigma_mixed']}+{r1['sigma_fp64']}")
    assert r0["converged"] and r1["converged"]
    assert r0["iterations"] == r1["iterations"]
    assert np.abs(e0 - e1).max() < 1e-10
    assert r1["sigma_mixed"] + r1["sigma_fp64"] > r0["sigma_mixed"] + r0["sigma_fp64"]
    _same_vectors(v0, v1, 1e-6)


@pytest.mark.parametrize("functional", ["ACPF", "ACPF_TEST"])
def test_iteration_graph_generalized_problem(functional):
    """(H - E_ref) c = E G c: G in the residuals and the preconditioner, the
    metric columns <c_q, G c_k> beside those of h, all inside the replay."""
    _say("case: iteration graph, generalized problem", functional)
    prob = make_problem(4, 30, 40, 12)
    plan = tci.Plan.from_problem(prob)
    g = tci.acpf_gvalues(getattr(tci, functional), 10)
    kw = dict(roots=2, energy_tol=1e-9, norm_tol=1e-6, g_values=g)
    e0, r0 = plan.davidson(**kw)
    e1, r1 = plan.davidson(graph="iteration", **kw)
    _say(f"  eager {e0} ({r0['iterations']} it)  graph {e1} ({r1['iterations']} it)")
    assert r0["converged"] and r1["converged"]
    assert r1["iterations"] == r0["iterations"]
    assert np.abs(e0 - e1).max() < 1e-10


def test_iteration_graphs_across_solves_on_one_plan():
    """One plan, solves with 2, 3 and again 2 roots and both split precisions
    between FP64 ones: each solve equals its eager twin, whichever graphs the
    earlier solves captured (a split-precision switch drops them all)."""
    _say("case: iteration graphs across solves")
    prob = make_problem(5, 60, 80, 24, domain_range=(20, 60), weak_below=25, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    for roots, precision in [(2, "fp64"), (3, "bf16x3"), (2, "fp16x3"), (3, "fp64"), (2, "bf16x3")]:
        kw = dict(roots=roots, energy_tol=1e-9, norm_tol=1e-6, precision=precision, switch_tol=1e-6)
        e0, r0 = plan.davidson(**kw)
        e1, r1 = plan.davidson(graph="iteration", **kw)
        _say(f"  {roots} roots {precision}: {r0['iterations']} / {r1['iterations']} iterations")
        assert r0["converged"] and r1["converged"]
        assert r1["iterations"] == r0["iterations"]
        assert np.abs(e0 - e1).max() < 1e-10
