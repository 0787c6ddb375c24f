"""GPU Davidson: converged roots vs dense eigh and vs the oracle Davidson."""
import os

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import Problem, make_problem
from oracle.sigma import Integrals, sigma as oracle_sigma, diagonal as oracle_diagonal, dense_hamiltonian
from oracle.davidson import davidson as oracle_davidson

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_davidson_golden_roots():
    print("case: davidson golden", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "full_n3_v12.npz"))
    prob = Problem(int(z["n_int"]), int(z["n_v"]), int(z["n_chol"]), z["eps"], z["L"], z["pairs"],
                   z["dom_off"], z["dom_idx"])
    plan = tci.Plan.from_problem(prob)
    e, res = plan.davidson(roots=3, energy_tol=1e-11, norm_tol=1e-8)
    assert res["converged"]
    assert np.abs(e - plan.info.e_ref - z["eigvals"][:3]).max() < 1e-8


def test_davidson_matches_oracle_trajectory():
    print("case: davidson vs oracle", flush=True)
    prob = make_problem(5, 30, 40, 3)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    E, X, hist = oracle_davidson(lambda v: oracle_sigma(prob, v, ints), oracle_diagonal(prob, ints),
                                 roots=2, energy_tol=1e-6, norm_tol=1e-6)
    e, res = plan.davidson(roots=2, energy_tol=1e-6, norm_tol=1e-6)
    assert res["converged"]
    assert np.abs(e - plan.info.e_ref - E).max() < 1e-8
    assert res["iterations"] == hist[-1][0]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_davidson_six_roots_vs_dense(precision):
    """Six roots: a pass over the basis serves at most four roots, so the
    subspace columns, the residuals and the Gram-Schmidt sweeps each run in
    two groups (four, then two) and the new vectors are orthonormalised
    across both groups."""
    print("case: davidson six roots", precision, flush=True)
    prob = make_problem(3, 10, 12, 7)
    ref = np.linalg.eigvalsh(dense_hamiltonian(prob, Integrals(prob)))[:6]
    plan = tci.Plan.from_problem(prob)
    e, res = plan.davidson(roots=6, energy_tol=1e-11, norm_tol=1e-8, precision=precision)
    assert res["converged"]
    assert np.abs(e - plan.info.e_ref - ref).max() < 1e-8


@pytest.mark.parametrize("roots,max_vectors,precision", [(2, 10, "fp64"), (3, 15, "fp64"), (2, 12, "bf16x3")])
def test_davidson_budget_collapse(roots, max_vectors, precision):
    """A subspace budget of 5-6 vectors per root forces collapses to the
    current and previous Ritz vectors of every root (davidson_restart.f90),
    including at budgets below 7 per root, where the collapse builds its
    vectors in the store beyond the budget.  Roots against the dense spectrum
    and the oracle Davidson with the same budget."""
    print("case: davidson collapse", roots, max_vectors, precision, flush=True)
    prob = make_problem(3, 10, 12, 7)
    ints = Integrals(prob)
    ref = np.linalg.eigvalsh(dense_hamiltonian(prob, ints))[:roots]
    E, X, hist = oracle_davidson(lambda v: oracle_sigma(prob, v, ints), oracle_diagonal(prob, ints),
                                 roots=roots, energy_tol=1e-10, norm_tol=1e-7, max_vectors=max_vectors)
    plan = tci.Plan.from_problem(prob)
    e, res = plan.davidson(roots=roots, energy_tol=1e-10, norm_tol=1e-7, max_vectors=max_vectors,
                           precision=precision)
    print(f"  {res['iterations']} iterations, {res['collapses']} collapses, "
          f"root error {np.abs(e - plan.info.e_ref - ref).max():.2e}", flush=True)
    assert res["converged"] and res["collapses"] >= 1
    assert np.abs(e - plan.info.e_ref - ref).max() < 1e-8
    assert np.abs(E - ref).max() < 1e-8
