"""Generalized Davidson (ACPF-type functionals) on the GPU against the dense
generalized eigenproblem (H - E_ref) c = E G c solved by LAPACK (scipy).

The problems use molecule-sized integrals (l_scale < 1): with l_scale = 1 the
correlation energy is hundreds of Eh, the diagonal preconditioner is poor and
the iteration needs more than 50 steps (the same trajectory as a numpy
restatement of the algorithm)."""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.acpf import gvalues, g_diagonal, solve_dense
from oracle.problem import make_problem
from oracle.sigma import Integrals, dense_hamiltonian

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("args,flag,roots", [
    (dict(n_int=3, n_v=10, n_chol=20, seed=71, l_scale=0.03), 1, 1),
    (dict(n_int=3, n_v=12, n_chol=20, seed=72, domain_range=(5, 12), weak_below=6, l_scale=0.03), 2, 2),
    (dict(n_int=3, n_v=10, n_chol=20, seed=73, l_scale=0.05), 3, 2),
])
def test_gen_davidson_matches_dense(args, flag, roots):
    _say("case: generalized davidson", args, flag, roots)
    prob = make_problem(**args)
    A = dense_hamiltonian(prob, Integrals(prob))
    g = gvalues(flag, 2 * prob.n_int)
    ref = solve_dense(A, g_diagonal(prob, g), roots)
    plan = tci.Plan.from_problem(prob)
    e, res = plan.davidson(roots=roots, energy_tol=1e-11, norm_tol=1e-8, g_values=g)
    _say(f"  gpu {e - plan.info.e_ref}  dense {ref}  ({res['iterations']} it)")
    assert res["converged"]
    assert np.abs(e - plan.info.e_ref - ref).max() < 1e-8


def test_gen_davidson_unit_g_is_standard():
    _say("case: generalized davidson with G = 1")
    prob = make_problem(4, 20, 30, 74, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    e0, r0 = plan.davidson(roots=2, energy_tol=1e-10, norm_tol=1e-7)
    e1, r1 = plan.davidson(roots=2, energy_tol=1e-10, norm_tol=1e-7, g_values=tci.acpf_gvalues(tci.ACPF_TEST, 8))
    assert r0["converged"] and r1["converged"]
    assert np.abs(e0 - e1).max() < 1e-9


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_gen_davidson_mixed_precision(precision):
    _say("case: generalized davidson, mixed", precision)
    prob = make_problem(5, 40, 60, 75, domain_range=(15, 40), weak_below=18, l_scale=0.03)
    plan = tci.Plan.from_problem(prob)
    g = tci.acpf_gvalues(tci.ACPF, 10)
    kw = dict(roots=2, energy_tol=1e-9, norm_tol=1e-6, g_values=g)
    e64, r64 = plan.davidson(**kw)
    emx, rmx = plan.davidson(precision=precision, switch_tol=1e-6, **kw)
    _say(f"  fp64 {e64} ({r64['iterations']} it)  mixed {emx} ({rmx['sigma_mixed']} + {rmx['sigma_fp64']} sigma)")
    assert r64["converged"] and rmx["converged"] and rmx["sigma_mixed"] > 0
    assert np.abs(emx - e64).max() < 1e-6
    # ACPF lies below the CI energy (g < 1 on the excited CSFs)
    eci, _ = plan.davidson(roots=1, energy_tol=1e-9, norm_tol=1e-6)
    assert e64[0] < eci[0]
