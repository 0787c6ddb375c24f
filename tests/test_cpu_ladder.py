"""The C restatement of the reference's ladder loop (oracle/cpu_ladder.c, the
timed CPU baseline) against the oracle's dense ladder and its literal Python
restatement of the same loop, on full and local domains, with and without
explicit one-electron integrals, over the whole range of outer virtuals and
split into pieces (as the bounded baseline sample runs it)."""
import numpy as np
import pytest

from oracle.problem import make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma
from oracle.scepper import ladder_packed_reference
from oracle.cpu_sigma import cpu_ladder, cpu_sigma


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("args", [
    dict(n_int=3, n_v=8, n_chol=9, seed=1),
    dict(n_int=4, n_v=20, n_chol=15, seed=2, domain_range=(5, 20), weak_below=7),
    dict(n_int=5, n_v=30, n_chol=40, seed=3, h_offdiag=0.1),
])
def test_cpu_ladder_matches_oracle(args):
    prob = make_problem(**args)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 1)
    lad = cpu_ladder(prob, x, threads=4)
    assert _rel(lad, sigma(prob, x, ints, terms={"ladder"})) < 1e-13
    assert _rel(lad, ladder_packed_reference(prob, x)) < 1e-13
    assert _rel(cpu_sigma(prob, x, ints, threads=2), sigma(prob, x, ints)) < 1e-13
    # pieces of the outer loop add up to the whole
    part = np.zeros_like(lad)
    a = np.arange(prob.n_v)
    for sel in (a[::3], a[1::3], a[2::3]):
        cpu_ladder(prob, x, part, a_list=sel, threads=3)
    assert _rel(part, lad) < 1e-14
