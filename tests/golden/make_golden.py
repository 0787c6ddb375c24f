"""Writes the committed golden fixtures of the sigma path (run from the repo root):

    python tests/golden/make_golden.py

Each .npz holds one seeded synthetic problem (orbital energies, Cholesky
records, pair walks and domains), CI vectors, the oracle's sigma for them, the
exact diagonal, and (small cases) the lowest eigenvalues of the dense H.  The
oracle itself is pinned against brute-force determinant CI in
tests/test_oracle_pins.py.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.problem import make_problem, random_ci_vector  # noqa: E402
from oracle.sigma import Integrals, sigma, diagonal, dense_hamiltonian  # noqa: E402

CASES = {
    "full_n3_v12": dict(n_int=3, n_v=12, n_chol=20, seed=11),
    "local_n4_v30": dict(n_int=4, n_v=30, n_chol=24, seed=12, domain_range=(6, 20), weak_below=8),
}


def write(name, args, eig=True):
    prob = make_problem(**args)
    ints = Integrals(prob)
    xs = np.stack([random_ci_vector(prob, 100 + k) for k in range(2)])
    sig = np.stack([sigma(prob, x, ints) for x in xs])
    out = dict(n_int=prob.n_int, n_v=prob.n_v, n_chol=prob.n_chol, eps=prob.eps, L=prob.L,
               pairs=prob.pairs, dom_off=prob.dom_off, dom_idx=prob.dom_idx, x=xs, sigma=sig,
               diag=diagonal(prob, ints))
    if eig:
        out["eigvals"] = np.linalg.eigvalsh(dense_hamiltonian(prob, ints))[:4]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{name}.npz")
    np.savez_compressed(path, **out)
    print(path, prob.layout().n_csf)


if __name__ == "__main__":
    write("full_n3_v12", CASES["full_n3_v12"])
    write("local_n4_v30", CASES["local_n4_v30"])
