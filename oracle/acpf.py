"""ACPF-type functionals: g values, the diagonal metric G and the dense
generalized eigenproblem (H - E_ref) c = E G c.

TEST INFRASTRUCTURE ONLY (see the package header): the checker for
``tci_acpf_gvalues`` and ``tci_davidson_gen_run``.

* ``gvalues`` restates get_gvalues (tiger_fortran/gtimesc.f90:308-379):
  n = correlated electrons; ACPF (flag 1) g3..g5 = 2/n; ACPF-2 (flag 2)
  g3 = g4 = (4/n)(1 - 1/(2(n-1))), g5 = 2/n; MRAQCC (flag 3) g3..g5 =
  (4/n)(1 - 1/(2(n-1))); custom (flag 4) g3 = g4 = custom, g5 = 2/n; test
  (flag 9) all 1; g1 = g2 = 1 always.  Pinned by the g values the reference's
  own test outputs print (tests/Tests/Helium2_ACPF/Helium2_ACPF.output:417-423,
  methanol-equ-MRACPF/methanol.output:794-800,
  Acetylene_SuperMol_MRACPF2/Acetylene_SuperMol_MRACPF2.output:743-749);
  MRAQCC shares the ACPF-2 g3 expression, so the acetylene output pins it too.
* ``g_diagonal`` restates build_g_diagonal (gtimesc.f90:108-184) for the
  closed-shell single-reference CSF space: the valence CSF is the reference
  (g1), the N-1 walks take g4 (:140-148), the N-2 walks g5 (:152-178).
* ``solve_dense`` is the dense generalized symmetric-definite eigenproblem
  that davidson_liu_gen_eig (davidson.f90:253-469) iterates towards; it is
  pinned by scipy's LAPACK driver and by the G = 1 case reducing to the
  standard problem.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

from .problem import Problem


def gvalues(flag: int, n_corr_elec: int, custom: float = 0.0) -> np.ndarray:
    n = float(n_corr_elec)
    g = np.ones(5)
    acpf = 2.0 / n
    acpf2 = (4.0 / n) * (1.0 - 1.0 / (2.0 * (n - 1.0)))
    if flag == 1:
        g[2:5] = acpf
    elif flag == 2:
        g[2:4] = acpf2
        g[4] = acpf
    elif flag == 3:
        g[2:5] = acpf2
    elif flag == 4:
        g[2:4] = custom
        g[4] = acpf
    elif flag == 9:
        pass
    else:
        raise ValueError(f"unknown functional {flag}")
    return g


def g_diagonal(prob: Problem, g: np.ndarray) -> np.ndarray:
    lay = prob.layout()
    d = np.full(lay.n_csf, g[4])
    d[0] = g[0]
    d[1:1 + prob.n_int * prob.n_v] = g[3]
    return d


def solve_dense(A: np.ndarray, gdiag: np.ndarray, roots: int) -> np.ndarray:
    """lowest eigenvalues E of A c = E G c with A = H - E_ref (dense) and G = diag(gdiag)"""
    w = scipy.linalg.eigh(A, np.diag(gdiag), eigvals_only=True)
    return w[:roots]
