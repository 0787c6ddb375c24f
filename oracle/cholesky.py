"""AO -> MO transformation of Cholesky vectors.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restated from the reference (0-based indices here):

  first_half_transform   F/cholesky_transform.f90:305-364   per vector J and MO
                         i: H[J][mu][i] += C[nu][i] L^J(mu,nu) and, for mu != nu,
                         H[J][nu][i] += C[mu][i] L^J(mu,nu), over the AO pairs
                         (mu, nu) the decomposition keeps
  second_half_transform  :371-380   MO[J][j][i] = sum_mu C[mu][j] H[J][mu][i]
                         (DGEMM of the half-transformed vectors with C)
  write_MO_cho_vec       :387-411   record of pair ij = i(i+1)/2 + j (i >= j)
                         holds MO[.][j][i] over J; its norm is the 2-norm of
                         the record (cho_norms)

C is n_ao x n_mo (C[mu][p], the reference's molecular_orbitals(mu, p)).
"""
from __future__ import annotations

import numpy as np


def ao_to_mo(ao_pairs, ao_vec, C):
    """(records, norms): records[ij][J] for ij = i(i+1)/2 + j, i >= j, and the
    2-norm of every record.  ao_pairs: (n_pairs, 2) AO index pairs; ao_vec:
    (n_chol, n_pairs) values of each Cholesky vector on those pairs."""
    ao_pairs = np.asarray(ao_pairs)
    ao_vec = np.asarray(ao_vec, dtype=np.float64)
    C = np.asarray(C, dtype=np.float64)
    n_chol = ao_vec.shape[0]
    n_ao, n_mo = C.shape
    H = np.zeros((n_chol, n_ao, n_mo))
    mu, nu = ao_pairs[:, 0], ao_pairs[:, 1]
    for J in range(n_chol):
        v = ao_vec[J]
        for k in range(len(mu)):
            a, b = mu[k], nu[k]
            if a != b:
                H[J, a, :] += C[b, :] * v[k]
                H[J, b, :] += C[a, :] * v[k]
            else:
                H[J, a, :] += C[a, :] * v[k]
    MO = np.einsum("mj,Jmi->Jji", C, H)  # MO[J][j][i] = sum_mu C[mu][j] H[J][mu][i]
    n_rec = n_mo * (n_mo + 1) // 2
    rec = np.zeros((n_rec, n_chol))
    for i in range(n_mo):
        for j in range(i + 1):
            rec[i * (i + 1) // 2 + j] = MO[:, j, i]
    return rec, np.sqrt(np.abs((rec * rec).sum(axis=1)))
