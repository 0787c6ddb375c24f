"""Pair-block pack/unpack and the reference's packed 4-external ladder.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pack/unpack restates F/scepmaker_mod.f90:
  scepper_diag_p1 (:121-166)   packed ab v + aa w -> symmetric M, M_aa = sqrt2*w
  scepper_diag_m1 (:168-212)   packed ab v -> antisymmetric M, upper (row<col) = +v
  unscepper_diag_sigma_atomic_p1 (:633-685)  v += M_ji + M_ij ; w += sqrt2*M_ii
  unscepper_diag_sigma_atomic_m12 (:687-733) v += M_ji - M_ij
Packed order: (2,1),(3,1),(3,2),... i.e. numpy.tril_indices(d, -1), index
i(i-1)/2 + j for domain positions i > j (F/scepmaker_mod.f90:84-96).
"""
from __future__ import annotations

import math
import numpy as np

from .problem import Problem

SQRT2 = math.sqrt(2.0)


def tri(d: int):
    """(rows, cols) of the packed order: rows > cols."""
    return np.tril_indices(d, -1)


def scepper_p1(v: np.ndarray, w: np.ndarray, d: int) -> np.ndarray:
    """F/scepmaker_mod.f90:121-166."""
    M = np.zeros((d, d))
    r, c = tri(d)
    M[c, r] = v
    M[r, c] = v
    M[np.arange(d), np.arange(d)] = SQRT2 * w
    return M


def scepper_m1(v: np.ndarray, d: int) -> np.ndarray:
    """F/scepmaker_mod.f90:168-212 (mode -1): scep(j,i)=+v, scep(i,j)=-v, i>j."""
    M = np.zeros((d, d))
    r, c = tri(d)
    M[c, r] = v
    M[r, c] = -v
    return M


def unscepper_p1(M: np.ndarray):
    """Adjoint of scepper_p1 (F/scepmaker_mod.f90:633-685): returns (v, w)."""
    d = M.shape[0]
    r, c = tri(d)
    return M[c, r] + M[r, c], SQRT2 * np.diagonal(M).copy()


def unscepper_m12(M: np.ndarray) -> np.ndarray:
    """Adjoint of scepper_m1 (F/scepmaker_mod.f90:687-733)."""
    d = M.shape[0]
    r, c = tri(d)
    return M[c, r] - M[r, c]


def ladder_packed_reference(prob: Problem, x: np.ndarray) -> np.ndarray:
    """Literal restatement of the 4-external ladder of htimesc stage C.

    Loop structure of all_external_3_4_loops_cho_2 / ext34_parallel_load /
    ext_3_4_seg_loops_vec_lmo_res_2 / sort_ints_34_vec_lmo
    (F/three_and_four_seg_mod_1.f90:318-469, 472-632, 3240-3518, 3520-3627):
    for every virtual pair (a,c), c<=a, assemble I(x,y) = (a x|c y) from L^J
    (DGEMM at :556/:589), then for every N-2 walk whose domain holds a and c
    (the abcd record of build_ext_34seg_dryrun, :3142-3238) form
      mod2 = I(D_j,D_i)+I(D_i,D_j), mod3 = I(D_j,D_i)-I(D_i,D_j), mod4 = I(D_i,D_i)
    and add the DGEMV products to the target CSF (:3354-3487).
    Returns the ladder part of sigma only (tiny problems).
    """
    lay = prob.layout()
    n_int, n_v = prob.n_int, prob.n_v
    Lf = prob.L_full()
    Lv = Lf[:, n_int:, n_int:]                       # (n_chol, n_v, n_v)
    sig = np.zeros(lay.n_csf)
    pos = []
    for P in range(prob.n_pairs):
        p = -np.ones(n_v, np.int64)
        p[prob.domain(P)] = np.arange(prob.dsize(P))
        pos.append(p)
    for a in range(n_v):
        for c in range(a + 1):
            I = Lv[:, a, :].T @ Lv[:, c, :]            # I[x,y] = (a x|c y)
            for P in range(prob.n_pairs):
                pa, pc = pos[P][a], pos[P][c]
                if pa < 0 or pc < 0:
                    continue
                D = prob.domain(P)
                d = len(D)
                r, cc = tri(d)
                t1 = I[D[cc], D[r]]
                t2 = I[D[r], D[cc]]
                mod2, mod3 = t1 + t2, t1 - t2
                mod4 = I[D, D]
                m2 = d * (d - 1) // 2
                aa = lay.aa_off[P]
                ab = lay.ab_off[P]
                c_aa = x[aa:aa + d]
                c_ab_s = x[ab:ab + m2]
                if a != c:
                    tgt = pa * (pa - 1) // 2 + pc
                    sig[ab + tgt] += c_ab_s @ mod2 + SQRT2 * (c_aa @ mod4)
                    if lay.nsp[P] == 2:
                        c_ab_t = x[ab + m2:ab + 2 * m2]
                        sig[ab + m2 + tgt] += -(c_ab_t @ mod3)
                else:
                    # triplet part adds (1/sqrt2) C_ab.mod3 == 0 identically
                    sig[aa + pa] += c_aa @ mod4 + (c_ab_s @ mod2) / SQRT2
    return sig
