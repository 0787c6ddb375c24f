"""Dense pair-block restatement of sigma = (H - E_ref) c.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Closed-shell single-reference MRSDCI space of TigerCI (valence + N-1 + N-2
walks; CSF layout in oracle/problem.py).  H is the bare Hamiltonian whose
closed-shell Fock matrix is diag(eps) (core_hamiltonian in problem.py); sigma
excludes E_ref on the diagonal.

Dense representation (all pair indices ordered, i,j over internal orbitals):
  C^{ij} (n_v x n_v), C^{ji} = C^{ij T};  for a strong walk P=(i,j) the dense
  block on its domain is
     i>j:  C^{ij} = -S+(x_ab+, x_aa)/2 + S-(x_ab-)/(2 sqrt3)
     i=j:  C^{ii} = S+(x_ab, x_aa)/sqrt2
  (the minus sign is the phase of the reference's SGGA spin functions: the
  restated coupling coefficients of the valence <-> N-2 loop give
  <V|H|aa, ab singlet> = -<K, S+> and <V|H|ab triplet> = +sqrt3 <K, S->
  relative to the valence CSF; tests/test_spin.py)
  (S+ = scepper_diag_p1, S- = scepper_diag_m1); singles s_i^a = x_i^a/sqrt2
  (x_i^a at 1 + (n_int-1-i) n_v + a: N-1 walks by descending hole);
  reference c0 = x_0.  Row index of C^{ij} pairs with i.
Contravariant residual R^{ij}_{ab} (derived from the spin-orbital CISD
projections; pinned by brute-force determinant CI in tests/):
  R^{ij} = c0 K^{ij}                                   doubles<-reference
         + (e_a+e_b-e_i-e_j) C^{ij}                    same-configuration
         + sum_cd (ac|bd) C^{ij}_cd                     4-external ladder
         + sum_kl (ki|lj) C^{kl}                        4-internal
         + Z^{ij} + Z^{ji T}                           pair-pair (ia|jb)/(ij|ab)
         + sum_c [(ac|bj) s_i^c + (bc|ai) s_j^c]        3-external singles
         - sum_k [(ki|bj) s_k^a + (kj|ai) s_k^b]        3-internal singles
  Z^{ij} = sum_k [ Ct^{ik} K^{kj} - C^{ik} J^{kj} - J^{kj} C^{ik} ]
  Ct^{ik} = 2 C^{ik} - C^{ik T},  K^{kj}_{cb} = (kc|jb),  J^{kj}_{ac} = (kj|ac)
Orthonormal sigma: i>j: sigma+ = -S+^T R, sigma- = sqrt3 S-^T R; i=j:
S+^T R / sqrt2 (unscepper adjoints).  Singles:
  sigma_i^a = sqrt2 [ (e_a-e_i) s_i^a + sum_jb (2(ai|jb)-(ab|ij)) s_j^b
                      + sum_jbc (ab|jc) Ct^{ij}_bc - sum_jkb (ji|kb) Ct^{jk}_ab ]
Reference: sigma_0 = sum_ij <K^{ij}, Ct^{ij}>.

Reference routines whose terms these are (htimesc.f90:91-120 stages):
  diagtimesc.f90:46-275 (same-configuration), two_seg_mod_3.f90:274-406,
  2177-2684 (external one-body-like Hab, part of Z with k in {i,j}),
  three_and_four_seg_mod_1.f90:636-1376 ((ia|bc) singles<->doubles),
  :318-632,3240-3627 (ladder), three_and_four_seg_mod_2.f90:1000-1408,
  7137-8004 (pair-pair J/K), :8007-8162 (doubles<->reference), :305-448
  ((ij|kl)), :2929-4308 ((ai|jk)), :4942-5620 ((ij|aj)).
"""
from __future__ import annotations

import math
import numpy as np

from .problem import Problem, Layout, fock
from .scepper import (scepper_p1, scepper_m1, unscepper_p1, unscepper_m12,
                      SQRT2)

SQRT3 = math.sqrt(3.0)
ALL_TERMS = frozenset({"ref", "fock", "ladder", "hole", "ring", "s3ext",
                       "s3int", "ss", "fov"})


class Integrals:
    """Integral blocks assembled from L^J (make_iajb.f90:7, make_ijab.f90:7,
    make_all_integrals.f90:174 roles)."""

    def __init__(self, prob: Problem):
        no, nv = prob.n_int, prob.n_v
        Lf = prob.L_full()
        self.Loo = Lf[:, :no, :no]
        self.Lov = Lf[:, :no, no:]
        self.Lvv = Lf[:, no:, no:]
        nc = prob.n_chol
        Lov2 = self.Lov.reshape(nc, no * nv)
        self.K = (Lov2.T @ Lov2).reshape(no, nv, no, nv).transpose(0, 2, 1, 3).copy()  # (ia|jb) -> K[i,j,a,b]
        self.J = (self.Loo.reshape(nc, -1).T @ self.Lvv.reshape(nc, -1)).reshape(no, no, nv, nv)  # (ij|ab)
        self.O4 = (self.Loo.reshape(nc, -1).T @ self.Loo.reshape(nc, -1)).reshape(no, no, no, no)  # (ki|lj)
        # O3[k,i,b,j] = (ki|bj)
        self.O3 = (self.Loo.reshape(nc, -1).T @ self.Lov.transpose(0, 2, 1).reshape(nc, -1)
                   ).reshape(no, no, nv, no)
        self.F = fock(prob)                           # closed-shell Fock matrix f = h + G


def unpack(prob: Problem, x: np.ndarray, lay: Layout | None = None):
    """CI vector -> (c0, s[n_int,n_v], C[n_int,n_int,n_v,n_v])."""
    lay = lay or prob.layout()
    no, nv = prob.n_int, prob.n_v
    c0 = float(x[0])
    s = x[1:1 + no * nv].reshape(no, nv)[::-1] / SQRT2      # N-1 walks by descending hole
    C = np.zeros((no, no, nv, nv))
    for P in range(prob.n_pairs):
        i, j = prob.pairs[P]
        D = prob.domain(P)
        d = len(D)
        m2 = d * (d - 1) // 2
        w = x[lay.aa_off[P]:lay.aa_off[P] + d]
        vs = x[lay.ab_off[P]:lay.ab_off[P] + m2]
        if i == j:
            blk = scepper_p1(vs, w, d) / SQRT2
        else:
            vt = x[lay.ab_off[P] + m2:lay.ab_off[P] + 2 * m2]
            blk = -scepper_p1(vs, w, d) / 2.0 + scepper_m1(vt, d) / (2.0 * SQRT3)
        C[i, j][np.ix_(D, D)] = blk
        if i != j:
            C[j, i] = C[i, j].T
    return c0, s, C


def pack(prob: Problem, sig0: float, sig_s: np.ndarray, R: dict, lay: Layout | None = None):
    """(sigma_0, sigma_singles[n_int,n_v] (orthonormal), R[(i,j)] dense) -> vector."""
    lay = lay or prob.layout()
    out = np.zeros(lay.n_csf)
    out[0] = sig0
    out[1:1 + prob.n_int * prob.n_v] = sig_s[::-1].ravel()
    for P in range(prob.n_pairs):
        i, j = prob.pairs[P]
        D = prob.domain(P)
        d = len(D)
        m2 = d * (d - 1) // 2
        RP = R[(i, j)][np.ix_(D, D)]
        v, w = unscepper_p1(RP)
        if i == j:
            v, w = v / SQRT2, w / SQRT2
        else:
            v, w = -v, -w
        out[lay.aa_off[P]:lay.aa_off[P] + d] = w
        out[lay.ab_off[P]:lay.ab_off[P] + m2] = v
        if i != j:
            out[lay.ab_off[P] + m2:lay.ab_off[P] + 2 * m2] = SQRT3 * unscepper_m12(RP)
    return out


def sigma(prob: Problem, x: np.ndarray, ints: Integrals | None = None,
          terms=ALL_TERMS) -> np.ndarray:
    """sigma = (H - E_ref) x for the CSF layout of ``prob``."""
    ints = ints or Integrals(prob)
    lay = prob.layout()
    no, nv = prob.n_int, prob.n_v
    F = ints.F
    Foo, Fov, Fvv = F[:no, :no], F[:no, no:], F[no:, no:]
    c0, s, C = unpack(prob, x, lay)
    Ct = 2.0 * C - C.transpose(0, 1, 3, 2)
    K, J = ints.K, ints.J

    # ---- doubles: contravariant residual over all ordered pairs ----------
    R = np.zeros_like(C)
    if "ref" in terms:
        R += c0 * K
    if "fock" in terms:
        # sum_c f_ac C_cb + C_ac f_cb - sum_k f_ki C^{kj} - C^{ik} f_kj; with a
        # diagonal Fock matrix this is (e_a+e_b-e_i-e_j) C^{ij}
        R += (np.einsum('ac,ijcb->ijab', Fvv, C, optimize=True)
              + np.einsum('ijac,cb->ijab', C, Fvv, optimize=True)
              - np.einsum('ki,kjab->ijab', Foo, C, optimize=True)
              - np.einsum('kj,ikab->ijab', Foo, C, optimize=True))
    if "fov" in terms:
        # doubles <- singles through the occupied-virtual Fock block
        R += np.einsum('ia,jb->ijab', s, Fov) + np.einsum('ia,jb->ijab', Fov, s)
    if "hole" in terms:
        R += np.einsum('kilj,klab->ijab', ints.O4, C, optimize=True)
    if "ring" in terms:
        Z = (np.einsum('ikac,kjcb->ijab', Ct, K, optimize=True)
             - np.einsum('ikac,kjcb->ijab', C, J, optimize=True)
             - np.einsum('kjac,ikcb->ijab', J, C, optimize=True))
        R += Z + Z.transpose(1, 0, 3, 2)
    if "s3ext" in terms:
        # X[i,j,a,b] = sum_c (ac|bj) s_i^c
        Y = np.einsum('Jac,ic->Jai', ints.Lvv, s, optimize=True)
        X = np.einsum('Jai,Jjb->ijab', Y, ints.Lov, optimize=True)
        R += X + X.transpose(1, 0, 3, 2)
    if "s3int" in terms:
        T = np.einsum('kibj,ka->ijab', ints.O3, s, optimize=True)
        R -= T + T.transpose(1, 0, 3, 2)
    Rd = {}
    for P in range(prob.n_pairs):
        i, j = prob.pairs[P]
        blk = R[i, j].copy()
        if "ladder" in terms:
            D = prob.domain(P)
            Cp = C[i, j][np.ix_(D, D)]
            if np.any(Cp):                              # a zero block adds nothing
                Lp = ints.Lvv[:, D][:, :, D]            # (n_chol, d, d)
                lad = np.einsum('Jac,cd,Jbd->ab', Lp, Cp, Lp, optimize=True)
                blk[np.ix_(D, D)] += lad
        Rd[(i, j)] = blk

    # ---- singles -----------------------------------------------------------
    g = np.zeros((no, nv))
    if "fock" in terms:
        g += s @ Fvv - Foo @ s
    if "fov" in terms:
        # singles <- reference and singles <- doubles through f_ia
        g += c0 * Fov + np.einsum('jb,ijab->ia', Fov, Ct, optimize=True)
    if "ss" in terms:
        g += np.einsum('ijab,jb->ia', 2.0 * K - J, s, optimize=True)
    if "s3ext" in terms:
        W = np.einsum('Jjc,ijbc->Jib', ints.Lov, Ct, optimize=True)
        g += np.einsum('Jab,Jib->ia', ints.Lvv, W, optimize=True)
    if "s3int" in terms:
        g -= np.einsum('jibk,jkab->ia', ints.O3, Ct, optimize=True)
    sig_s = SQRT2 * g

    # ---- reference ---------------------------------------------------------
    sig0 = float(np.sum(K * Ct)) if "ref" in terms else 0.0
    if "fov" in terms:
        sig0 += 2.0 * float(np.sum(Fov * s))
    return pack(prob, sig0, sig_s, Rd, lay)


def diagonal(prob: Problem, ints: Integrals | None = None) -> np.ndarray:
    """Exact H_ii - E_ref for every CSF (preconditioner of build_next_vector_std_eig,
    davidson.f90:609-631; values diag_element_mod_2.f90:83 provides)."""
    ints = ints or Integrals(prob)
    lay = prob.layout()
    no, nv = prob.n_int, prob.n_v
    fd = np.diag(ints.F)                              # only the diagonal Fock elements enter H_ii
    eo, ev = fd[:no], fd[no:]
    Lvv = ints.Lvv
    vv_diag = np.einsum('Jaa,Jbb->ab', Lvv, Lvv)        # (aa|bb)
    vv_exch = np.einsum('Jab,Jab->ab', Lvv, Lvv)        # (ab|ab)
    Kd = np.einsum('mmaa->ma', ints.K).copy()          # (ma|ma)
    Jd = np.einsum('mmaa->ma', ints.J).copy()          # (mm|aa)
    out = np.zeros(lay.n_csf)
    out[0] = 0.0
    sing = (ev[None, :] - eo[:, None]) + 2.0 * Kd - Jd
    out[1:1 + no * nv] = sing[::-1].ravel()
    for P in range(prob.n_pairs):
        i, j = prob.pairs[P]
        D = prob.domain(P)
        d = len(D)
        r, c = np.tril_indices(d, -1)
        a, b = D[r], D[c]
        ea, eb = ev[a], ev[b]
        m2 = d * (d - 1) // 2
        if i == j:
            iiii = ints.O4[i, i, i, i]
            out[lay.aa_off[P]:lay.aa_off[P] + d] = (2 * ev[D] - 2 * eo[i] + vv_diag[D, D] + iiii
                                                     + 2 * Kd[i, D] - 4 * Jd[i, D])
            out[lay.ab_off[P]:lay.ab_off[P] + m2] = (ea + eb - 2 * eo[i] + vv_diag[a, b] + vv_exch[a, b]
                                                     + iiii + Kd[i, a] + Kd[i, b]
                                                     - 2 * Jd[i, a] - 2 * Jd[i, b])
        else:
            iijj = ints.O4[i, i, j, j]
            ijij = ints.O4[i, j, i, j]
            de = ea + eb - eo[i] - eo[j]
            ksum = Kd[i, a] + Kd[i, b] + Kd[j, a] + Kd[j, b]
            jsum = Jd[i, a] + Jd[i, b] + Jd[j, a] + Jd[j, b]
            out[lay.aa_off[P]:lay.aa_off[P] + d] = (2 * ev[D] - eo[i] - eo[j] + vv_diag[D, D]
                                                     + iijj + ijij + Kd[i, D] + Kd[j, D]
                                                     - 2 * Jd[i, D] - 2 * Jd[j, D])
            out[lay.ab_off[P]:lay.ab_off[P] + m2] = (de + vv_diag[a, b] + vv_exch[a, b] + iijj + ijij
                                                     + 0.5 * ksum - jsum)
            out[lay.ab_off[P] + m2:lay.ab_off[P] + 2 * m2] = (de + vv_diag[a, b] - vv_exch[a, b]
                                                              + iijj - ijij + 1.5 * ksum - jsum)
    return out


def dense_hamiltonian(prob: Problem, ints: Integrals | None = None) -> np.ndarray:
    """Column-by-column H - E_ref from sigma on unit vectors (tiny problems)."""
    ints = ints or Integrals(prob)
    n = prob.layout().n_csf
    H = np.zeros((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[k] = 1.0
        H[:, k] = sigma(prob, e, ints)
        e[k] = 0.0
    return H
