"""Brute-force Slater-determinant CI: the independent pin of the sigma oracle.

Spin orbitals 2p (alpha) and 2p+1 (beta) for spatial orbital p; a determinant
is a bitmask of occupied spin orbitals, canonical order = ascending index.
Matrix elements follow the Slater-Condon rules with spin-orbital integrals
<PQ|RS> = (pr|qs) delta(sP,sR) delta(sQ,sS).  CSFs are built by applying the
spin-summed generators E_pq = sum_s a+_{ps} a_{qs} to the closed-shell
reference, so nothing here depends on the closed-form sigma expressions.
"""
from __future__ import annotations

import itertools
import numpy as np


def _apply(op: list, det: int):
    """Apply a string of (create?, spinorb) operators, rightmost first.
    Returns (sign, det) or (0, None)."""
    sign = 1
    for create, p in reversed(op):
        bit = 1 << p
        occ = bool(det & bit)
        if create == occ:
            return 0, None
        below = bin(det & (bit - 1)).count("1")
        if below & 1:
            sign = -sign
        det ^= bit
    return sign, det


def apply_E(p: int, q: int, state: dict) -> dict:
    """E_pq |state> for spatial p,q."""
    out: dict = {}
    for det, c in state.items():
        for s in (0, 1):
            sg, d2 = _apply([(True, 2 * p + s), (False, 2 * q + s)], det)
            if sg:
                out[d2] = out.get(d2, 0.0) + sg * c
    return out


class DetCI:
    def __init__(self, n_int: int, n_orb: int, h: np.ndarray, eri: np.ndarray):
        self.n_int, self.n_orb = n_int, n_orb
        self.h, self.eri = h, eri
        self.ref = sum(1 << k for k in range(2 * n_int))
        occ = list(range(2 * n_int))
        vir = list(range(2 * n_int, 2 * n_orb))
        dets = {self.ref}
        for i in occ:
            for a in vir:
                if (i ^ a) & 1 == 0:
                    dets.add(self.ref ^ (1 << i) ^ (1 << a))
        for i, j in itertools.combinations(occ, 2):
            for a, b in itertools.combinations(vir, 2):
                if (i & 1) + (j & 1) == (a & 1) + (b & 1):
                    dets.add(self.ref ^ (1 << i) ^ (1 << j) ^ (1 << a) ^ (1 << b))
        self.dets = sorted(dets)
        self.index = {d: k for k, d in enumerate(self.dets)}

    # spin-orbital integrals
    def _g(self, P, Q, R, S):
        if (P & 1) != (R & 1) or (Q & 1) != (S & 1):
            return 0.0
        return self.eri[P >> 1, R >> 1, Q >> 1, S >> 1]

    def _ga(self, P, Q, R, S):
        return self._g(P, Q, R, S) - self._g(P, Q, S, R)

    def _h(self, P, Q):
        return self.h[P >> 1, Q >> 1] if (P & 1) == (Q & 1) else 0.0

    def element(self, d2: int, d1: int) -> float:
        diff = d1 ^ d2
        nd = bin(diff).count("1")
        occ1 = [k for k in range(2 * self.n_orb) if d1 >> k & 1]
        if nd == 0:
            e = sum(self._h(P, P) for P in occ1)
            e += 0.5 * sum(self._ga(P, Q, P, Q) for P in occ1 for Q in occ1)
            return e
        if nd == 2:
            I = (diff & d1).bit_length() - 1
            A = (diff & d2).bit_length() - 1
            sg, d = _apply([(True, A), (False, I)], d1)
            assert d == d2
            e = self._h(A, I) + sum(self._ga(A, K, I, K) for K in occ1)
            return sg * e
        if nd == 4:
            holes = [k for k in range(2 * self.n_orb) if (diff & d1) >> k & 1]
            parts = [k for k in range(2 * self.n_orb) if (diff & d2) >> k & 1]
            I, J = holes
            A, B = parts
            sg, d = _apply([(True, A), (True, B), (False, J), (False, I)], d1)
            assert d == d2
            return sg * self._ga(A, B, I, J)
        return 0.0

    def hamiltonian(self) -> np.ndarray:
        n = len(self.dets)
        H = np.zeros((n, n))
        for x in range(n):
            for y in range(x, n):
                v = self.element(self.dets[x], self.dets[y])
                H[x, y] = H[y, x] = v
        return H

    def vector(self, state: dict) -> np.ndarray:
        v = np.zeros(len(self.dets))
        for d, c in state.items():
            if abs(c) > 0:
                v[self.index[d]] += c
        return v

    def csf_matrix(self, c0: float, s: np.ndarray, C: np.ndarray) -> np.ndarray:
        """Determinant expansion of c0|0> + sum s_i^a E_ai|0> + 1/2 sum C^{ij}_ab E_ai E_bj|0>."""
        no = self.n_int
        nv = self.n_orb - no
        ref = {self.ref: 1.0}
        out = {self.ref: c0}
        for i in range(no):
            for a in range(nv):
                if s[i, a] != 0.0:
                    for d, c in apply_E(no + a, i, ref).items():
                        out[d] = out.get(d, 0.0) + s[i, a] * c
        for i in range(no):
            for j in range(no):
                for a in range(nv):
                    for b in range(nv):
                        if C[i, j, a, b] != 0.0:
                            st = apply_E(no + a, i, apply_E(no + b, j, ref))
                            for d, c in st.items():
                                out[d] = out.get(d, 0.0) + 0.5 * C[i, j, a, b] * c
        return self.vector(out)

    def s_plus_norm2(self, v: np.ndarray) -> float:
        """||S+ v||^2 with S+ = sum_p a+_{p alpha} a_{p beta}; equals <S^2> for M_s=0."""
        out: dict = {}
        for k, d in enumerate(self.dets):
            if v[k] == 0.0:
                continue
            for p in range(self.n_orb):
                sg, d2 = _apply([(True, 2 * p), (False, 2 * p + 1)], d)
                if sg:
                    out[d2] = out.get(d2, 0.0) + sg * v[k]
        return float(sum(c * c for c in out.values()))
