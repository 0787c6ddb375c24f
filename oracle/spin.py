"""Spin functions and symmetric-group representation matrices of the
reference's SGGA (symmetric group graphical approach).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates F/spin_mod.f90:
  branching_diagram (:99-231)   vertex weights of the branching diagram
  sft_graph (:232-521)          the spin functions (paths) in the reference's
                                order, the action of every elementary
                                transposition (ET) (i, i+1) on each of them
                                (sft) and the intermediate spin
                                multiplicities of every path (spin_matrix)
  make_cycles (:523-698)        matrices of the cycles (l..k), l < k, built
                                column by column from the ETs, including the
                                (-1)^P factor of the permutation
  make_transpositions (:700-860) matrices of the transpositions (i, j),
                                (i, j) = (j+1..i)^T (j..i)
Indices follow the reference: open shells 1..n, cycle label
(k-1)(k-2)/2 + l for (l..k), transposition label (i-1)(i-2)/2 + j for i > j.
The matrices are pinned in tests/test_spin.py against the genealogical spin
functions built explicitly in spin-determinant space.
"""
from __future__ import annotations

import math

import numpy as np


def fsn(n: int, mult: int = 1) -> int:
    """Number of spin functions of n open shells with multiplicity mult
    (f(n,S) = (2S+1) n! / ((n/2+S+1)! (n/2-S)!))."""
    if n < 0 or (n + mult - 1) % 2:
        return 0
    S2 = mult - 1
    if S2 > n:
        return 0
    return mult * math.factorial(n) // (math.factorial((n + S2) // 2 + 1) * math.factorial((n - S2) // 2))


class SpinTables:
    """branching_diagram + sft_graph + make_cycles + make_transpositions for
    `n` open shells and spin multiplicity `mult`."""

    def __init__(self, n: int, mult: int = 1):
        assert n >= 2
        self.n, self.mult = n, mult
        up = (n + mult - 1) // 2
        down = n - up
        self.up, self.down = up, down
        # ---- branching_diagram: branch(a, b), a = 1..up, b = 1..down+1 -----
        br = np.zeros((up + 2, down + 3), dtype=np.int64)
        a, b = up, down + 1
        amax, bmax = up, down + 1
        while True:
            if a == up and b == down + 1:
                br[a, b] = 1
            elif b == down + 1 or a + 1 == b:
                br[a, b] = br[a + 1, b]
            elif b == 1 and a < up:
                br[a, b] = br[a + 1, b] + br[a, b + 1]
            elif a == up:
                br[a, b] = 1
            else:
                br[a, b] = br[a + 1, b] + br[a, b + 1]
            if a == 1 and b == 1:
                break
            if b == down + 1 or a + 1 == b or a == b:
                if bmax > 1 and amax == up:
                    bmax -= 1
                    a, b = amax, bmax
                    continue
                elif bmax == 1 and amax <= up:
                    amax -= 1
                    a, b = amax, bmax
                    continue
            else:
                a -= 1
                b += 1
        self.branch = br
        self.nf = fsn(n, mult)
        self._sft_graph()
        self._make_cycles()
        self._make_transpositions()

    # ---- sft_graph -------------------------------------------------------------
    def _sft_graph(self):
        n, up, down, br = self.n, self.up, self.down, self.branch
        nf = self.nf
        sft = np.zeros((n, nf + 1), dtype=np.int64)          # sft[i, w], i = 1..n-1, w = 1..nf
        spin = np.zeros((n + 1, nf + 1), dtype=np.int64)     # spin_matrix[0..n, w]
        paths = {}
        path = [False] * (n + 1)                             # path[1..n]
        pab = [[0, 0] for _ in range(n + 1)]                 # pathAB[0..n]
        path[n] = True
        shells = n - 1
        pab[n] = [0, 1]
        pab[shells] = [1, 1]
        weight = 1
        spin[0, 1:] = self.mult
        spin[n, 1:] = 1
        while True:
            while True:                                      # get_spin_function
                if pab[shells][0] == up and pab[shells][1] == down + 1 and shells == 0:
                    break
                elif pab[shells][1] == down + 1 or pab[shells][1] == pab[shells][0] + 1:
                    path[shells] = True
                    shells -= 1
                    pab[shells] = [pab[shells + 1][0] + 1, pab[shells + 1][1]]
                else:
                    path[shells] = False
                    shells -= 1
                    pab[shells] = [pab[shells + 1][0], pab[shells + 1][1] + 1]
            paths[weight] = list(path)
            for i in range(n - 1, 0, -1):
                prev, cur = path[i + 1], path[i]
                a, b = pab[i]
                spin[i, weight] = spin[i + 1, weight] - 1 if not prev else spin[i + 1, weight] + 1
                if prev and not cur and a == b:
                    sft[i, weight] = weight
                elif prev and not cur:
                    nw = weight - br[a - 1, b + 1]
                    if b + 2 <= down + 1:
                        nw += br[a - 1, b + 2]
                    sft[i, weight] = nw
                elif not prev and cur:
                    nw = weight + br[a, b]
                    if b <= a and b < down + 1:
                        nw -= br[a, b + 1]
                    sft[i, weight] = -nw
                else:
                    sft[i, weight] = 0
            shells = 1
            done = False
            while True:                                      # look_for_branch_point
                if shells == n:
                    done = True
                    break
                elif not path[shells] and pab[shells][0] < up:
                    path[shells] = True
                    weight += br[pab[shells][0], pab[shells][1] + 1]
                    shells -= 1
                    pab[shells] = [pab[shells + 1][0] + 1, pab[shells + 1][1]]
                    break
                else:
                    shells += 1
                    if pab[shells][1] < down + 1 and pab[shells][1] <= pab[shells][0] and path[shells]:
                        weight -= br[pab[shells][0], pab[shells][1] + 1]
            if done:
                break
        self.sft, self.spin_matrix, self.paths = sft, spin, paths

    # ---- make_cycles -------------------------------------------------------------
    def _make_cycles(self):
        n, nf = self.n, self.nf
        ncyc = n * (n - 1) // 2
        cyc = np.zeros((ncyc + 1, nf + 1, nf + 1))           # cycles[label, row, col], 1-based
        for i in range(1, nf + 1):
            for k in range(2, n + 1):
                reorder = [0] * (nf + 2)
                column = [0.0] * (nf + 2)
                reorder[1] = i
                column[1] = 1.0
                ns = 1
                for l in range(k - 1, 0, -1):
                    label = (k - 1) * (k - 2) // 2 + l
                    ns0 = ns                                 # Fortran do-loop bound: fixed at entry
                    for j in range(1, ns0 + 1):
                        sm = self.spin_matrix[l + 1, reorder[j]]
                        akl = 1.0 / sm
                        net = column[j]
                        sv = self.sft[l, reorder[j]]
                        if sv == 0:
                            column[j] = -column[j]
                        elif sv > 0:
                            column[j] = akl * net
                            if sv != reorder[j]:
                                ns += 1
                                reorder[ns] = abs(sv)
                                column[ns] = -math.sqrt(1.0 - akl ** 2) * net
                        else:
                            column[j] = -akl * net
                            ns += 1
                            reorder[ns] = abs(sv)
                            column[ns] = -math.sqrt(1.0 - akl ** 2) * net
                    for jj in range(1, ns + 1):
                        v = column[jj] if abs(column[jj]) >= 1e-9 else 0.0
                        cyc[label, reorder[jj], i] = v
        self.cycles = cyc

    def cycle(self, l: int, k: int) -> np.ndarray:
        """Matrix of the cycle (l..k), l < k (nf x nf, 0-based rows/cols)."""
        return self.cycles[(k - 1) * (k - 2) // 2 + l, 1:, 1:]

    # ---- make_transpositions -----------------------------------------------------
    def _make_transpositions(self):
        n, nf = self.n, self.nf
        tr = {}
        for i in range(1, n):
            tr[(i + 1, i)] = self.cycle(i, i + 1).copy()
        for i in range(3, n + 1):
            for j in range(i - 2, 0, -1):
                tr[(i, j)] = self.cycle(j + 1, i).T @ self.cycle(j, i)
        self.transp = tr

    def transposition(self, i: int, j: int) -> np.ndarray:
        """Matrix of the transposition (i, j) (symmetric; includes (-1)^P)."""
        return self.transp[(max(i, j), min(i, j))]

    def intermediate_spins(self, w: int):
        """2S of the coupling after each open shell for spin function w
        (1-based, the reference's order): vertex n is empty (2S = 0), vertex
        0 holds all open shells."""
        return [int(self.spin_matrix[v, w]) - 1 for v in range(self.n, -1, -1)]


def contracted_cycle(T: SpinTables, i: int, j: int, k: int, l: int, d1: int, d2: int, left: int,
                     right: int) -> np.ndarray:
    """(l..k)(j..i)|d1,d2 between a configuration of `left` open shells (bra)
    and one of `right` open shells (ket): contracted_cycle_stateless
    (F/spin_mod.f90:861-1312).  T holds the tables of the largest number of
    open shells; the result is fsn(left) x fsn(right)."""
    assert left >= right
    d1a, d2b = max(d1, d2), min(d1, d2)
    nsfl, nsfr = fsn(left), fsn(right)
    temp = right
    if 0 < d1a < left - 1:
        temp += 2
    if 0 < d2b < left - 3:
        temp += 2
    nsfri = fsn(temp)
    nf = T.nf
    cyc = T.cycles                      # [label, row, col], 1-based
    con = np.zeros((nf + 1, nf + 1))
    if j - i < 0:
        lab = (max(i, j) - 1) * (max(i, j) - 2) // 2 + min(i, j)
        con[1:nsfl + 1, 1:nsfri + 1] = cyc[lab, 1:nsfl + 1, 1:nsfri + 1]
    elif j - i > 0:
        lab = (max(i, j) - 1) * (max(i, j) - 2) // 2 + min(i, j)
        con[1:nsfl + 1, 1:nsfl + 1] = cyc[lab, 1:nsfl + 1, 1:nsfl + 1].T
    else:
        for o in range(1, nsfri + 1):
            con[o, o] = 1.0
    sft, sm = T.sft, T.spin_matrix
    if d1a != 0 and d2b == 0:
        cs = np.zeros_like(con)
        col = 0
        for f in range(1, nsfri + 1):
            sv = sft[d1a, f]
            if sv <= 0:
                continue
            col += 1
            if sv == f:
                cs[1:nsfl + 1, col] = math.sqrt(2.0) * con[1:nsfl + 1, f]
            else:
                ak = 1.0 / sm[d1a + 1, f]
                m = -1.0 if sv > 0 else 1.0
                sv = abs(sv)
                cs[1:nsfl + 1, col] = (-m * math.sqrt(1 - m * ak)) * con[1:nsfl + 1, f] + \
                    (m * math.sqrt(1 + m * ak)) * con[1:nsfl + 1, sv]
        con = cs
    elif d1a != 0 and d2b != 0:
        cs = np.zeros_like(con)
        col = 0
        for f in range(1, nsfri + 1):
            if col == nsfr:
                break
            s1, s2 = sft[d1a, f], sft[d2b, f]
            if s1 <= 0 or s2 <= 0:
                continue
            col += 1
            if s1 == f and s2 == f:
                cs[1:nsfl + 1, col] = 2.0 * con[1:nsfl + 1, f]
            else:
                a1 = 1.0 / sm[d1a + 1, f]
                a2 = 1.0 / sm[d2b + 1, f]
                m1 = -1.0 if s1 > 0 else 1.0
                m2 = -1.0 if s2 > 0 else 1.0
                s1, s2 = abs(s1), abs(s2)
                c_f = (-m1 * math.sqrt(1 - m1 * a1)) * (-m2 * math.sqrt(1 - m2 * a2))
                c_2 = (-m1 * math.sqrt(1 - m1 * a1)) * (m2 * math.sqrt(1 + m2 * a2))
                c_1 = (m1 * math.sqrt(1 + m1 * a1)) * (-m2 * math.sqrt(1 - m2 * a2))
                c_12 = (m1 * math.sqrt(1 + m1 * a1)) * (m2 * math.sqrt(1 + m2 * a2))
                cs[1:nsfl + 1, col] = (c_f * con[1:nsfl + 1, f] + c_1 * con[1:nsfl + 1, s1]
                                       + c_2 * con[1:nsfl + 1, s2] + c_12 * con[1:nsfl + 1, sft[d1a, s2]])
        con = cs
    out = np.zeros((nsfl, nsfr))
    if l - k < 0:
        lab = (max(k, l) - 1) * (max(k, l) - 2) // 2 + min(k, l)
        out = cyc[lab, 1:nsfl + 1, 1:nsfl + 1] @ con[1:nsfl + 1, 1:nsfr + 1]
    elif l - k > 0:
        lab = (max(k, l) - 1) * (max(k, l) - 2) // 2 + min(k, l)
        out = cyc[lab, 1:nsfl + 1, 1:nsfl + 1].T @ con[1:nsfl + 1, 1:nsfr + 1]
    else:
        out = con[1:nsfl + 1, 1:nsfr + 1].copy()
    return out
