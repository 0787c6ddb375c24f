"""Walk order of the reference's CI vector: the distinct row table (DRT) of
the internal space and its lexical walk indices.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference lays out the CSFs of each class (valence, N-1, N-2) walk by
walk in the order of the walk's lexical index (index_vector_internal,
F/graph_mod.f90:1032-1589, "do i = 1, abs(y0(vertex(num_internal,
num_elec-2)))"), where the index of a walk is 1 + the sum of its arc weights
(F/locist_mod.f90:394; arc weights recorded by tree_searchX,
F/new_tree_search.f90:122-173).  Restated here:

  graph_border   (F/graph_mod.f90:101-262)   cumulative-occupation borders of
                 every level, trimmed to single and double excitations of the
                 references
  drt_maker      (F/graph_mod.f90:263-380)   vertex weights y0 and arc weights
                 y1, y2 (a 0-arc weighs 0)
  tree_searchX   (F/new_tree_search.f90:45-202) depth-first enumeration of
                 the walks between two vertices, trying steps 2, 1, 0 at every
                 level

Orbitals are numbered 1..n (internal first) as in the reference; a walk is
its occupation vector over the internal levels.
"""
from __future__ import annotations

import itertools


class DRT:
    """Borders and weights of the graph of a closed-shell or multireference
    space (references: occupation vectors over all orbitals)."""

    def __init__(self, n_orb: int, n_elec: int, references, spin_mult: int = 1, exilevel: int = 2):
        self.n = n_orb
        min_singles = spin_mult - 1
        right_singles = min_singles
        right_doubly = (n_elec - right_singles) // 2
        left_singles = max(0, min_singles)
        left_doubly = (n_elec - left_singles) // 2
        left_unocc = n_orb - left_singles - left_doubly
        left = [0] * (n_orb + 1)
        right = [0] * (n_orb + 1)
        for i in range(1, n_orb + 1):
            add = 0 if i <= left_unocc else (1 if i <= left_unocc + left_singles else 2)
            left[i] = add if i == 1 else left[i - 1] + add
            add = 2 if i <= right_doubly else (1 if i <= right_doubly + right_singles else 0)
            right[i] = add if i == 1 else right[i - 1] + add
        cum = [r[0] for r in references]
        for i in range(2, n_orb + 1):
            cum = [c + r[i - 1] for c, r in zip(cum, references)]
            left[i] = max(min(cum) - exilevel, left[i])
            right[i] = min(max(cum) + exilevel, right[i])
        self.left, self.right = left, right
        # drt_maker: y0 (paths from the top to the vertex; negative when the
        # vertex below is outside the graph), y1 / y2 (weights of the arcs
        # leaving the vertex; -1 when the arc leaves the graph)
        self.y0, self.y1, self.y2 = {}, {}, {}
        self.y0[(0, 0)] = 1 if self._belong(1, 0) else -1
        self.y1[(0, 0)] = 0 if self._belong(1, 1) else -1
        self.y2[(0, 0)] = 0 if self._belong(1, 2) else -1
        for i in range(1, n_orb + 1):
            for j in range(right[i], left[i] - 1, -1):
                w = 0
                if self.add0(i - 1, j):
                    w += abs(self.y0[(i - 1, j)])
                if self.add1(i - 1, j - 1):
                    w += abs(self.y0[(i - 1, j - 1)])
                if self.add2(i - 1, j - 2):
                    w += abs(self.y0[(i - 1, j - 2)])
                self.y0[(i, j)] = w if self._belong(i + 1, j) else -w
                if not self._belong(i + 1, j + 1):
                    w1 = -1
                else:
                    w1 = abs(self.y0[(i, j + 1)]) if self.add0(i, j + 1) else 0
                self.y1[(i, j)] = w1
                if not self._belong(i + 1, j + 2):
                    w2 = -1
                else:
                    w2 = 0
                    if self.add1(i, j + 1):
                        w2 += abs(self.y0[(i, j + 1)])
                    if self.add0(i, j + 2):
                        w2 += abs(self.y0[(i, j + 2)])
                self.y2[(i, j)] = w2

    # ci_utilities_mod.f90:49-125
    def _belong(self, i: int, j: int) -> bool:
        if i < 0 or j < 0 or i > self.n:
            return False
        return self.left[i] <= j <= self.right[i]

    def add0(self, i, j):
        return i >= 0 and j >= 0 and self._belong(i, j) and self.y0.get((i, j), 0) > 0

    def add1(self, i, j):
        return i >= 0 and j >= 0 and self._belong(i, j) and self.y1.get((i, j), -1) >= 0

    def add2(self, i, j):
        return i >= 0 and j >= 0 and self._belong(i, j) and self.y2.get((i, j), -1) >= 0

    def weight(self, occ) -> int:
        """1 + sum of the arc weights of the walk with these occupations of
        levels 1..len(occ) from the top vertex."""
        j, w = 0, 0
        for i, d in enumerate(occ):
            if d == 1:
                w += abs(self.y1[(i, j)])
            elif d == 2:
                w += abs(self.y2[(i, j)])
            j += d
        return w + 1

    def tree_search(self, bottom_level: int, bottom_elec: int):
        """Walks from the top vertex to (bottom_level, bottom_elec) in the
        order tree_searchX finds them: steps 2, 1, 0 tried at every level."""
        out = []

        def rec(i, j, occ):
            if i == bottom_level:
                if j == bottom_elec:
                    out.append(tuple(occ))
                return
            for d, ok in ((2, self.add2), (1, self.add1), (0, self.add0)):
                if ok(i, j) and j + d <= bottom_elec and self._belong(i + 1, j + d):
                    rec(i + 1, j + d, occ + [d])

        rec(0, 0, [])
        return out


def check_excitation(occ, references, n_elec: int) -> bool:
    """check_excitation (F/graph_mod.f90:963-1002): the walk is at most doubly
    excited from some reference: sum |occ - ref| over the internal levels, plus
    1 or 2 for one or two electrons outside the internal space, is at most 4."""
    ext = n_elec - sum(occ)
    return any(sum(abs(int(o) - int(r)) for o, r in zip(occ, ref)) + ext <= 4 for ref in references)


def check_frozen(occ, n_frozen: int) -> bool:
    """check_frozen (F/graph_mod.f90:1004-1026): the first n_frozen internal
    levels stay doubly occupied."""
    return all(int(occ[j]) == 2 for j in range(n_frozen))


def internal_walks(n_int: int, n_v: int, n_elec: int, references, n_ext: int, n_frozen: int = 0):
    """(walks, kept): every walk of the internal graph with n_elec - n_ext
    electrons, in lexical (index) order, and whether tree_search_internal
    keeps it (F/graph_mod.f90:700-745: kept walks get their singles count in
    the index vector, the others kill_number).  ``references`` are internal
    occupations."""
    g = DRT(n_int + n_v, n_elec, [list(r) + [0] * n_v for r in references])
    walks = sorted(g.tree_search(n_int, n_elec - n_ext), key=g.weight)
    kept = [check_excitation(w, references, n_elec) and check_frozen(w, n_frozen) for w in walks]
    return walks, kept


def closed_shell_walk_order(n_int: int, n_v: int = 4):
    """(singles, pairs): the N-1 walks as their hole orbital (0-based) and the
    N-2 walks as their hole pair (i, j), i >= j (0-based), in the order of the
    reference's CI vector for the closed-shell single reference with n_int
    internal orbitals (the virtual count does not change the internal
    order)."""
    n_orb = n_int + n_v
    ref = [2] * n_int + [0] * n_v
    g = DRT(n_orb, 2 * n_int, [ref])
    n1 = sorted(g.tree_search(n_int, 2 * n_int - 1), key=g.weight)
    n2 = sorted(g.tree_search(n_int, 2 * n_int - 2), key=g.weight)
    singles = [occ.index(1) for occ in n1]
    pairs = []
    for occ in n2:
        holes = [k for k, d in enumerate(occ) for _ in range(2 - d)]
        pairs.append((max(holes), min(holes)))
    return singles, pairs


def brute_force_walks(n_int: int, n_elec: int):
    """Every occupation vector of n_int levels with n_elec electrons (0/1/2)."""
    return [occ for occ in itertools.product((0, 1, 2), repeat=n_int) if sum(occ) == n_elec]
