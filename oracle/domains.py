"""Allowed virtuals of internal walks under the weak-pairs (WP) and
truncation-of-virtuals (TOV) approximations.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restated from the reference (orbital indices 0-based here, internal levels
0..n_int-1, virtuals 0..n_v-1):

  find_hole_pairs        F/locist_mod.f90:533-607   the hole pairs (i <= j) of a
                         walk against every reference it is a double excitation
                         of; a walk that is a reference, or a single excitation
                         of any reference, has none and is "not double"
  find_allowed_virtuals  F/locist_mod.f90:453-518   not double: every virtual;
                         double: the union over the WP-interacting hole pairs of
                         the virtuals a with TOV(i, j, a), then sorted
  add_new_virtual_list   F/allowed_virtuals_mod.f90:163-219   one table of unique
                         sorted lists for N-1 and N-2 walks, numbered in order of
                         first appearance; each walk points at its list
  WP_throw_out           F/locist_mod.f90:210-245   a double-only walk is removed
                         when none of its hole pairs interacts under WP

The WP and TOV predicates themselves (are_interacting_WP / _TOV,
F/wp_tov_mod.f90:254-323) come from orbital spheres and cylinders of the
molecule, which belong to the front end; here they are input tables
wp[i][j] and tov[i][j][a].
"""
from __future__ import annotations

import numpy as np


def is_reference(occ, references) -> bool:
    return any(tuple(occ) == tuple(r) for r in references)


def find_hole_pairs(occ, references):
    """(only_double, pairs): pairs are (i, j), i <= j, one per reference the
    walk is a double excitation of, in reference order."""
    occ = [int(x) for x in occ]
    pairs = []
    for ref in references:
        if is_reference(occ, references):
            return False, []
        diff = [int(r) - o for r, o in zip(ref, occ)]
        holes = sum(x for x in diff if x > 0)
        particles = -sum(x for x in diff if x < 0)
        if holes > 2 or particles > 2:
            continue
        if holes == 1:
            return False, []
        first = second = -1
        for j, x in enumerate(diff):
            if x == 2:
                first = second = j
            if x == 1:
                if first < 0:
                    first = j
                else:
                    second = j
        pairs.append((first, second))
    if not pairs:
        raise ValueError("walk is not a reference, single or double excitation: %s" % (occ,))
    return True, pairs


def find_allowed_virtuals(occ, references, wp, tov, n_v: int):
    """(only_double, sorted allowed virtuals) of one walk."""
    only_double, pairs = find_hole_pairs(occ, references)
    if not only_double:
        return False, list(range(n_v))
    virt = []
    for i, j in pairs:
        if not wp[i][j]:
            continue
        for a in range(n_v):
            if tov[i][j][a] and a not in virt:
                virt.append(a)
    return True, sorted(virt)


def wp_removed(occ, references, wp) -> bool:
    """WP_throw_out: a double-only walk none of whose hole pairs interacts."""
    if is_reference(occ, references):
        return False
    only_double, pairs = find_hole_pairs(occ, references)
    if not only_double:
        return False
    return not any(wp[i][j] for i, j in pairs)


def allowed_virtual_table(walks, references, wp, tov, n_v: int):
    """The reference's table of unique lists for walks processed in the given
    order: (status, list_of_walk, lists).  status per walk: 0 reference, 1
    single excitation (every virtual), 2 double excitation kept, 3 double
    excitation removed by WP (its list is empty); list_of_walk[w] is the index
    of the walk's list in ``lists`` (-1 for a reference)."""
    status, index, lists = [], [], []
    for occ in walks:
        if is_reference(occ, references):
            status.append(0)
            index.append(-1)
            continue
        only_double, virt = find_allowed_virtuals(occ, references, wp, tov, n_v)
        status.append(1 if not only_double else (3 if wp_removed(occ, references, wp) else 2))
        try:
            k = lists.index(virt)
        except ValueError:
            k = len(lists)
            lists.append(virt)
        index.append(k)
    return status, index, lists


def random_tables(n_int: int, n_v: int, seed: int, p_wp: float = 0.7, p_tov: float = 0.5):
    """Seeded symmetric WP and TOV tables (wp[i][i] true: an orbital always
    correlates with itself, F/wp_tov_mod.f90:196-205)."""
    rng = np.random.default_rng(seed)
    wp = rng.random((n_int, n_int)) < p_wp
    wp = wp | wp.T
    np.fill_diagonal(wp, True)
    tov = rng.random((n_int, n_int, n_v)) < p_tov
    tov = tov | tov.transpose(1, 0, 2)
    return wp.astype(np.uint8), tov.astype(np.uint8)
