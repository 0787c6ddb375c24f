"""Ladder loop records of the iteration-1 dry run.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates all_external_3_4_loops_cho_2 (F/three_and_four_seg_mod_1.f90:318-469,
the iteration-1 branch) with build_ext_34seg_dryrun (:3142-3238): for every
virtual pair a >= c (orbital indices, a outer), unless ignorable_pair(c, a),
and for every kept N-2 walk in tree-search order whose allowed-virtual list
holds both, one record

    (a(a-1)/2 + c, head_weight, fsn(s), fsn(s-2) (0 when s < 2), fsn(s+2),
     position of a in the list, position of c)

with s the walk's open shells and 1-based positions.
"""
from __future__ import annotations

from oracle.spin import fsn


def ladder_records(n_int, n_v, head_weight, num_singles, lists, ignorable=None, mult=1):
    """lists: per walk, the sorted 0-based virtuals; ignorable: n_v x n_v
    booleans or None.  Returns the list of 7-tuples."""
    out = []
    pos = [{a: q + 1 for q, a in enumerate(lst)} for lst in lists]
    for a in range(n_v):
        for c in range(a + 1):
            if ignorable is not None and ignorable[c][a]:
                continue
            ag, cg = n_int + 1 + a, n_int + 1 + c
            for w, pw in enumerate(pos):
                if a in pw and c in pw:
                    s = num_singles[w]
                    out.append((ag * (ag - 1) // 2 + cg, head_weight[w], fsn(s, mult),
                                fsn(s - 2, mult) if s >= 2 else 0, fsn(s + 2, mult), pw[a], pw[c]))
    return out
