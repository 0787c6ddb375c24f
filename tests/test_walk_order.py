"""The walk order of the CI vector: the oracle's restatement of the
reference's DRT (graph_border, drt_maker) and walk enumeration (tree_searchX)
gives every walk a distinct lexical index 1..count; the reference lays out
the N-1 and N-2 walks in that order.  For a closed-shell reference the order
is: N-1 walks by descending hole, N-2 walks by descending i, then descending
j.  The oracle's CI-vector layout and the problem generators follow it; the
GPU parity tests compare the library with this oracle, so a library whose
walk order differs fails them."""
import itertools

import numpy as np
import pytest

from oracle.drt import DRT, closed_shell_walk_order, brute_force_walks
from oracle.problem import make_problem
from tigerci_amd.synthetic import make_synthetic


@pytest.mark.parametrize("n_int,n_v", [(1, 2), (2, 3), (3, 4), (5, 30), (8, 5)])
def test_lexical_indices_are_a_bijection(n_int, n_v):
    g = DRT(n_int + n_v, 2 * n_int, [[2] * n_int + [0] * n_v])
    for ne, holes in ((2 * n_int, 0), (2 * n_int - 1, 1), (2 * n_int - 2, 2)):
        walks = g.tree_search(n_int, ne)
        # the tree search finds exactly the occupations with at most `holes`
        # holes (the SD band of the graph), each once
        bf = [o for o in brute_force_walks(n_int, ne)]
        assert sorted(walks) == sorted(bf)
        w = sorted(g.weight(o) for o in walks)
        assert w == list(range(1, len(walks) + 1))
        # y0 of the end vertex counts its walks
        assert abs(g.y0[(n_int, ne)]) == len(walks)


def test_two_reference_space_weights():
    """A two-reference space (closed shell plus the HOMO^2 -> LUMO^2 double,
    the shape of a 2-reference MRSDCI): every internal walk of each electron
    count gets a distinct index 1..y0."""
    n_int, n_act_v = 4, 3
    ref1 = [2, 2, 2, 2, 0, 0, 0]
    ref2 = [2, 2, 2, 0, 2, 0, 0]
    g = DRT(7, 8, [ref1, ref2])
    for ne in range(8, 4, -1):
        for lev in (n_int, 5):
            if (lev, ne) not in g.y0 or g.y0[(lev, ne)] == 0:
                continue
            walks = g.tree_search(lev, ne)
            assert sorted(g.weight(o) for o in walks) == list(range(1, len(walks) + 1))


@pytest.mark.parametrize("n_int", [1, 2, 3, 4, 6])
def test_closed_shell_order(n_int):
    singles, pairs = closed_shell_walk_order(n_int)
    assert singles == list(range(n_int - 1, -1, -1))
    assert pairs == [(i, j) for i in range(n_int - 1, -1, -1) for j in range(i, -1, -1)]


def test_layout_follows_the_reference_order():
    prob = make_problem(5, 7, 3, 1, domain_range=(2, 7), weak_below=3)
    singles, pairs = closed_shell_walk_order(prob.n_int)
    strong = [tuple(p) for p in prob.pairs.tolist()]
    assert strong == [p for p in pairs if p in set(strong)]
    syn = make_synthetic(5, 7, 3, 1, domain_range=(2, 7), weak_below=3)
    assert [tuple(p) for p in syn.pairs.tolist()] == strong
    assert np.array_equal(syn.dom_idx, prob.dom_idx)
    # unpack puts x[1 + r n_v + a] at the hole of the r-th N-1 walk
    from oracle.sigma import unpack
    x = np.zeros(prob.layout().n_csf)
    for r, i in enumerate(singles):
        x[1 + r * prob.n_v + 2] = 10.0 * (i + 1)
    _, s, _ = unpack(prob, x)
    for i in range(prob.n_int):
        assert s[i, 2] == pytest.approx(10.0 * (i + 1) / np.sqrt(2.0))
