"""Oracle pins for the internal walk table and the allowed-virtual lists
(oracle/drt.py: internal_walks, check_excitation, check_frozen;
oracle/domains.py), against the counts the reference's own test outputs print.

* Methanol MRACPF (tests/Tests/methanol-equ-MRACPF/methanol.output:386-439):
  3 references over 10 internal levels (8 inactive, 2 active), 18 electrons,
  2 frozen levels, 40 virtuals, no local truncation.  The DRT holds 55 / 98 /
  125 valence / N-1 / N-2 walks; tree_search_internal keeps 36 / 62 / 76
  (19 / 36 / 49 "thrown out"); the configurations by singles count and the CSF
  totals (36, 3920, 147160; 151116) follow with fsn; 76 "holes", 40 virtuals
  each.
* Methane LSDCI (tests/Tests/Methane_LSDCI/Methane_LSDCI.output:344-390): one
  closed-shell reference over 5 internal levels, 30 virtuals, WP keeps every
  pair, and the TOV spheres (:253-291) keep the 16 virtuals centred on the
  hydrogens (orbitals 20-35) away from the carbon core pair only: 16 N-2
  "excitations removed to 1 virtual", 360 "to 2 virtuals", 134 / 2131 / 4350
  configurations with 0 / 2 / 4 singles, 10965 N-2 CSFs, 11116 CSFs in total,
  15 holes with 28.9333 virtuals on average.  The two removal counts fix the
  distribution: the 2-virtual count 465 - k(k+1)/2 summed over walks is
  strictly concave in the removed count, so 16 removed in total with 360
  pair removals means one walk loses all 16.
"""
import numpy as np
import pytest

from oracle.drt import DRT, internal_walks, check_excitation, check_frozen
from oracle.domains import (allowed_virtual_table, find_allowed_virtuals, find_hole_pairs, random_tables,
                            wp_removed)
from oracle.spin import fsn


def _open(w):
    return sum(1 for x in w if x == 1)


def _configs(walks, kept, n_ext, lists=None, n_v=0):
    """{singles: configurations} of one class (valence, N-1, N-2)."""
    out = {}
    for k, (w, ok) in enumerate(zip(walks, kept)):
        if not ok:
            continue
        o = _open(w)
        d = n_v if lists is None else len(lists[k])
        if n_ext == 0:
            out[o] = out.get(o, 0) + 1
        elif n_ext == 1:
            out[o + 1] = out.get(o + 1, 0) + d
        else:
            out[o] = out.get(o, 0) + d
            out[o + 2] = out.get(o + 2, 0) + d * (d - 1) // 2
    return out


def _csfs(cfg):
    return sum(fsn(s) * n for s, n in cfg.items())


METHANOL_REFS = [[2] * 9 + [0], [2] * 8 + [1, 1], [2] * 8 + [0, 2]]


def test_methanol_mr_walk_table_matches_reference_output():
    n_int, n_v, n_elec = 10, 40, 18
    res = {}
    for n_ext in (0, 1, 2):
        walks, kept = internal_walks(n_int, n_v, n_elec, METHANOL_REFS, n_ext, n_frozen=2)
        res[n_ext] = (walks, kept)
    assert [len(res[e][0]) for e in (0, 1, 2)] == [55, 98, 125]
    assert [len(res[e][0]) - sum(res[e][1]) for e in (0, 1, 2)] == [19, 36, 49]
    cv = _configs(*res[0], 0)
    c1 = _configs(*res[1], 1, n_v=n_v)
    c2 = _configs(*res[2], 2, n_v=n_v)
    assert cv == {0: 8, 2: 28}
    assert c1 == {2: 1040, 4: 1440}
    assert c2 == {0: 520, 2: 12060, 4: 38040, 6: 11700}
    assert (_csfs(cv), _csfs(c1), _csfs(c2)) == (36, 3920, 147160)
    assert _csfs(cv) + _csfs(c1) + _csfs(c2) == 151116
    assert sum(res[0][1]) + sum(res[1][1]) + sum(res[2][1]) == 174
    # no local truncation: every kept N-2 walk ("hole") keeps all 40 virtuals
    wp = np.ones((n_int, n_int), np.uint8)
    tov = np.ones((n_int, n_int, n_v), np.uint8)
    n2 = [w for w, k in zip(*res[2]) if k]
    assert len(n2) == 76
    assert all(len(find_allowed_virtuals(w, METHANOL_REFS, wp, tov, n_v)[1]) == 40 for w in n2)


def _methane_tables():
    n_int, n_v = 5, 30
    wp = np.ones((n_int, n_int), np.uint8)
    tov = np.ones((n_int, n_int, n_v), np.uint8)
    tov[0, 0, 14:30] = 0  # orbitals 20-35: 0.8 bohr spheres on the hydrogens, apart from the C 1s sphere
    return n_int, n_v, wp, tov


def test_methane_lsdci_truncation_matches_reference_output():
    n_int, n_v, wp, tov = _methane_tables()
    refs = [[2] * 5]
    walks1, kept1 = internal_walks(n_int, n_v, 10, refs, 1)
    walks2, kept2 = internal_walks(n_int, n_v, 10, refs, 2)
    w1 = [w for w, k in zip(walks1, kept1) if k]
    w2 = [w for w, k in zip(walks2, kept2) if k]
    status, index, lists = allowed_virtual_table(w1 + w2, refs, wp, tov, n_v)
    l1 = [lists[i] for i in index[:len(w1)]]
    l2 = [lists[i] for i in index[len(w1):]]
    assert sum(n_v - len(x) for x in l1) == 0
    assert sum(n_v - len(x) for x in l2) == 16
    assert sum(n_v * (n_v + 1) // 2 - len(x) * (len(x) + 1) // 2 for x in l2) == 360
    c2 = _configs(w2, [True] * len(w2), 2, l2)
    assert c2 == {0: 134, 2: 2131, 4: 4350}
    assert _csfs(c2) == 10965
    assert 1 + len(w1) * n_v + _csfs(c2) == 11116
    assert len(w2) == 15 and abs(sum(len(x) for x in l2) / 15 - 28.9333333333333) < 1e-12
    assert status[:len(w1)] == [1] * len(w1) and status[len(w1):] == [2] * len(w2)
    assert len(lists) == 2 and lists[0] == list(range(n_v)) and lists[1] == list(range(14))


def test_hole_pairs_of_closed_shell_walks():
    """Single closed-shell reference: every N-2 walk is a double excitation
    with the hole pair (i, j) of its two missing electrons; every N-1 walk a
    single excitation (all virtuals); the list of a double walk is its TOV row
    when WP holds and empty otherwise."""
    n_int, n_v = 6, 9
    ref = [[2] * n_int]
    wp, tov = random_tables(n_int, n_v, 5)
    walks, _ = internal_walks(n_int, n_v, 2 * n_int, ref, 2)
    for w in walks:
        holes = [k for k, x in enumerate(w) for _ in range(2 - x)]
        dbl, pairs = find_hole_pairs(w, ref)
        assert dbl and pairs == [(min(holes), max(holes))]
        i, j = pairs[0]
        _, virt = find_allowed_virtuals(w, ref, wp, tov, n_v)
        assert virt == ([a for a in range(n_v) if tov[i, j, a]] if wp[i, j] else [])
        assert wp_removed(w, ref, wp) == (not wp[i, j])
    for w in internal_walks(n_int, n_v, 2 * n_int, ref, 1)[0]:
        assert find_allowed_virtuals(w, ref, wp, tov, n_v) == (False, list(range(n_v)))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_multireference_lists_against_independent_definition(seed):
    """Against a statement by determinant counting: a walk is 'not double' if
    it equals a reference or is reachable from one by moving one electron out
    of a level of the reference (one hole); otherwise each reference with two
    holes (and at most two particles) contributes the TOV row of its hole pair
    under WP; lists are sorted and unique, and the table numbers unique lists
    in order of first appearance."""
    rng = np.random.default_rng(seed)
    n_int, n_v, n_elec = 6, 7, 10
    base = [2] * 5 + [0]
    refs = [base]
    while len(refs) < 3:
        r = list(base)
        i, a = int(rng.integers(0, 5)), int(rng.integers(3, 6))
        if r[i] > 0 and r[a] < 2 and i != a:
            r[i] -= 1
            r[a] += 1
            if r not in refs:
                refs.append(r)
    wp, tov = random_tables(n_int, n_v, seed + 10)
    walks = []
    for n_ext in (1, 2):
        w, k = internal_walks(n_int, n_v, n_elec, refs, n_ext)
        walks += [x for x, ok in zip(w, k) if ok]
    status, index, lists = allowed_virtual_table(walks, refs, wp, tov, n_v)
    seen = []
    for w, st, k in zip(walks, status, index):
        ex = []
        single = False
        for r in refs:
            d = np.asarray(r) - np.asarray(w)
            h, p = int(d[d > 0].sum()), int(-d[d < 0].sum())
            if h == 1 and p <= 2:
                single = True
            if h == 2 and p <= 2:
                idx = [q for q in range(n_int) for _ in range(max(int(d[q]), 0))]
                ex.append((idx[0], idx[1]))
        want = list(range(n_v)) if single else sorted({a for (i, j) in ex if wp[i, j] for a in range(n_v) if tov[i, j, a]})
        assert lists[k] == want
        assert st == (1 if single else (2 if any(wp[i, j] for i, j in ex) else 3))
        if lists[k] not in seen:
            assert k == len(seen)
            seen.append(lists[k])
    assert len(seen) == len(lists)


def test_check_excitation_and_frozen_levels():
    refs = [[2, 2, 1, 1], [2, 2, 2, 0]]
    assert check_excitation([2, 2, 1, 1], refs, 6)
    assert check_excitation([2, 1, 1, 1], refs, 6)           # |diff| 1 + 1 external
    assert check_excitation([1, 1, 1, 1], refs, 6)           # 2 + 2 external
    assert not check_excitation([1, 1, 1, 0], refs, 6)       # at least 3 + 3
    assert check_frozen([2, 1, 1, 1], 1) and not check_frozen([2, 1, 1, 1], 2)
