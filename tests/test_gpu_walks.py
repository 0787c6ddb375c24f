"""GPU parity of the internal walk table (tci_internal_walks) and the
allowed-virtual lists (tci_allowed_virtuals) with the oracle restatements
(oracle/drt.py internal_walks, oracle/domains.py allowed_virtual_table):
bit-exact, on the reference's methanol MR and methane local spaces and on
seeded multireference spaces."""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.drt import internal_walks as oracle_walks
from oracle.domains import allowed_virtual_table, random_tables

pytestmark = pytest.mark.gpu

METHANOL_REFS = [[2] * 9 + [0], [2] * 8 + [1, 1], [2] * 8 + [0, 2]]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_walks(n_int, n_v, n_elec, refs, n_frozen, n_ext):
    occ, kept = tci.internal_walks(n_int, n_v, n_elec, refs, n_ext, n_frozen)
    w, k = oracle_walks(n_int, n_v, n_elec, refs, n_ext, n_frozen)
    assert occ.tolist() == [list(x) for x in w]
    assert kept.tolist() == k
    return occ, kept


@pytest.mark.parametrize("n_ext", [0, 1, 2])
def test_methanol_walk_table(n_ext):
    _say("case: methanol MR walks, n_ext", n_ext)
    occ, kept = _check_walks(10, 40, 18, METHANOL_REFS, 2, n_ext)
    assert (len(occ), int(kept.sum())) == [(55, 36), (98, 62), (125, 76)][n_ext]


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_multireference_walk_tables(seed):
    rng = np.random.default_rng(seed)
    n_int = int(rng.integers(5, 12))
    n_occ = n_int - 2
    base = [2] * n_occ + [0, 0]
    refs = [base]
    for _ in range(int(rng.integers(0, 4))):
        r = list(base)
        i, a = int(rng.integers(0, n_occ)), int(rng.integers(n_occ, n_int))
        r[i] -= 1
        r[a] += 1
        if r not in refs:
            refs.append(r)
    n_frozen = int(rng.integers(0, 2))
    _say("case: random walks", seed, n_int, refs, n_frozen)
    for n_ext in (0, 1, 2):
        _check_walks(n_int, int(rng.integers(3, 30)), 2 * n_occ, refs, n_frozen, n_ext)


def test_methane_local_lists():
    _say("case: methane LSDCI allowed virtuals")
    n_int, n_v = 5, 30
    wp = np.ones((n_int, n_int), np.uint8)
    tov = np.ones((n_int, n_int, n_v), np.uint8)
    tov[0, 0, 14:30] = 0
    refs = [[2] * 5]
    w1, _ = tci.internal_walks(n_int, n_v, 10, refs, 1)
    w2, _ = tci.internal_walks(n_int, n_v, 10, refs, 2)
    walks = np.concatenate([w1, w2])
    st, low, lists = tci.allowed_virtuals(refs, walks, wp, tov)
    ost, olow, olists = allowed_virtual_table(walks.tolist(), refs, wp, tov, n_v)
    assert st.tolist() == ost and low.tolist() == olow and lists == olists
    l2 = [lists[k] for k in low[len(w1):]]
    assert sum(n_v - len(x) for x in l2) == 16 and sum(len(x) for x in l2) == 434


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_random_multireference_lists(seed):
    rng = np.random.default_rng(seed)
    n_int, n_v = 8, int(rng.integers(5, 70))
    base = [2] * 6 + [0, 0]
    refs = [base, [2] * 5 + [1, 1, 0], [2] * 5 + [1, 0, 1]]
    wp, tov = random_tables(n_int, n_v, seed, p_wp=0.6, p_tov=0.4)
    walks = []
    for n_ext in (1, 2):
        occ, kept = tci.internal_walks(n_int, n_v, 12, refs, n_ext)
        walks += [x for x, k in zip(occ.tolist(), kept) if k]
    _say("case: random lists", seed, n_v, len(walks))
    st, low, lists = tci.allowed_virtuals(refs, np.asarray(walks, np.int8), wp, tov)
    ost, olow, olists = allowed_virtual_table(walks, refs, wp, tov, n_v)
    assert st.tolist() == ost and low.tolist() == olow and lists == olists
    assert set(ost) >= {1, 2}


def test_rejects_walk_outside_the_sd_space():
    refs = [[2, 2, 2, 2]]
    walks = np.asarray([[2, 2, 1, 1], [0, 0, 2, 2]], np.int8)  # the second is a quadruple excitation
    with pytest.raises(tci.TigerSigmaError, match="not a reference, single or double"):
        tci.allowed_virtuals(refs, walks, np.ones((4, 4), np.uint8), np.ones((4, 4, 3), np.uint8))
