"""GPU parity of the ladder loop records (tci_ladder_records) with the oracle
restatement (oracle/records.py), bit-exact: the N-2 walks of the methane
LSDCI space with their TOV-truncated lists, the methanol MR space, and
seeded lists with ignorable pairs."""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.domains import allowed_virtual_table
from oracle.drt import internal_walks
from oracle.records import ladder_records

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _n2_walks(n_int, n_v, n_elec, refs, n_frozen=0):
    walks, kept = internal_walks(n_int, n_v, n_elec, refs, 2, n_frozen=n_frozen)
    hw = [w for w, k in enumerate(kept) if k]          # head weight = lexical index - 1
    occ = [walks[w] for w in hw]
    return hw, occ, [sum(1 for x in o if x == 1) for o in occ]


def _check(n_int, n_v, hw, ns, lists, ignorable=None):
    got = tci.ladder_records(n_int, n_v, hw, ns, lists, ignorable)
    ref = np.asarray(ladder_records(n_int, n_v, hw, ns, lists, ignorable), dtype=np.int64).reshape(-1, 7)
    assert got.shape == ref.shape
    assert np.array_equal(got.astype(np.int64), ref)
    return got


def test_methane_lsdci_records():
    _say("case: ladder records methane LSDCI")
    n_int, n_v = 5, 30
    wp = np.ones((n_int, n_int), np.uint8)
    tov = np.ones((n_int, n_int, n_v), np.uint8)
    tov[0, 0, 14:30] = 0
    refs = [[2] * 5]
    hw, occ, ns = _n2_walks(n_int, n_v, 10, refs)
    _, index, lists = allowed_virtual_table(occ, refs, wp, tov, n_v)
    wl = [lists[i] for i in index]
    got = _check(n_int, n_v, hw, ns, wl)
    assert len(got) == sum(len(x) * (len(x) + 1) // 2 for x in wl)


def test_methanol_mr_records():
    _say("case: ladder records methanol MR")
    n_int, n_v = 10, 40
    refs = [[2] * 9 + [0], [2] * 8 + [1, 1], [2] * 8 + [0, 2]]
    hw, occ, ns = _n2_walks(n_int, n_v, 18, refs, n_frozen=2)
    assert len(hw) == 76
    got = _check(n_int, n_v, hw, ns, [list(range(n_v))] * len(hw))
    assert len(got) == 76 * 40 * 41 // 2


@pytest.mark.parametrize("seed", [3, 4])
def test_seeded_lists_with_ignorable_pairs(seed):
    _say("case: ladder records seeded", seed)
    rng = np.random.default_rng(seed)
    n_int, n_v, nw = 7, 37, 23
    lists = [sorted(rng.choice(n_v, size=int(rng.integers(0, n_v + 1)), replace=False).tolist()) for _ in range(nw)]
    hw = sorted(rng.choice(1000, size=nw, replace=False).tolist())
    ns = rng.integers(0, 7, nw).tolist()
    ig = rng.random((n_v, n_v)) < 0.2
    _check(n_int, n_v, hw, ns, lists, ig)
    _check(n_int, n_v, hw, ns, lists, None)


def test_ladder_records_reject_bad_lists():
    with pytest.raises(tci.TigerSigmaError):
        tci.ladder_records(3, 5, [0], [0], [[2, 1]])          # unsorted
    with pytest.raises(tci.TigerSigmaError):
        tci.ladder_records(3, 5, [0], [0], [[0, 5]])          # out of range
