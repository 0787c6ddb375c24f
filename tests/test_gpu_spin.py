"""GPU parity of the spin tables (tci_spin_tables) and contracted cycles
(tci_contracted_cycles) with the oracle restatement (oracle/spin.py:
SpinTables, contracted_cycle), which tests/test_spin.py pins against
genealogical spin functions built in determinant space and against the
valence <-> N-2 coupling coefficients (-1, -sqrt 3) that brute-force
determinant CI fixes.  Integer tables bit-exact, matrices to 1e-12."""
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.spin import SpinTables, contracted_cycle, fsn

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,mult", [(2, 1), (2, 3), (3, 2), (4, 1), (4, 3), (5, 2), (5, 4), (6, 1), (6, 3),
                                    (7, 2), (8, 1), (8, 5), (9, 2), (10, 1), (10, 3), (11, 2), (12, 1)])
def test_spin_tables_match_oracle(n, mult):
    _say("case: spin tables", n, mult)
    T = SpinTables(n, mult)
    g = tci.spin_tables(n, mult)
    assert g["nf"] == T.nf == fsn(n, mult)
    assert np.array_equal(g["sft"], T.sft[1:n, 1:].astype(np.int32))
    assert np.array_equal(g["spin_matrix"], T.spin_matrix[:, 1:].astype(np.int32))
    for k in range(2, n + 1):
        for l in range(1, k):
            assert np.allclose(g["cycles"][(l, k)], T.cycle(l, k), atol=1e-12, rtol=0), (l, k)
    for i in range(2, n + 1):
        for j in range(1, i):
            assert np.allclose(g["transpositions"][(i, j)], T.transposition(i, j), atol=1e-12, rtol=0), (i, j)


def test_valence_doubles_coupling_coefficient_on_gpu():
    """The record that gives <V|H|D> for the four open shells j < i < a < b:
    (-1, -sqrt 3), as the oracle's pin fixes it."""
    (cc,) = tci.contracted_cycles(4, [(3, 4, 1, 3, 3, 1, 4, 0)])
    assert cc.shape == (2, 1)
    assert np.allclose(cc.ravel(), [-1.0, -np.sqrt(3.0)], atol=1e-14)


def _oracle_defined(T, rec):
    """False where the restatement would index a spin function by a negative
    sft entry (Python would wrap around; the reference never forms such a
    record)."""
    i, j, k, l, d1, d2, left, right = rec
    d1a, d2b = max(d1, d2), min(d1, d2)
    if d1a == 0 or d2b == 0:
        return True
    temp = right + (2 if 0 < d1a < left - 1 else 0) + (2 if 0 < d2b < left - 3 else 0)
    for f in range(1, fsn(temp) + 1):
        s1, s2 = T.sft[d1a, f], T.sft[d2b, f]
        if s1 > 0 and s2 > 0 and not (s1 == f and s2 == f) and T.sft[d1a, s2] < 0:
            return False
    return True


@pytest.mark.parametrize("n_open", [4, 6, 8, 10])
def test_contracted_cycles_match_oracle(n_open):
    """Seeded records over every shape: cycles either way or none, one or two
    compressed shells, right = left, left - 2, left - 4."""
    _say("case: contracted cycles", n_open)
    T = SpinTables(n_open, 1)
    rng = np.random.default_rng(n_open)
    recs = []
    while len(recs) < 300:
        left = int(rng.choice([x for x in range(2, n_open + 1, 2)]))
        right = int(rng.choice([x for x in (left, left - 2, left - 4) if x >= 0]))
        i, j, k, l = (int(v) for v in rng.integers(1, left + 1, 4))
        d1 = int(rng.integers(0, left))
        d2 = int(rng.integers(0, left)) if rng.random() < 0.6 else 0
        temp = right + (2 if 0 < max(d1, d2) < left - 1 else 0) + (2 if 0 < min(d1, d2) < left - 3 else 0)
        if temp > n_open:
            continue
        rec = (i, j, k, l, d1, d2, left, right)
        if _oracle_defined(T, rec):
            recs.append(rec)
    got = tci.contracted_cycles(n_open, recs)
    for rec, g in zip(recs, got):
        ref = contracted_cycle(T, *rec)
        assert g.shape == ref.shape, rec
        assert np.allclose(g, ref, atol=1e-12, rtol=0), rec


def test_spin_entry_points_reject_bad_input():
    with pytest.raises(tci.TigerSigmaError):
        tci.spin_tables(13, 1)
    with pytest.raises(tci.TigerSigmaError):
        tci.spin_tables(4, 2)  # parity of n and 2S disagree
    with pytest.raises(tci.TigerSigmaError):
        tci.contracted_cycles(4, [(1, 2, 1, 2, 0, 0, 6, 4)])  # left beyond the tables
    with pytest.raises(tci.TigerSigmaError):
        tci.contracted_cycles(4, [(1, 2, 1, 2, 0, 0, 4, 4)], spin_mult=3)
    with pytest.raises(tci.TigerSigmaError):
        tci.contracted_cycles(-3, [(1, 2, 1, 2, 0, 0, 4, 4)])
