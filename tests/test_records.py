"""Oracle pins for the ladder loop records (oracle/records.py): the record
count has the closed form sum over walks of d(d+1)/2 (every unordered pair,
diagonal included, of a walk's allowed virtuals), each record's positions
point back at its virtuals in the walk's list, records come in the
reference's order (pair a outer, then c, then walk), and an ignorable pair
removes exactly its records."""
import numpy as np

from oracle.records import ladder_records
from oracle.spin import fsn


def _lists(seed, n_walks, n_v):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_walks):
        d = int(rng.integers(0, n_v + 1))
        out.append(sorted(rng.choice(n_v, size=d, replace=False).tolist()))
    return out


def test_count_order_and_positions():
    n_int, n_v = 5, 9
    lists = _lists(1, 7, n_v)
    hw = [3 * w + 1 for w in range(7)]
    ns = [0, 2, 4, 1, 3, 2, 0]
    rec = ladder_records(n_int, n_v, hw, ns, lists)
    assert len(rec) == sum(len(x) * (len(x) + 1) // 2 for x in lists)
    keys = []
    for r in rec:
        ab, h, f0, fm, fp, pa, pc = r
        w = hw.index(h)
        a, c = lists[w][pa - 1], lists[w][pc - 1]
        ag, cg = n_int + 1 + a, n_int + 1 + c
        assert ab == ag * (ag - 1) // 2 + cg and a >= c
        s = ns[w]
        assert (f0, fm, fp) == (fsn(s), fsn(s - 2) if s >= 2 else 0, fsn(s + 2))
        keys.append((a, c, w))
    assert keys == sorted(keys)


def test_ignorable_pairs_drop_their_records():
    n_int, n_v = 3, 8
    lists = _lists(2, 5, n_v)
    hw, ns = list(range(5)), [2] * 5
    full = ladder_records(n_int, n_v, hw, ns, lists)
    ig = np.zeros((n_v, n_v), dtype=bool)
    ig[1, 6] = ig[0, 0] = True
    part = ladder_records(n_int, n_v, hw, ns, lists, ignorable=ig)
    drop = {(n_int + 7) * (n_int + 6) // 2 + n_int + 2, (n_int + 1) * n_int // 2 + n_int + 1}
    assert part == [r for r in full if r[0] not in drop]
