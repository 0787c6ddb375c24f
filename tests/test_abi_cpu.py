"""CPU tests of the C-ABI library: it loads, exports every declared entry point,
and rejects malformed problems with TCI_ERR_ARG before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

import tigerci_amd as tci

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tigersigma.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(tci_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = tci.load_library()
    names = declared_functions()
    assert len(names) >= 9
    for n in names:
        assert hasattr(lib, n), n


def _problem(**over):
    from oracle.problem import make_problem
    p = make_problem(2, 4, 3, 1)
    d = dict(n_int=p.n_int, n_v=p.n_v, n_chol=p.n_chol, eps=p.eps, chol=p.L, pairs=p.pairs,
             dom_off=p.dom_off, dom_idx=p.dom_idx)
    d.update(over)
    return d


@pytest.mark.parametrize("bad", [
    dict(pairs=np.array([[0, 1], [1, 1], [1, 0]])),            # j > i
    dict(dom_idx=np.array([0, 2, 1, 3] * 3, np.int32)),        # unsorted domain
    dict(dom_idx=np.array([0, 1, 2, 9] * 3, np.int32)),        # out of range
    dict(n_chol=0, chol=np.zeros(0)),
])
def test_invalid_problem_rejected(bad):
    kw = _problem(**bad)
    with pytest.raises(tci.TigerSigmaError) as e:
        tci.Plan(**kw)
    assert "error 1" in str(e.value)


def test_missing_library_is_loud(tmp_path):
    with pytest.raises(tci.TigerSigmaError):
        tci.load_library(str(tmp_path / "nope.so"))


def test_golden_fixtures_match_oracle():
    """The committed fixtures are reproduced by the oracle."""
    from oracle.problem import Problem
    from oracle.sigma import sigma, diagonal
    for name in ("full_n3_v12", "local_n4_v30"):
        z = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"))
        prob = Problem(int(z["n_int"]), int(z["n_v"]), int(z["n_chol"]), z["eps"], z["L"], z["pairs"],
                       z["dom_off"], z["dom_idx"])
        for x, s in zip(z["x"], z["sigma"]):
            assert np.abs(sigma(prob, x) - s).max() < 1e-12 * np.abs(s).max()
        assert np.abs(diagonal(prob) - z["diag"]).max() < 1e-12 * np.abs(z["diag"]).max()


def test_synthetic_generator_matches_oracle():
    """The bench's chunked generator reproduces the oracle's problem bit for bit."""
    from oracle.problem import make_problem
    from tigerci_amd.synthetic import make_synthetic
    for args in [dict(n_int=3, n_v=9, n_chol=7, seed=5),
                 dict(n_int=5, n_v=30, n_chol=11, seed=6, domain_range=(4, 25), weak_below=8)]:
        a = make_problem(**args)
        b = make_synthetic(**args, chunk_rows=17)
        for f in ("eps", "L", "pairs", "dom_off", "dom_idx"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
