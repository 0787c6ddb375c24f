"""Pins of the CPU oracle: brute-force determinant CI, adjointness, invariants.

These tests check the oracle (test infrastructure) against mathematics that
does not depend on its closed forms: Slater-Condon rules on explicitly built
spin-adapted CSFs, and the reference's own packed ladder loop restated
literally.  A dropped term, wrong sign, wrong factor or transposed operand in
oracle/sigma.py changes at least one element of H and fails here.
"""
import numpy as np
import pytest

from oracle.problem import (make_problem, random_ci_vector, eri_full,
                            core_hamiltonian, reference_energy)
from oracle.sigma import (Integrals, sigma, unpack, dense_hamiltonian, diagonal,
                          ALL_TERMS)
from oracle.scepper import (scepper_p1, scepper_m1, unscepper_p1, unscepper_m12,
                            ladder_packed_reference)
from oracle.davidson import davidson
from tests.detci import DetCI

CASES = [
    dict(n_int=2, n_v=3, n_chol=6, seed=1),
    dict(n_int=3, n_v=3, n_chol=8, seed=2),
    dict(n_int=3, n_v=5, n_chol=10, seed=3, domain_range=(2, 5), weak_below=3),
    # explicit one-electron integrals with off-diagonal Fock blocks (oo, ov, vv)
    dict(n_int=2, n_v=3, n_chol=6, seed=4, h_offdiag=0.3),
    dict(n_int=3, n_v=4, n_chol=8, seed=5, h_offdiag=0.2),
    dict(n_int=3, n_v=5, n_chol=10, seed=6, domain_range=(2, 5), weak_below=3, h_offdiag=0.25),
]


def _brute(prob):
    eri = eri_full(prob)
    h = core_hamiltonian(prob, eri)
    ci = DetCI(prob.n_int, prob.n_orb, h, eri)
    Hd = ci.hamiltonian()
    e0 = Hd[ci.index[ci.ref], ci.index[ci.ref]]
    n = prob.layout().n_csf
    U = np.zeros((len(ci.dets), n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        U[:, k] = ci.csf_matrix(*unpack(prob, e))
    return ci, Hd, e0, U


@pytest.mark.parametrize("case", CASES)
def test_sigma_equals_determinant_ci(case):
    prob = make_problem(**case, decay=4.0)
    ci, Hd, e0, U = _brute(prob)
    n = U.shape[1]
    assert abs(e0 - reference_energy(prob)) < 1e-10
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-13          # orthonormal CSFs
    assert max(ci.s_plus_norm2(U[:, k]) for k in range(n)) < 1e-24   # singlets
    Hb = U.T @ (Hd - e0 * np.eye(Hd.shape[0])) @ U
    Ho = dense_hamiltonian(prob)
    assert np.abs(Hb - Ho).max() < 1e-11 * max(1.0, np.abs(Hb).max())
    assert np.abs(np.diag(Hb) - diagonal(prob)).max() < 1e-11 * np.abs(Hb).max()


def test_each_term_is_needed():
    """Dropping any single term breaks agreement with the determinant CI
    (off-diagonal Fock blocks present, so the occupied-virtual terms count)."""
    prob = make_problem(n_int=3, n_v=3, n_chol=8, seed=2, decay=4.0, h_offdiag=0.3)
    ci, Hd, e0, U = _brute(prob)
    Hb = U.T @ (Hd - e0 * np.eye(Hd.shape[0])) @ U
    ints = Integrals(prob)
    n = U.shape[1]
    for drop in sorted(ALL_TERMS):
        terms = ALL_TERMS - {drop}
        H = np.stack([sigma(prob, np.eye(n)[k], ints, terms) for k in range(n)], 1)
        assert np.abs(H - Hb).max() > 1e-6, drop


def test_pack_unpack_adjoint():
    rng = np.random.default_rng(0)
    d = 7
    m2 = d * (d - 1) // 2
    v, w, M = rng.standard_normal(m2), rng.standard_normal(d), rng.standard_normal((d, d))
    vv, ww = unscepper_p1(M)
    assert abs(np.sum(scepper_p1(v, w, d) * M) - (v @ vv + w @ ww)) < 1e-12
    assert abs(np.sum(scepper_m1(v, d) * M) - v @ unscepper_m12(M)) < 1e-12
    S = scepper_p1(v, w, d)
    assert np.allclose(S, S.T) and np.allclose(np.diag(S), np.sqrt(2) * w)
    A = scepper_m1(v, d)
    assert np.allclose(A, -A.T) and A[0, 1] == v[0]
    vb, wb = unscepper_p1(S)
    assert np.allclose(vb, 2 * v) and np.allclose(wb, 2 * w)
    assert np.allclose(unscepper_m12(A), 2 * v)


@pytest.mark.parametrize("case", [dict(n_int=3, n_v=6, n_chol=10, seed=1),
                                  dict(n_int=4, n_v=9, n_chol=12, seed=3,
                                       domain_range=(3, 9), weak_below=4)])
def test_packed_ladder_matches_dense(case):
    prob = make_problem(**case)
    x = random_ci_vector(prob, 5)
    a = sigma(prob, x, terms={"ladder"})
    b = ladder_packed_reference(prob, x)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


def test_hermitian_and_davidson_roots():
    prob = make_problem(n_int=3, n_v=6, n_chol=10, seed=1)
    ints = Integrals(prob)
    H = dense_hamiltonian(prob, ints)
    assert np.abs(H - H.T).max() < 1e-12 * np.abs(H).max()
    w = np.linalg.eigvalsh(H)
    E, X, hist = davidson(lambda v: sigma(prob, v, ints), diagonal(prob, ints),
                          roots=3, energy_tol=1e-11, norm_tol=1e-9)
    assert np.abs(E - w[:3]).max() < 1e-8
    for r in range(3):
        assert np.linalg.norm(H @ X[r] - E[r] * X[r]) < 1e-4


def test_davidson_collapse_path():
    prob = make_problem(n_int=3, n_v=6, n_chol=10, seed=4)
    ints = Integrals(prob)
    w = np.linalg.eigvalsh(dense_hamiltonian(prob, ints))
    E, X, hist = davidson(lambda v: sigma(prob, v, ints), diagonal(prob, ints),
                          roots=2, energy_tol=1e-11, norm_tol=1e-9, max_vectors=10, maxit=200)
    assert np.abs(E - w[:2]).max() < 1e-8


def test_csf_counts_closed_form():
    """CSF count = 1 + n_int n_v + sum_P [d + nsp d(d-1)/2] (fsn(0)=fsn(2)=1, fsn(4)=2)."""
    prob = make_problem(10, 100, 4, 1)
    lay = prob.layout()
    assert lay.n_csf == 1 + 1000 + 500500
    prob = make_problem(16, 300, 4, 1)
    assert prob.layout().n_csf - 1 - 16 * 300 == 11522400


def test_csf_counts_match_reference_methane_output():
    """Counts printed by the reference for methane SDCI, 5 internal and 30
    virtual orbitals, no truncation (tests/Tests/Methane_SDCI/
    Methane_SDCI.output:380-396): valence 1; N-1 150; N-2 configurations
    150 (0 singles, 1 spin function) + 2475 (2 singles, 1) + 4350 (4 singles,
    2) = 11325 CSFs; 7126 configurations, 11476 CSFs in total."""
    prob = make_problem(5, 30, 4, 1)
    lay = prob.layout()
    n_v = prob.n_v
    assert lay.single_off + prob.n_int * n_v == lay.aa_off[0]
    assert lay.single_off - lay.ref_off == 1 and prob.n_int * n_v == 150
    conf0 = conf2 = conf4 = 0
    for P, (i, j) in enumerate(prob.pairs):
        d = int(lay.dsz[P])
        if i == j:                    # a=b: 0 open shells; a<b: 2
            conf0 += d
            conf2 += d * (d - 1) // 2
        else:                         # a=b: 2 open shells; a<b: 4
            conf2 += d
            conf4 += d * (d - 1) // 2
            assert lay.nsp[P] == 2
    assert (conf0, conf2, conf4) == (150, 2475, 4350)
    assert lay.n_csf - lay.aa_off[0] == conf0 + conf2 + 2 * conf4 == 11325
    assert 1 + 150 + conf0 + conf2 + conf4 == 7126
    assert lay.n_csf == 11476
