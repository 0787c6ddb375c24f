"""The reference's SGGA spin tables and the phase of the CSFs.

1. oracle/spin.py restates branching_diagram, sft_graph, make_cycles and
   make_transpositions (F/spin_mod.f90).  Their transposition matrices must
   equal the action of the spin permutations on genealogical (Yamanouchi-
   Kotani) spin functions built here in spin-determinant space with
   Condon-Shortley coupling coefficients, coupling the last open shell first,
   times the (-1)^P of a transposition: an independent construction.
2. The reference's coupling coefficient of the valence <-> N-2 loop (loop type
   2: chain C8, member 6, Px = P3 = (k+1..i)(l..j)|k+1,l,
   three_and_four_seg_mod_2.f90:1524-1533, with contracted_cycle_stateless,
   spin_mod.f90:861-1312) for the four open shells j < i < a < b: its
   magnitudes (1, sqrt3) are those that brute-force determinant CI fixes for
   <V|H|D>, and only the loop indices (4,3,2,1) give them.  Its signs fix
   the phase of the aa / singlet / triplet CSFs of an i > j walk relative to
   the valence, which the oracle (and the library, through the GPU parity
   tests) carries: the valence component is
   cc1 <K^T, S+(c)> + cc2 <K^T, S-(c)> (do_nm2_v_multiplications_3_lmo,
   :8007-8162, K transposed for loop types other than 4).
"""
import numpy as np
import pytest

from oracle.spin import SpinTables, contracted_cycle, fsn
from oracle.problem import make_problem
from oracle.sigma import Integrals, sigma
from oracle.scepper import scepper_p1, scepper_m1


def _genealogical(order, twoS, n, M2):
    """Spin function of n electrons coupled in `order` with intermediate 2S
    values twoS (after each coupled electron), projection 2M = M2; a vector
    over the 2^n spin determinants (bit e-1 set: electron e is alpha)."""
    def couple(m, m2):
        if m == 0:
            return {(): 1.0} if m2 == 0 else {}
        s2, p2 = twoS[m - 1], (twoS[m - 2] if m >= 2 else 0)
        S, M = s2 / 2, m2 / 2
        out = {}
        for sig in (1, -1):
            mp2 = m2 - sig
            if abs(mp2) > p2:
                continue
            if s2 == p2 + 1:    # |S,M> = sqrt((S+M)/2S)|S-1/2,M-1/2>up + sqrt((S-M)/2S)|S-1/2,M+1/2>down
                c = np.sqrt((S + sig * M) / (2 * S))
            else:               # |S,M> = -sqrt((S-M+1)/(2S+2))|..>up + sqrt((S+M+1)/(2S+2))|..>down
                c = -np.sqrt((S - M + 1) / (2 * S + 2)) if sig > 0 else np.sqrt((S + M + 1) / (2 * S + 2))
            for k, v in couple(m - 1, mp2).items():
                out[k + (sig,)] = out.get(k + (sig,), 0.0) + c * v
        return out
    vec = np.zeros(2 ** n)
    for spins, v in couple(n, M2).items():
        idx = sum(1 << (e - 1) for pos, e in enumerate(order) if spins[pos] > 0)
        vec[idx] += v
    return vec


def _swap(n, i, j):
    P = np.zeros((2 ** n, 2 ** n))
    for x in range(2 ** n):
        bi, bj = (x >> (i - 1)) & 1, (x >> (j - 1)) & 1
        y = (x & ~(1 << (i - 1)) & ~(1 << (j - 1))) | (bj << (i - 1)) | (bi << (j - 1))
        P[y, x] = 1.0
    return P


@pytest.mark.parametrize("n,mult", [(2, 1), (3, 2), (4, 1), (4, 3), (5, 2), (6, 1), (6, 3), (7, 2)])
def test_transpositions_match_genealogical_functions(n, mult):
    T = SpinTables(n, mult)
    assert T.nf == fsn(n, mult)
    V = np.array([_genealogical(list(range(n, 0, -1)),
                                [int(T.spin_matrix[v, w]) - 1 for v in range(n - 1, -1, -1)], n, mult - 1)
                  for w in range(1, T.nf + 1)]).T
    assert np.allclose(V.T @ V, np.eye(T.nf), atol=1e-13)
    for i in range(2, n + 1):
        for j in range(1, i):
            U = V.T @ _swap(n, i, j) @ V
            assert np.allclose(-U, T.transposition(i, j), atol=1e-12), (i, j)


def test_coupling_order_matters():
    """Coupling the first open shell first gives other matrices (n = 6): the
    check above does pin the reference's order."""
    T = SpinTables(6, 1)
    V = np.array([_genealogical(list(range(1, 7)), [int(T.spin_matrix[v, w]) - 1 for v in range(5, -1, -1)], 6, 0)
                  for w in range(1, T.nf + 1)]).T
    assert not all(np.allclose(-(V.T @ _swap(6, i, j) @ V), T.transposition(i, j))
                   for i in range(2, 7) for j in range(1, i))


def _valence_nm2_cc(ibar, jbar, kbar, lbar):
    T = SpinTables(4, 1)
    return contracted_cycle(T, kbar + 1, ibar, lbar, jbar, kbar + 1, lbar, 4, 0).ravel()


def test_valence_doubles_coupling_coefficient():
    cc = _valence_nm2_cc(4, 3, 2, 1)
    assert np.allclose(cc, [-1.0, -np.sqrt(3.0)], atol=1e-14)
    # every other admissible set of loop indices misses the magnitudes
    for lb in range(0, 4):
        for kb in range(lb, 4):
            for jb in range(kb + 1, 4):
                if (jb + 1, jb, kb, lb) != (4, 3, 2, 1):
                    c = _valence_nm2_cc(jb + 1, jb, kb, lb)
                    assert not np.allclose(np.abs(c), [1.0, np.sqrt(3.0)])


def test_csf_phase_of_pair_walks():
    """sigma_0 of a unit vector on an i > j walk equals the reference's
    valence <-> N-2 formula; flipping the triplet (or singlet) sign of the
    oracle's CSFs breaks it."""
    cc = _valence_nm2_cc(4, 3, 2, 1)
    prob = make_problem(3, 5, 6, 4)
    ints = Integrals(prob)
    lay = prob.layout()
    checked = 0
    for P, (i, j) in enumerate(prob.pairs):
        if i == j:
            continue
        D = prob.domain(P)
        d = len(D)
        m2 = d * (d - 1) // 2
        Kt = ints.K[i, j][np.ix_(D, D)].T            # K^{ij}_ab = (ia|jb), transposed for loop type 2
        for k in range(lay.aa_off[P], lay.ab_off[P] + 2 * m2):
            x = np.zeros(lay.n_csf)
            x[k] = 1.0
            w, vs, vt = x[lay.aa_off[P]:lay.aa_off[P] + d], x[lay.ab_off[P]:lay.ab_off[P] + m2], \
                x[lay.ab_off[P] + m2:lay.ab_off[P] + 2 * m2]
            ref = cc[0] * np.sum(Kt * scepper_p1(vs, w, d)) + cc[1] * np.sum(Kt * scepper_m1(vt, d))
            assert abs(sigma(prob, x, ints)[0] - ref) < 1e-12 * max(1.0, abs(ref))
            checked += 1
    assert checked > 20


def test_csf_phase_of_single_walks():
    """The valence <-> N-1 element of the reference: in
    one_intern_seg_compl_vec_lmo (F/two_seg_mod_3.f90:3018-3105, loop type 2
    of one_internal_seg_2_cho, :617-745) Ka = G(a) from make_gij_a_pao
    (:3332-3427), h_ia + sum_{k != i} [2(ia|kk) - (ik|ka)] + (ia|ii), added on
    the diagonal by add_in_Gij_vec (case 3, :849-879), times
    Px = contracted_cycle_stateless(jbar, ibar, 0, 0, jbar, 0, 2, 0) with
    jbar = 1 (the open shell i) and ibar = 2 (one_info_simplifier, :747-843).
    sigma_0 of every unit vector on a single, and sigma of the valence unit
    vector on every single, equal Px G(a): the N-1 CSFs carry the reference's
    phase relative to the valence (a flipped single CSF would give -Px G)."""
    px = contracted_cycle(SpinTables(2, 1), 1, 2, 0, 0, 1, 0, 2, 0)
    assert px.shape == (1, 1) and abs(px[0, 0] - np.sqrt(2.0)) < 1e-14
    prob = make_problem(3, 5, 6, 4, h_offdiag=0.2)
    ints = Integrals(prob)
    lay = prob.layout()
    no, nv = prob.n_int, prob.n_v
    Lf = prob.L_full()

    def eri(p, q, r, s):
        return float(np.dot(Lf[:, p, q], Lf[:, r, s]))

    xv = np.zeros(lay.n_csf)
    xv[lay.ref_off] = 1.0
    s_v = sigma(prob, xv, ints)
    big = 0
    for i in range(no):
        for a in range(nv):
            A = no + a
            g = prob.h[i, A] + eri(i, A, i, i)
            g += sum(2.0 * eri(i, A, k, k) - eri(i, k, k, A) for k in range(no) if k != i)
            ref = px[0, 0] * g
            k = lay.single_off + (no - 1 - i) * nv + a
            x = np.zeros(lay.n_csf)
            x[k] = 1.0
            assert abs(sigma(prob, x, ints)[0] - ref) < 1e-12 * max(1.0, abs(ref))
            assert abs(s_v[k] - ref) < 1e-12 * max(1.0, abs(ref))
            big += abs(ref) > 1e-2
    assert big > no * nv // 2
