"""The algebra of the resident-(ab|cd) ladder (tigerci_amd/csrc/ladder_vv.hip),
restated in numpy with the kernels' own index formulas, against the
Cholesky-direct form R_P = sum_J L^J[D,D] C_P L^J[D,D].  CPU only: it pins the
packing, the symmetric/antisymmetric split and the signs of the gather before
any of it runs on the device (the device path itself is checked against the
oracle in the GPU tests)."""
import numpy as np
import pytest


def tri(a):
    return a * (a + 1) // 2


def untri(p):
    c = int((np.sqrt(8.0 * p + 1.0) - 1.0) * 0.5)
    while tri(c + 1) <= p:
        c += 1
    while tri(c) > p:
        c -= 1
    return c, p - tri(c)


def resident_ladder(L, blocks, doms, ij):
    """R blocks from V+- = (ac|bd) +- (ad|bc) and the X+- rows, as the kernels
    build them (scatter over r >= s, gather over all (r, s))."""
    n_v = L.shape[1]
    W = np.einsum("jac,jbd->acbd", L, L)  # (ac|bd)
    np_, nm = tri(n_v), tri(n_v - 1)
    Vp = np.zeros((np_, np_))
    Vm = np.zeros((nm, nm))
    for a in range(n_v):
        for b in range(a + 1):
            for col in range(np_):
                c, d = untri(col)
                x, y = W[a, c, b, d], W[a, d, b, c]
                Vp[tri(a) + b, col] = x + y
                if a > b and c > d:
                    Vm[tri(a - 1) + b, tri(c - 1) + d] = x - y
    colm, ncm = [], 0
    for (i, j) in ij:
        colm.append(ncm if i != j else -1)
        ncm += i != j
    Xp = np.zeros((len(blocks), np_))
    Xm = np.zeros((max(ncm, 1), nm))
    for P, (C, dom) in enumerate(zip(blocks, doms)):
        d = len(dom)
        for t in range(tri(d)):
            r, c = untri(t)
            gr, gc = dom[r], dom[c]
            hr, rh = C[r, c], C[c, r]
            if gr < gc:
                gr, gc, hr, rh = gc, gr, rh, hr
            Xp[P, tri(gr) + gc] = 0.5 * hr if r == c else 0.5 * (hr + rh)
            if colm[P] >= 0 and gr != gc:
                Xm[colm[P], tri(gr - 1) + gc] = 0.5 * (hr - rh)
    Sp, Sm = Xp @ Vp, Xm @ Vm
    out = []
    for P, dom in enumerate(doms):
        d = len(dom)
        R = np.zeros((d, d))
        for r in range(d):
            for c in range(d):
                gr, gc = dom[r], dom[c]
                hi, lo = max(gr, gc), min(gr, gc)
                v = Sp[P, tri(hi) + lo]
                if colm[P] >= 0 and hi != lo:
                    w = Sm[colm[P], tri(hi - 1) + lo]
                    v += w if gr > gc else -w
                R[r, c] = v
        out.append(R)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_resident_ladder_matches_cholesky_direct(seed):
    rng = np.random.default_rng(seed)
    n_v, n_chol = 9, 5
    L = rng.standard_normal((n_chol, n_v, n_v))
    L = L + L.transpose(0, 2, 1)
    ij = [(2, 1), (1, 1), (3, 0), (2, 2)]
    doms, blocks = [], []
    for (i, j) in ij:
        d = int(rng.integers(2, n_v + 1))
        dom = rng.permutation(n_v)[:d]  # unsorted global indices
        C = rng.standard_normal((d, d))
        if i == j:
            C = C + C.T  # diagonal walks carry symmetric blocks only
        doms.append(dom)
        blocks.append(C)
    got = resident_ladder(L, blocks, doms, ij)
    for C, dom, R in zip(blocks, doms, got):
        Ld = L[:, dom][:, :, dom]
        ref = sum(Ld[J] @ C @ Ld[J] for J in range(n_chol))
        np.testing.assert_allclose(R, ref, rtol=1e-12, atol=1e-12)


def test_untri_inverts_tri_over_large_range():
    for p in list(range(2000)) + [tri(18527) + 18527, tri(45149) + 3, 10**9]:
        c, d = untri(p)
        assert 0 <= d <= c and tri(c) + d == p
