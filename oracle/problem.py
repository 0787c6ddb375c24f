"""Synthetic closed-shell MRSDCI problem and the CSF layout of the CI vector.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Orbitals: internal (doubly occupied in the closed-shell reference) 0..n_int-1,
then virtual n_int..n_int+n_v-1.  Virtual indices inside domains are 0-based
over the virtual space only.

Cholesky vectors follow the reference MO-integral record layout: one record
of ``n_chol`` doubles per orbital pair p>=q (C/ThreeIndexWriter.cpp:18-35,
cho_vectors(numcho, npair) in F/cholesky_structs.f90:35-39), pair index
p(p+1)/2+q in 0-based form (the reference's 1-based index2(p,q) =
max(max-1)/2+min of F/utilities_mod.f90:199 shifted by one).

CI vector layout (F/graph_mod.f90:1041-1092, 1430-1450) for the closed-shell
single reference, walks in the reference's lexical walk order (oracle/drt.py):
  [valence]  1 coefficient (the reference determinant, fsn(0)=1)
  [N-1]      per internal hole i, holes in descending order (i = n_int-1
             first): n_v coefficients (fsn(2)=1 spin function, virt-minor;
             singles use all virtuals)
  [N-2]      per strong pair walk P=(i,j), i>=j, in the order of ``pairs``
             (the generators emit the reference's order: i descending, then
             j descending, weak pairs left out):
               aa block: d_P coefficients           (fsn(s)=1)
               ab block: nsp * d_P(d_P-1)/2          (nsp=fsn(s+2): 1 for i==j,
                                                      2 for i>j: singlet-coupled
                                                      block first, then triplet)
             packed ab order (2,1),(3,1),(3,2),... over domain positions, as
             written by scepper (F/scepmaker_mod.f90:84-96).
Weak pairs (domain smaller than the weak threshold) carry no doubles
(WP removal, F/locist_mod.f90:204-240).
"""
from __future__ import annotations

import dataclasses
import numpy as np


def pair_index(p: int, q: int) -> int:
    """0-based packed index of orbital pair (p,q): max(max+1)/2+min."""
    if p < q:
        p, q = q, p
    return p * (p + 1) // 2 + q


@dataclasses.dataclass
class Problem:
    n_int: int
    n_v: int
    n_chol: int
    eps: np.ndarray            # (n_orb,) orbital energies, internal first
    L: np.ndarray              # (npair_all, n_chol) float64, pair-major records
    pairs: np.ndarray          # (n_pairs, 2) int32 strong pairs (i>=j), in CI-vector order
    dom_off: np.ndarray        # (n_pairs+1,) int64 offsets into dom_idx
    dom_idx: np.ndarray        # (sum d_P,) int32 sorted virtual indices
    # one-electron integrals h_pq (n_orb, n_orb), symmetric; None: the core
    # Hamiltonian whose closed-shell Fock matrix is diag(eps)
    h: np.ndarray | None = None

    # ---- derived -------------------------------------------------------
    @property
    def n_orb(self) -> int:
        return self.n_int + self.n_v

    @property
    def n_pairs(self) -> int:
        return int(self.pairs.shape[0])

    def domain(self, P: int) -> np.ndarray:
        return self.dom_idx[self.dom_off[P]:self.dom_off[P + 1]]

    def dsize(self, P: int) -> int:
        return int(self.dom_off[P + 1] - self.dom_off[P])

    def layout(self) -> "Layout":
        return Layout.build(self)

    def L_full(self) -> np.ndarray:
        """(n_chol, n_orb, n_orb) symmetric L^J_pq (test helper)."""
        n = self.n_orb
        iu = np.tril_indices(n)
        idx = iu[0] * (iu[0] + 1) // 2 + iu[1]
        out = np.zeros((self.n_chol, n, n))
        out[:, iu[0], iu[1]] = self.L[idx].T
        out[:, iu[1], iu[0]] = self.L[idx].T
        return out


@dataclasses.dataclass
class Layout:
    """Offsets of every block of the CI vector (0-based)."""
    n_csf: int
    ref_off: int
    single_off: int                  # singles of hole i at single_off + (n_int-1-i)*n_v
    aa_off: np.ndarray               # (n_pairs,)
    ab_off: np.ndarray               # (n_pairs,)
    nsp: np.ndarray                  # (n_pairs,) 1 or 2
    dsz: np.ndarray                  # (n_pairs,)

    @staticmethod
    def build(prob: Problem) -> "Layout":
        n_pairs = prob.n_pairs
        aa_off = np.zeros(n_pairs, np.int64)
        ab_off = np.zeros(n_pairs, np.int64)
        nsp = np.zeros(n_pairs, np.int64)
        dsz = np.zeros(n_pairs, np.int64)
        off = 1 + prob.n_int * prob.n_v
        for P in range(n_pairs):
            i, j = prob.pairs[P]
            d = prob.dsize(P)
            dsz[P] = d
            nsp[P] = 1 if i == j else 2
            aa_off[P] = off
            off += d
            ab_off[P] = off
            off += nsp[P] * (d * (d - 1) // 2)
        return Layout(int(off), 0, 1, aa_off, ab_off, nsp, dsz)


def make_problem(n_int: int, n_v: int, n_chol: int, seed: int,
                 domain_range: tuple[int, int] | None = None,
                 weak_below: int = 0,
                 decay: float = 80.0, l_scale: float = 1.0,
                 h_offdiag: float = 0.0) -> Problem:
    """Seeded synthetic problem.

    * eps_i ~ U[-1.5,-0.3], eps_a ~ U[0.5,5.0]
    * L^J_pq = L^J_qp ~ N(0,1) * exp(-J/decay) * l_scale over all orbitals
      (l_scale < 1 gives molecule-sized two-electron integrals and a
      diagonally dominant H; the same draws for every l_scale)
    * full pair domains unless ``domain_range=(lo,hi)``: then each pair's
      domain is a sorted random subset of size U[lo,hi] (inclusive) and pairs
      with size < ``weak_below`` carry no doubles.
    * ``h_offdiag > 0``: explicit one-electron integrals h = h_canonical + V,
      V symmetric with N(0,1) * h_offdiag entries (stream [seed,4]), so the
      closed-shell Fock matrix diag(eps) + V has off-diagonal occupied-occupied,
      occupied-virtual and virtual-virtual blocks (localized orbitals).
    Independent sub-streams keep each piece reproducible on its own.
    """
    n_orb = n_int + n_v
    rng_e = np.random.default_rng([seed, 0])
    eps = np.concatenate([rng_e.uniform(-1.5, -0.3, n_int),
                          rng_e.uniform(0.5, 5.0, n_v)])
    npair_all = n_orb * (n_orb + 1) // 2
    rng_l = np.random.default_rng([seed, 1])
    scale = np.exp(-np.arange(n_chol) / decay) * l_scale
    L = rng_l.standard_normal((npair_all, n_chol)) * scale[None, :]
    pairs, doms = [], []
    rng_d = np.random.default_rng([seed, 2])
    for i in range(n_int):
        for j in range(i + 1):
            if domain_range is None:
                dom = np.arange(n_v, dtype=np.int32)
            else:
                size = int(rng_d.integers(domain_range[0], domain_range[1] + 1))
                size = min(size, n_v)
                if size < weak_below:
                    continue
                dom = np.sort(rng_d.choice(n_v, size, replace=False)).astype(np.int32)
            pairs.append((i, j))
            doms.append(dom)
    # the reference's walk order (i descending, then j descending); domains
    # are drawn in ascending order above, so a pair keeps its domain
    order = sorted(range(len(pairs)), key=lambda k: (-pairs[k][0], -pairs[k][1]))
    pairs = [pairs[k] for k in order]
    doms = [doms[k] for k in order]
    pairs_a = np.array(pairs, dtype=np.int32).reshape(-1, 2)
    dom_off = np.zeros(len(doms) + 1, np.int64)
    dom_off[1:] = np.cumsum([len(d) for d in doms])
    dom_idx = (np.concatenate(doms) if doms else np.zeros(0)).astype(np.int32)
    prob = Problem(n_int, n_v, n_chol, eps, L, pairs_a, dom_off, dom_idx)
    if h_offdiag > 0.0:
        V = np.random.default_rng([seed, 4]).standard_normal((n_orb, n_orb)) * h_offdiag
        prob.h = core_hamiltonian(prob) + 0.5 * (V + V.T)
    return prob


def packed_lower(M: np.ndarray) -> np.ndarray:
    """Symmetric matrix -> packed lower triangle, element (p,q), p>=q, at
    p(p+1)/2+q: the layout of the reference's one-electron integral array
    pp(num_orbitals(num_orbitals+1)/2) (F/get_integrals_mod.f90 getpp,
    index2 of F/utilities_mod.f90 shifted to 0-based)."""
    r, c = np.tril_indices(M.shape[0])
    out = np.zeros(M.shape[0] * (M.shape[0] + 1) // 2)
    out[r * (r + 1) // 2 + c] = M[r, c]
    return out


def random_ci_vector(prob: Problem, seed: int) -> np.ndarray:
    """N(0,1) CI vector over every CSF, normalised to 1."""
    n = prob.layout().n_csf
    c = np.random.default_rng([seed, 3]).standard_normal(n)
    return c / np.linalg.norm(c)


# ---- integral blocks from L (pq|rs) = sum_J L^J_pq L^J_rs ------------------
# (F/make_all_integrals.f90:174-282 makeOneIntegral; F/make_abcd.f90:498-526)

def eri_full(prob: Problem) -> np.ndarray:
    """(n_orb,)*4 chemist-notation ERIs, tiny problems only."""
    Lf = prob.L_full()
    n = prob.n_orb
    M = Lf.reshape(prob.n_chol, n * n)
    return (M.T @ M).reshape(n, n, n, n)


def closed_shell_g(prob: Problem) -> np.ndarray:
    """G_pq = sum_k [2(pq|kk) - (pk|kq)] over the doubly occupied internal
    orbitals k, from the Cholesky vectors (no four-index array)."""
    Lf = prob.L_full()
    o = slice(0, prob.n_int)
    dJ = np.einsum('Jkk->J', Lf[:, o, o])
    Lo = Lf[:, :, o]                                   # (J, p, k)
    return 2.0 * np.einsum('J,Jpq->pq', dJ, Lf) - np.einsum('Jpk,Jqk->pq', Lo, Lo)


def core_hamiltonian(prob: Problem, eri: np.ndarray | None = None) -> np.ndarray:
    """h_pq: ``prob.h`` when given, otherwise the one whose closed-shell Fock
    matrix is diag(eps): f_pq = h_pq + G_pq  ->  h = diag(eps) - G."""
    if prob.h is not None:
        return np.array(prob.h, dtype=np.float64)
    if eri is None:
        return np.diag(prob.eps) - closed_shell_g(prob)
    o = slice(0, prob.n_int)
    G = 2.0 * np.einsum('pqkk->pq', eri[:, :, o, o]) - np.einsum('pkkq->pq', eri[:, o, o, :])
    return np.diag(prob.eps) - G


def fock(prob: Problem) -> np.ndarray:
    """Closed-shell Fock matrix f = h + G (diag(eps) when no h is given)."""
    if prob.h is None:
        return np.diag(prob.eps).astype(np.float64)
    return np.array(prob.h, dtype=np.float64) + closed_shell_g(prob)


def reference_energy(prob: Problem, eri: np.ndarray | None = None) -> float:
    """E_ref = sum_i (h_ii + f_ii) of the closed-shell reference."""
    h = core_hamiltonian(prob, eri)
    f = fock(prob)
    o = np.arange(prob.n_int)
    return float(np.sum(h[o, o] + f[o, o]))
