"""Seeded synthetic sigma workloads (the benchmark shapes of this package).

Every array follows the tci_problem layout of include/tigersigma.h:
  * eps_i ~ U[-1.5,-0.3] (internal), eps_a ~ U[0.5,5.0] (virtual)
  * L^J_pq = L^J_qp ~ N(0,1) exp(-J/80) l_scale, one record of n_chol values
    per orbital pair p>=q (record p(p+1)/2+q)
  * pair walks (i,j), i>=j, in the reference's CI-vector walk order (i
    descending, then j descending: (n-1,n-1),(n-1,n-2),...,(0,0)); full
    domains, or seeded random sorted subsets of size U[lo,hi] (drawn in the
    order (0,0),(1,0),(1,1),(2,0),...) with pairs below ``weak_below``
    dropped (singles only)
Independent numpy sub-streams [seed,0..3] make each piece reproducible alone;
L is drawn in row chunks so an FP32 copy never needs an FP64 intermediate of
the full size.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

WORKLOADS = {
    # name: (n_int, n_v, n_chol, domain_range, weak_below, chol dtype, l_scale)
    "sigma_10x100x400_full": (10, 100, 400, None, 0, np.float64, 1.0),
    "sigma_16x300x1200_full": (16, 300, 1200, None, 0, np.float64, 1.0),
    "local_40x1000x3000_dom40-400": (40, 1000, 3000, (40, 400), 80, np.float32, 1.0),
    # Davidson solves: the same shapes with L scaled to molecule-sized
    # integrals (diagonally dominant H, reference weight ~0.95)
    "davidson_16x300x1200_full": (16, 300, 1200, None, 0, np.float64, 0.03),
    "davidson_40x1000x3000_dom40-400": (40, 1000, 3000, (40, 400), 80, np.float32, 0.03),
    # the shape of a cc-pVDZ polyene C12H14 local MRSDCI (34 correlated
    # internal, 192 virtual orbitals), local pair domains of 40..192 virtuals,
    # pairs below 60 singles-only
    "davidson_34x192x1200_dom40-192": (34, 192, 1200, (40, 192), 60, np.float64, 0.01),
}


@dataclass
class Synthetic:
    n_int: int
    n_v: int
    n_chol: int
    eps: np.ndarray
    L: np.ndarray        # (npair_all, n_chol)
    pairs: np.ndarray
    dom_off: np.ndarray
    dom_idx: np.ndarray


def make_synthetic(n_int: int, n_v: int, n_chol: int, seed: int, domain_range=None, weak_below: int = 0,
                   decay: float = 80.0, dtype=np.float64, chunk_rows: int = 1 << 14,
                   l_scale: float = 1.0) -> Synthetic:
    n_orb = n_int + n_v
    rng_e = np.random.default_rng([seed, 0])
    eps = np.concatenate([rng_e.uniform(-1.5, -0.3, n_int), rng_e.uniform(0.5, 5.0, n_v)])
    npair = n_orb * (n_orb + 1) // 2
    rng_l = np.random.default_rng([seed, 1])
    scale = np.exp(-np.arange(n_chol) / decay) * l_scale
    L = np.empty((npair, n_chol), dtype=dtype)
    for r0 in range(0, npair, chunk_rows):
        r1 = min(npair, r0 + chunk_rows)
        L[r0:r1] = rng_l.standard_normal((r1 - r0, n_chol)) * scale[None, :]
    pairs, doms = [], []
    rng_d = np.random.default_rng([seed, 2])
    for i in range(n_int):
        for j in range(i + 1):
            if domain_range is None:
                dom = np.arange(n_v, dtype=np.int32)
            else:
                size = min(int(rng_d.integers(domain_range[0], domain_range[1] + 1)), n_v)
                if size < weak_below:
                    continue
                dom = np.sort(rng_d.choice(n_v, size, replace=False)).astype(np.int32)
            pairs.append((i, j))
            doms.append(dom)
    order = sorted(range(len(pairs)), key=lambda k: (-pairs[k][0], -pairs[k][1]))
    pairs = [pairs[k] for k in order]
    doms = [doms[k] for k in order]
    dom_off = np.zeros(len(doms) + 1, np.int64)
    dom_off[1:] = np.cumsum([len(d) for d in doms])
    dom_idx = (np.concatenate(doms) if doms else np.zeros(0)).astype(np.int32)
    return Synthetic(n_int, n_v, n_chol, eps, L, np.array(pairs, np.int32).reshape(-1, 2), dom_off, dom_idx)


def workload(name: str, seed: int = 1234, l_scale: float | None = None) -> Synthetic:
    n_int, n_v, n_chol, dr, weak, dt, ls = WORKLOADS[name]
    ls = ls if l_scale is None else l_scale
    return make_synthetic(n_int, n_v, n_chol, seed, domain_range=dr, weak_below=weak, dtype=dt, l_scale=ls)


def random_vector(n: int, seed: int) -> np.ndarray:
    c = np.random.default_rng([seed, 3]).standard_normal(n)
    return c / np.linalg.norm(c)
