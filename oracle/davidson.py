"""Davidson-Liu restatement (standard eigenproblem).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows F/davidson.f90:
  davidson_liu              :47-239  (maxit = 50*roots, per-root loop)
  expand_krylov_subspace    :478-494 (h(k,idx) = <c_k, Hc_idx>)
  build_residual_std_eig    :503-526 (r = sum_k y_k (Hc_k - E c_k))
  is_converged              :564-602 (norm < norm_tol, or |dE| < e_tol and it > 1,
                                      never in/after a restart iteration, stop at maxit)
  build_next_vector_std_eig :609-631 (r <- -r/(H_ii - E); 0 where H_ii - E == 0)
  build_wavefunction        :675-694 (x = sum y_k c_k, renormalised)
  orthonormalize_vector     :702-723 (MGS, renormalise after every projection,
                                      zero if norm < 1e-14)
  init_from_diag            F/initial_davidson_guess.f90:99-132 (unit vectors at the
                                      lowest diagonal elements)
Subspace collapse (F/davidson_restart.f90:26-130) is triggered by a vector budget
``max_vectors`` in place of the reference's assumed disk space; it keeps, per root,
the current and previous Ritz vectors (orthonormalised).
"""
from __future__ import annotations

import numpy as np


def davidson(matvec, diag: np.ndarray, roots: int = 1, energy_tol: float = 1e-6,
             norm_tol: float = 1e-6, maxit: int | None = None,
             max_vectors: int | None = None, log=None):
    n = diag.shape[0]
    maxit = maxit or 50 * roots
    order = np.argsort(diag, kind="stable")[:roots]
    V, S = [], []
    for r in range(roots):
        v = np.zeros(n)
        v[order[r]] = 1.0
        V.append(v)
    root2idx = list(range(roots))
    converged = [False] * roots
    energy = np.zeros(roots)
    h = np.zeros((0, 0))
    restarts = []
    history = []
    y = None
    for it in range(1, maxit + 1):
        for r in range(roots):
            if converged[r]:
                continue
            idx = root2idx[r]
            s = matvec(V[idx])
            while len(S) <= idx:
                S.append(None)
            S[idx] = s
            m = len(V)
            hn = np.zeros((m, m))
            hn[:h.shape[0], :h.shape[1]] = h
            for k in range(idx + 1):
                e = V[k] @ s
                hn[k, idx] = hn[idx, k] = e
            h = hn
        m = len(V)
        w, Y = np.linalg.eigh(h[:m, :m])
        dE = np.abs(w[:roots] - energy)
        energy = w[:roots].copy()
        prevY = y
        y = Y[:, :roots].copy()
        restarted = False
        if max_vectors is not None and m > 4 * roots and m >= max_vectors:
            V, S, h, y = _collapse(V, S, y, prevY, converged, roots)
            restarted = True
        restarts.append(restarted)
        m = len(V)
        for r in range(roots):
            if converged[r]:
                continue
            x = sum(y[k, r] * V[k] for k in range(m))
            x /= np.linalg.norm(x)
            res = sum(y[k, r] * (S[k] - energy[r] * V[k]) for k in range(m))
            rn = float(np.linalg.norm(res))
            history.append((it, r, float(energy[r]), rn, float(dE[r])))
            if log:
                log(it, r, energy[r], rn, dE[r])
            if _is_converged(dE[r], rn, norm_tol, energy_tol, it, restarts, maxit):
                converged[r] = True
                continue
            pre = diag - energy[r]
            nv = np.where(pre == 0.0, 0.0, -res / np.where(pre == 0.0, 1.0, pre))
            for k in range(len(V)):
                nv = nv - (nv @ V[k]) * V[k]
                nrm = np.linalg.norm(nv)
                nv = np.zeros(n) if abs(nrm) < 1e-14 else nv / nrm
            V.append(nv)
            root2idx[r] = len(V) - 1
        if all(converged):
            break
    X = np.stack([sum(y[k, r] * V[k] for k in range(len(y))) for r in range(roots)])
    X /= np.linalg.norm(X, axis=1)[:, None]
    return energy, X, history


def _is_converged(dE, norm, norm_tol, energy_tol, it, restarts, maxit):
    if restarts[it - 1] or (it > 1 and restarts[it - 2]):
        return False
    if norm < norm_tol:
        return True
    if abs(dE) < energy_tol:
        return it > 1
    return it == maxit


def _collapse(V, S, y, prevY, converged, roots):
    """Keep span{current, previous} Ritz vectors per root (davidson_restart.f90:76-130)."""
    m = len(V)
    newV, newS = [], []
    for r in range(roots):
        for coef in ([y[:, r]] + ([prevY[:, r]] if prevY is not None else [])):
            cc = np.zeros(m)
            cc[:len(coef)] = coef
            v = sum(cc[k] * V[k] for k in range(m))
            s = sum(cc[k] * S[k] for k in range(m))
            for q in range(len(newV)):
                o = newV[q] @ v
                v, s = v - o * newV[q], s - o * newS[q]
            nrm = np.linalg.norm(v)
            if nrm > 1e-8:
                newV.append(v / nrm)
                newS.append(s / nrm)
    k = len(newV)
    h = np.array([[newV[a] @ newS[b] for b in range(k)] for a in range(k)])
    h = 0.5 * (h + h.T)
    w, Y = np.linalg.eigh(h)
    return newV, newS, h, Y[:, :roots].copy()
