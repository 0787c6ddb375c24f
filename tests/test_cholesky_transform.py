"""AO -> MO transformation of Cholesky vectors: the oracle restatement
(oracle/cholesky.py) pinned against an independent dense transformation on
the CPU, and the GPU entry point (tci_cholesky_ao_to_mo) against the oracle."""
import numpy as np
import pytest

from oracle.cholesky import ao_to_mo


def _case(seed, n_ao=13, n_mo=9, n_chol=7, keep=0.7):
    rng = np.random.default_rng(seed)
    pairs = [(m, n) for m in range(n_ao) for n in range(m + 1) if rng.random() < keep or m == n]
    pairs = [p[::-1] if rng.random() < 0.5 else p for p in pairs]  # either order in the list
    vec = rng.standard_normal((n_chol, len(pairs)))
    C = rng.standard_normal((n_ao, n_mo))
    return np.asarray(pairs, np.int32), vec, C


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_against_dense_transformation(seed):
    """Records equal C^T L^J C of the dense symmetric L^J built from the pair
    list, and the MO integrals sum_J L_pq L_rs equal the AO integrals
    transformed index by index (a transposed C or a lost mu != nu term fails)."""
    pairs, vec, C = _case(seed)
    n_chol, n_ao, n_mo = vec.shape[0], C.shape[0], C.shape[1]
    L = np.zeros((n_chol, n_ao, n_ao))
    for k, (m, n) in enumerate(pairs):
        L[:, m, n] = vec[:, k]
        L[:, n, m] = vec[:, k]
    dense = np.einsum("ap,Jab,bq->Jpq", C, L, C)
    rec, nrm = ao_to_mo(pairs, vec, C)
    for p in range(n_mo):
        for q in range(p + 1):
            np.testing.assert_allclose(rec[p * (p + 1) // 2 + q], dense[:, q, p], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(nrm, np.linalg.norm(rec, axis=1), rtol=1e-14)
    ao_eri = np.einsum("Jab,Jcd->abcd", L, L)
    mo_eri = np.einsum("ap,bq,cr,ds,abcd->pqrs", C, C, C, C, ao_eri)
    full = np.zeros((n_chol, n_mo, n_mo))
    for p in range(n_mo):
        for q in range(p + 1):
            full[:, p, q] = full[:, q, p] = rec[p * (p + 1) // 2 + q]
    np.testing.assert_allclose(np.einsum("Jpq,Jrs->pqrs", full, full), mo_eri, rtol=1e-10, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(13, 9, 7), (150, 140, 96), (40, 40, 1)])
def test_gpu_transform_matches_oracle(shape):
    import tigerci_amd as tci
    n_ao, n_mo, n_chol = shape
    pairs, vec, C = _case(5 + n_ao, n_ao, n_mo, n_chol, keep=0.6)
    print("case: AO->MO transform", shape, len(pairs), flush=True)
    rec, nrm = tci.cholesky_ao_to_mo(pairs, vec, C)
    ref, rnrm = ao_to_mo(pairs, vec, C)
    assert np.abs(rec - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(nrm, rnrm, rtol=1e-12)


@pytest.mark.gpu
def test_gpu_transform_rejects_duplicate_pairs():
    import tigerci_amd as tci
    with pytest.raises(tci.TigerSigmaError, match="listed twice"):
        tci.cholesky_ao_to_mo(np.asarray([[0, 1], [1, 0]], np.int32), np.ones((2, 2)), np.eye(2))
