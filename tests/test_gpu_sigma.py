"""GPU parity of the HIP sigma path against the oracle (FP64: <= 1e-10 relative 2-norm)."""
import os
import sys

import numpy as np
import pytest

import tigerci_amd as tci
from oracle.problem import Problem, make_problem, random_ci_vector
from oracle.sigma import Integrals, sigma as oracle_sigma, diagonal as oracle_diagonal

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _gpu_sigma(plan, x, **kw):
    c = tci.DeviceArray.from_numpy(x)
    out = tci.DeviceArray(plan.n_csf)
    plan.sigma(c, out, **kw)
    return out.numpy()


def test_mfma_lane_layouts():
    _say("case: mfma self-test")
    tci.selftest_mfma()


@pytest.mark.parametrize("name", ["full_n3_v12", "local_n4_v30"])
def test_sigma_golden(name):
    _say("case: golden", name)
    z = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"))
    prob = Problem(int(z["n_int"]), int(z["n_v"]), int(z["n_chol"]), z["eps"], z["L"], z["pairs"],
                   z["dom_off"], z["dom_idx"])
    plan = tci.Plan.from_problem(prob)
    for x, s in zip(z["x"], z["sigma"]):
        assert _rel(_gpu_sigma(plan, x), s) < TOL
    assert _rel(plan.diagonal(), z["diag"]) < 1e-12


@pytest.mark.parametrize("args", [
    dict(n_int=2, n_v=5, n_chol=7, seed=1),
    dict(n_int=5, n_v=40, n_chol=60, seed=2),
    dict(n_int=6, n_v=70, n_chol=50, seed=3, domain_range=(10, 70), weak_below=16),
    dict(n_int=3, n_v=300, n_chol=30, seed=4, domain_range=(200, 300)),
])
def test_sigma_random(args):
    _say("case: random", args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    assert plan.n_csf == prob.layout().n_csf
    for k in range(2):
        x = random_ci_vector(prob, 50 + k)
        assert _rel(_gpu_sigma(plan, x), oracle_sigma(prob, x, ints)) < TOL


def test_sigma_unit_vectors_and_flags():
    _say("case: unit vectors / flags")
    prob = make_problem(3, 8, 9, 5, domain_range=(3, 8), weak_below=4)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    n = plan.n_csf
    rng = np.random.default_rng(0)
    for k in rng.choice(n, 12, replace=False).tolist() + [0, 1, n - 1]:
        e = np.zeros(n)
        e[k] = 1.0
        ref = oracle_sigma(prob, e, ints)
        assert np.abs(_gpu_sigma(plan, e) - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())
    x = random_ci_vector(prob, 9)
    c = tci.DeviceArray.from_numpy(x)
    base = rng.standard_normal(n)
    out = tci.DeviceArray.from_numpy(base)
    plan.sigma(c, out, accumulate=True, add_eref=True)
    expect = base + oracle_sigma(prob, x, ints) + plan.info.e_ref * x
    assert _rel(out.numpy(), expect) < TOL
    # host htimesc: sigmavec += H civec
    sv = np.ones(n)
    tci.htimesc(x, sv, plan)
    assert _rel(sv, 1.0 + oracle_sigma(prob, x, ints) + plan.info.e_ref * x) < TOL


def test_e_ref_matches_determinant_energy():
    _say("case: e_ref")
    from oracle.problem import reference_energy
    prob = make_problem(4, 6, 10, 8)
    plan = tci.Plan.from_problem(prob)
    assert abs(plan.info.e_ref - reference_energy(prob)) < 1e-10 * abs(reference_energy(prob))


def test_sigma_config1_size():
    """The 10/100/400 full-domain synthetic sigma (500k CSFs), one build."""
    _say("case: n_int=10 n_v=100 n_chol=400")
    prob = make_problem(10, 100, 400, 2024)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 1)
    ints = Integrals(prob)
    ref = oracle_sigma(prob, x, ints)
    got = _gpu_sigma(plan, x)
    assert _rel(got, ref) < TOL
    # the CPU baseline's sigma (the ladder in C with the reference's loop
    # structure) agrees with the GPU FP64 sigma as well
    from oracle.cpu_sigma import cpu_sigma
    cs = cpu_sigma(prob, x, ints)
    _say(f"  GPU vs oracle {_rel(got, ref):.2e}, GPU vs CPU baseline sigma {_rel(got, cs):.2e}")
    assert _rel(got, cs) < TOL


def test_dense_pair_path_matches_gather_path():
    """Full domains take the dense ordered-pair products for the ring and
    4-internal terms; TCI_DENSE_PAIRS=0 forces the per-walk gather GEMMs.
    Both must agree with the oracle (and so with each other)."""
    _say("case: dense pair products vs gather path")
    prob = make_problem(5, 37, 50, 21)
    ints = Integrals(prob)
    x = random_ci_vector(prob, 5)
    ref = oracle_sigma(prob, x, ints)
    dense = tci.Plan.from_problem(prob)
    os.environ["TCI_DENSE_PAIRS"] = "0"
    try:
        gather = tci.Plan.from_problem(prob)
    finally:
        del os.environ["TCI_DENSE_PAIRS"]
    a, b = _gpu_sigma(dense, x), _gpu_sigma(gather, x)
    assert _rel(a, ref) < TOL and _rel(b, ref) < TOL
    am, bm = _gpu_sigma(dense, x, precision="bf16x3"), _gpu_sigma(gather, x, precision="bf16x3")
    assert _rel(am, ref) < 5e-5 and _rel(bm, ref) < 5e-5


@pytest.mark.parametrize("env", [{}, {"TCI_LADDER_MERGED": "0"}, {"TCI_LADDER_MERGED": "0", "TCI_DENSE_PAIRS": "0"}])
def test_full_domain_variants(env):
    """Full domains run the ladder as single products over all walks; the
    per-walk programs (TCI_LADDER_MERGED=0) must give the same sigma, in FP64
    and bf16x3, including a domain size that is not a multiple of 4 or 8."""
    _say("case: full-domain variants", env)
    for args in [dict(n_int=4, n_v=37, n_chol=30, seed=31), dict(n_int=3, n_v=64, n_chol=20, seed=32)]:
        prob = make_problem(**args)
        ints = Integrals(prob)
        x = random_ci_vector(prob, 6)
        ref = oracle_sigma(prob, x, ints)
        os.environ.update(env)
        try:
            plan = tci.Plan.from_problem(prob)
        finally:
            for k in env:
                del os.environ[k]
        assert _rel(_gpu_sigma(plan, x), ref) < TOL
        assert _rel(_gpu_sigma(plan, x, precision="bf16x3"), ref) < 5e-5


@pytest.mark.parametrize("dense", ["0", "1"])
def test_local_domain_pair_paths(dense):
    """Local domains with weak pairs: the ring, 3-external and 4-internal terms
    as products over embedded blocks (TCI_DENSE_PAIRS=1) and as per-walk
    gathered programs (=0) both reproduce the oracle."""
    _say("case: local-domain pair paths, TCI_DENSE_PAIRS =", dense)
    prob = make_problem(6, 45, 40, 41, domain_range=(9, 45), weak_below=14)
    ints = Integrals(prob)
    os.environ["TCI_DENSE_PAIRS"] = dense
    try:
        plan = tci.Plan.from_problem(prob)
    finally:
        del os.environ["TCI_DENSE_PAIRS"]
    for k in range(2):
        x = random_ci_vector(prob, 60 + k)
        ref = oracle_sigma(prob, x, ints)
        assert _rel(_gpu_sigma(plan, x), ref) < TOL
        assert _rel(_gpu_sigma(plan, x, precision="bf16x3"), ref) < 5e-5


@pytest.mark.parametrize("z", ["0", "1"])
def test_dense_pair_z_forms(z):
    """Z = C~R Kov - CR Jbig + Yt Lov of the dense pair products directly
    (TCI_Z_LOWRANK=0) and as U Lov - CR Jbig with U = Yt + C~R Lov^T (=1, the
    default; FP64 and bf16x3, fp16x3 keeps the direct form): both reproduce the
    oracle with local domains and weak pairs, in every precision."""
    _say("case: dense pair Z form, TCI_Z_LOWRANK =", z)
    prob = make_problem(6, 45, 40, 43, domain_range=(9, 45), weak_below=14)
    ints = Integrals(prob)
    os.environ["TCI_DENSE_PAIRS"] = "1"
    os.environ["TCI_Z_LOWRANK"] = z
    try:
        plan = tci.Plan.from_problem(prob)
    finally:
        del os.environ["TCI_DENSE_PAIRS"]
        del os.environ["TCI_Z_LOWRANK"]
    for k in range(2):
        x = random_ci_vector(prob, 64 + k)
        ref = oracle_sigma(prob, x, ints)
        assert _rel(_gpu_sigma(plan, x), ref) < TOL
        assert _rel(_gpu_sigma(plan, x, precision="bf16x3"), ref) < 5e-5
        assert _rel(_gpu_sigma(plan, x, precision="fp16x3"), ref) < 2e-6
        assert _rel(_gpu_sigma(plan, x, precision="bf16x3"), ref) < 5e-5  # back after the switch


@pytest.mark.parametrize("args", [
    dict(n_int=34, n_v=6, n_chol=10, seed=81),
    dict(n_int=34, n_v=8, n_chol=12, seed=83, domain_range=(3, 8), weak_below=2),
    dict(n_int=2, n_v=150, n_chol=20, seed=82),
])
def test_sigma_pair_block_strips(args):
    """The (ia|jb) and (ki|bj) terms of the pair blocks run in 16-row strips
    over 128 columns and 32 internal orbitals at a time: n_int = 34 takes two
    passes over k (also with local domains and weak pairs), d = 150 two
    column passes.  FP64 and bf16x3 against the oracle."""
    _say("case: pair-block strips", args)
    prob = make_problem(**args)
    ints = Integrals(prob)
    plan = tci.Plan.from_problem(prob)
    x = random_ci_vector(prob, 90)
    ref = oracle_sigma(prob, x, ints)
    assert _rel(_gpu_sigma(plan, x), ref) < TOL
    assert _rel(_gpu_sigma(plan, x, precision="bf16x3"), ref) < 5e-5


@pytest.mark.parametrize("waves", ["256", "256w8", "256w16", "128x192d2", "128d2"])
@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_split_products_on_256_tiles(precision, waves):
    """Every split pair product (U, W, Z, Z3, the 4-internal product) on the
    256 x 256 tiles of the 32 x 32-MFMA kernel (TCI_PAIRS_TILE=256, or 256w8
    and 256w16 for its eight- and sixteen-wave forms) or on 128 x 192 tiles
    with two direct-load stages (128x192d2), and the resident ladder's
    product on them too
    (TCI_VV_TILE): sigma against the
    oracle, with outputs that are partial tiles (n_v = 45, 6 internal), W's
    blocked columns and U's accumulation."""
    _say("case: split products on 256 tiles", precision)
    prob = make_problem(6, 45, 40, 47, domain_range=(9, 45), weak_below=14)
    ints = Integrals(prob)
    env = {"TCI_DENSE_PAIRS": "1", "TCI_PAIRS_TILE": waves, "TCI_VV_TILE": waves, "TCI_LADDER_VV": "1",
           "TCI_AUTOTUNE": "0"}
    os.environ.update(env)
    try:
        plan = tci.Plan.from_problem(prob)
        for k in range(2):
            x = random_ci_vector(prob, 66 + k)
            ref = oracle_sigma(prob, x, ints)
            assert _rel(_gpu_sigma(plan, x, precision=precision), ref) < (5e-5 if precision == "bf16x3" else 2e-6)
        assert "-> resident" in plan.tuning()
    finally:
        for k in env:
            del os.environ[k]
