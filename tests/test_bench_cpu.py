"""bench.py contract on the CPU: the one-GPU guard and the roofline arithmetic."""
import importlib.util
import json
import os
import subprocess
import sys
from types import SimpleNamespace

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def _syn(n_v, doms, n_chol, f32=False):
    off = np.concatenate([[0], np.cumsum(doms)])
    L = np.zeros(1, dtype=np.float32 if f32 else np.float64)
    return SimpleNamespace(n_v=n_v, n_chol=n_chol, dom_off=off, L=L)


def test_ladder_minimum_bytes():
    b = _bench()
    # two local walks (d = 10, 20) and one full walk (d = n_v = 30), 7 vectors
    syn = _syn(30, [10, 20, 30], 7)
    d2 = 100 + 400 + 900
    assert b.ladder_alg_bytes(syn, "bf16x3") == 7 * (2 * 4 * d2 + 2 * 4 * (100 + 400) + 2 * 4 * 900)
    assert b.ladder_alg_bytes(syn, "fp64") == 7 * (2 * 8 * d2 + 2 * 8 * (100 + 400) + 2 * 8 * 900)
    syn32 = _syn(30, [10, 20], 7, f32=True)
    assert b.ladder_alg_bytes(syn32, "fp64") == 7 * (2 * 8 * 500 + 2 * 4 * 500)
    # fused walks (split only) read L^J[D,D] once and move no T: walks of 70
    # and 130 fused, 50 two-phase, one full walk of n_v = 300
    syn = _syn(300, [50, 70, 130, 300], 3)
    fz = [(65, 128), (129, 192)]
    assert b.ladder_alg_bytes(syn, "bf16x3", fz) == 3 * (2 * 4 * (2500 + 90000) + 2 * 4 * 2500
                                                          + 4 * (4900 + 16900) + 2 * 4 * 90000)
    assert b.ladder_alg_bytes(syn, "fp64", fz) == b.ladder_alg_bytes(syn, "fp64")


def test_fused_ranges_from_tuning_record():
    b = _bench()
    t = ("split ladder: ... whole two-phase:51.3ms fused512x:60.9ms -> two-phase; 65-128 two-phase:51.2ms "
         "fused128-512:43.5ms -> fused128-512; order 0 1")
    assert b.fused_ranges(t) == [(65, 128)]
    t2 = "split ladder: ... whole two-phase:51.3ms fused512x:40.9ms -> fused512x; 65-128 two-phase:40ms -> two-phase"
    assert b.fused_ranges(t2) == [(129, 192)]
    assert b.fused_ranges(t2.replace("-> two-phase", "-> fused128-1024x")) == [(65, 128), (129, 192)]
    assert b.fused_ranges("fp64 ladder: -> 8192/64/p1 128") == []
    t3 = t + "; 129-192 current:43.4ms fused192-512:30.1ms -> fused192-512"
    assert b.fused_ranges(t3) == [(65, 128), (129, 192)]


def test_kernel_roofline_and_ladder_section():
    """A kernel's roofline is its algorithmic flops per launch over its mean
    launch time against the dense MFMA peak of its precision; the ladder
    section is F_lad over the section time, with the L^J[D,D]-once byte floor."""
    b = _bench()
    k = dict(name="xgemm_kernel<128,192>", launches=6, ms=18.0, flops=1.5e12, launches_per_sigma=6.0)
    r = b.kernel_roofline(k, "bf16x3", 3.0e9)
    t = 18.0 / 6 / 1e3
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    assert abs(r["achieved"] - 1.5e12 / t / 1e12) < 1e-9 and abs(r["frac"] - 1.5e12 / t / 2.5e15) < 1e-12
    assert abs(r["mfma_busy_frac"] - 3 * r["frac"]) < 1e-12 and abs(r["traffic_gbs"] - 3.0e9 / t / 1e9) < 1e-9
    r64 = b.kernel_roofline(dict(k, name="dgemm_kernel<64>"), "fp64", None)
    assert abs(r64["frac"] - 1.5e12 / t / 78.6e12) < 1e-12 and "mfma_busy_frac" not in r64
    syn = _syn(192, [150] * 100, 1000)
    info = SimpleNamespace(flops_ladder=4.0 * 100 * 150 ** 3 * 1000)
    sec = b.ladder_section(info, syn, "bf16x3", 0.05)
    assert abs(sec["frac_mfma"] - info.flops_ladder / 0.05 / 2.5e15) < 1e-12
    assert sec["byte_floor"] == 4.0 * 100 * 150 ** 2 * 1000           # L^J[D,D] once, 4 bytes split
    assert b.ladder_floor_bytes(syn, "fp64") == 8.0 * 100 * 150 ** 2 * 1000
    assert sec["schedule_bytes"] > sec["byte_floor"]                    # T round trip and two reads of L


def test_resident_ladder_record_and_section():
    """The resident-(ab|cd) ladder is read from the tuning record per
    precision; its section counts the packed products' flops, 2 n_p^2 n_walks
    + 2 n_m^2 n_walks(i != j), and V+- read once as the byte floor."""
    b = _bench()
    t = ("fp64 ladder: -> 8192/64/p1 128\n"
         "fp64 ladder (ab|cd) resident: 6.8e+11 flops against 6.2e+12 Cholesky-direct; 128/1:12.0ms -> 128/1 "
         "order 1; resident 12.1ms, Cholesky-direct 200.3ms -> resident\n"
         "bf16x3 ladder (ab|cd) resident: 128x192/1:3.1ms -> 128x192/1 order 1; resident 3.0ms, "
         "Cholesky-direct 2.0ms -> Cholesky-direct\n")
    assert b.resident_ladder(t, "fp64") and not b.resident_ladder(t, "bf16x3")
    assert not b.resident_ladder(t, "fp16x3")
    syn = _syn(10, [10, 6, 10], 7)
    syn.pairs = np.array([[2, 1], [1, 1], [3, 0]], dtype=np.int32)
    n_p, n_m = 55, 45
    sec = b.ladder_section(SimpleNamespace(flops_ladder=1.0), syn, "fp64", 0.01, resident=True)
    assert sec["algorithm"] == "resident (ab|cd)"
    assert sec["algorithmic_flops"] == 2.0 * n_p ** 2 * 3 + 2.0 * n_m ** 2 * 2
    assert sec["byte_floor"] >= 8.0 * (n_p ** 2 + n_m ** 2)
    sx = b.ladder_section(SimpleNamespace(flops_ladder=1.0), syn, "bf16x3", 0.01, resident=True)
    assert sx["byte_floor"] < sec["byte_floor"]


def test_kernel_traffic_lookup_keeps_pipes_in_names():
    """Kernel names such as "fp64 ladder (ab|cd) product ..." hold the field
    separator; the lookup joins the middle fields back and takes the bytes
    from the last one.  Every machine-readable line of the committed file is
    found again under its own name."""
    b = _bench()
    path = os.path.join(ROOT, "profiles", "kernel_traffic.txt")
    lines = [l.rstrip("\n") for l in open(path) if l.strip() and not l.startswith("#")]
    assert lines
    for l in lines:
        f = l.split("|")
        assert b.kernel_traffic(f[0], f[1], "|".join(f[2:-1])) == float(f[-1])
    assert any("(ab|cd)" in l for l in lines)
    assert b.kernel_traffic("davidson_34x192x1200_dom40-192", "fp64", "no such kernel") is None
