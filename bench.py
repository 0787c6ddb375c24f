#!/usr/bin/env python3
"""Benchmark of the MI355X sigma / Davidson path.

Metric: Davidson time-to-|dE|<1e-6 Eh, mixed precision vs FP64 on one GPU,
with sigma TFLOP/s and the ladder kernel's roofline.  One step = one complete
Davidson-Liu solve (lowest --roots roots, from the init_from_diag guess) of
the synthetic workload with the split-precision ladder and FP64 refinement;
the FP64 solve of the same problem is timed once beside it.  --mode sigma
times single sigma builds instead.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--precision bf16x3|fp64] [--mode davidson|sigma] [--roots R]

Prints one JSON line.  Inputs are generated on the host and resident in HBM
before the timed region starts (plan creation is outside it).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

METRIC = "Davidson time-to-|ΔE|<1e-6 Eh: mixed vs FP64 GPU; sigma TFLOP/s, HBM GB/s"
PEAK = {"fp64": 78.6e12, "bf16x3": 2.5e15, "fp16x3": 2.5e15}   # dense MFMA, MI355X
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fused_ranges(tuning: str) -> list:
    """Domain-size ranges whose split ladder ran as a fused kernel, read from
    the plan's tuning record ("-> fused128-N" for 65-128, "-> fusedN" or
    "-> fused192-N" for 129-192)."""
    out = []
    if re.search(r"-> fused128-\d", tuning):
        out.append((65, 128))
    if re.search(r"-> fused\d+x?(?![\d-])", tuning) or re.search(r"-> fused192-\d", tuning):
        out.append((129, 192))
    return out


def ladder_alg_bytes(syn, precision: str, fused=()) -> float:
    """Minimum HBM bytes of one sigma's ladder.  Two-phase walks: per Cholesky
    vector, the intermediate T is written and read once (d^2 elements each
    way) and the L^J[D,D] block is read by both phases; walks with the full
    domain share L^J itself.  Walks in a fused domain range (split precision,
    T kept on chip) read their L^J[D,D] block once.  Elements are 4 bytes in
    split form (hi + lo) and 8 bytes in FP64 (L^J 4 bytes when stored in
    FP32)."""
    d = np.diff(np.asarray(syn.dom_off)).astype(np.float64)
    full = d == syn.n_v
    fz = np.zeros_like(full)
    if precision != "fp64":
        for lo, hi in fused:
            fz |= (d >= lo) & (d <= hi) & ~full
    e_t = 8.0 if precision == "fp64" else 4.0
    e_l = (4.0 if syn.L.dtype == np.float32 else 8.0) if precision == "fp64" else 4.0
    two = ~full & ~fz
    per_j = (2.0 * e_t * (d[~fz] ** 2).sum() + 2.0 * e_l * (d[two] ** 2).sum() + e_l * (d[fz] ** 2).sum()
             + (2.0 * e_l * syn.n_v ** 2 if full.any() else 0.0))
    return float(per_j * syn.n_chol)


def roofline(info, syn, precision: str, seconds: float, traffic, kernel: str, fused=()) -> dict:
    """Roofline of the ladder section: the binding roof is the larger of the
    MFMA time (algorithmic flops, x3 MFMAs for a split product) and the HBM
    time of the algorithm's minimum bytes; achieved is that quantity over the
    measured section time."""
    flops = info.flops_ladder
    mfma_flops = flops * (1 if precision == "fp64" else 3)
    alg_bytes = ladder_alg_bytes(syn, precision, fused)
    t_mfma, t_hbm = mfma_flops / PEAK[precision], alg_bytes / HBM_PEAK
    r = {"kernel": kernel, "avg_launch_ms": seconds * 1e3, "algorithmic_flops": flops, "algorithmic_bytes": alg_bytes,
         "traffic": traffic, "traffic_gbs": traffic / seconds / 1e9 if traffic else None,
         "tflops_algorithmic": flops / seconds / 1e12, "mfma_busy_frac": mfma_flops / seconds / PEAK[precision]}
    if t_hbm > t_mfma:
        ach = alg_bytes / seconds
        r.update({"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK})
    else:
        ach = flops / seconds  # algorithmic flops against the published dense peak of the precision
        r.update({"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK[precision] / 1e12, "unit": "TFLOP/s",
                  "frac": ach / PEAK[precision]})
    return r


def measured_traffic(workload: str, precision: str):
    """HBM bytes of one sigma's ladder section, measured with rocprofv3 --pmc
    (profiles/ladder_traffic.txt), or None for an unmeasured workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "ladder_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and not line.startswith("#") and f[0] == workload and f[1] == precision:
            return float(f[2])
    return None


def cpu_baseline(syn, plan, n_sigma_fp64: int, budget_s: float = 15.0) -> dict:
    """Time the oracle's dense ladder restatement on a bounded sample of pair
    blocks of the same workload, and scale to one FP64 Davidson solve
    (n_sigma_fp64 sigma builds, each plan.info.flops_sigma flops)."""
    from threadpoolctl import threadpool_info
    from oracle.scepper import scepper_p1, scepper_m1
    rng = np.random.default_rng(7)
    n_int = syn.n_int
    L = syn.L
    order = np.argsort(-np.diff(syn.dom_off))
    sample = list(order[::max(1, len(order) // 16)])
    flops = 0.0
    t0 = time.perf_counter()
    used = 0
    while time.perf_counter() - t0 < budget_s:
        P = sample[used % len(sample)]
        D = syn.dom_idx[syn.dom_off[P]:syn.dom_off[P + 1]]
        d = len(D)
        a = n_int + D
        rec = (np.maximum.outer(a, a) * (np.maximum.outer(a, a) + 1) // 2 + np.minimum.outer(a, a))
        Lp = np.ascontiguousarray(L[rec.ravel()].T.reshape(syn.n_chol, d, d), dtype=np.float64)
        m2 = d * (d - 1) // 2
        C = scepper_p1(rng.standard_normal(m2), rng.standard_normal(d), d) / 2 + \
            scepper_m1(rng.standard_normal(m2), d) / 3.4641016151377544
        (Lp @ C @ Lp).sum(axis=0)
        flops += 4.0 * d ** 3 * syn.n_chol
        used += 1
    dt = time.perf_counter() - t0
    rate = flops / dt
    threads = max([x.get("num_threads", 1) for x in threadpool_info() if x.get("user_api") == "blas"] or [1])
    sigma_s = plan.info.flops_sigma / rate
    return {"value": sigma_s * n_sigma_fp64, "unit": "s", "cores": int(threads), "kind": "port",
            "sample": f"oracle dense ladder (numpy BLAS, FP64), {used} pair-block ladders drawn from "
                      f"{len(sample)} of {len(order)} walks "
                      f"({flops:.3g} flop in {dt:.1f} s = {rate / 1e9:.1f} GFLOP/s), scaled by flops to "
                      f"{n_sigma_fp64} FP64 sigma builds of {plan.info.flops_sigma:.3g} flop",
            "sigma_seconds": sigma_s}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=None,
                    help="default: davidson_34x192x1200_dom40-192 (davidson), sigma_16x300x1200_full (sigma)")
    ap.add_argument("--roots", type=int, default=3)
    ap.add_argument("--l-scale", type=float, default=None, help="override the workload's Cholesky-vector scale")
    ap.add_argument("--ladder-launch", default="grouped", choices=["grouped", "per-pair"],
                    help="split ladder of local domains: grouped launches or one launch per walk (comparison)")
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--graph", dest="graph", action="store_true", default=True,
                    help="replay every sigma build from a captured HIP graph (default; ladder times then come "
                         "from separate eager profiled builds)")
    ap.add_argument("--no-graph", dest="graph", action="store_false", help="launch every sigma build eagerly")
    ap.add_argument("--mode", default="davidson", choices=["davidson", "sigma"])
    ap.add_argument("--switch-tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB device-to-device (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "the sigma path runs on one GPU: it is not sharded across GPUs "
                                    "(by pair walk, with one all-gather of c per sigma, is how it would be); "
                                    "run with --gpus 1"}))
        return 2

    if a.ladder_launch == "per-pair":
        os.environ["TCI_LADDER_PER_PAIR"] = "1"
    import tigerci_amd as tci
    from tigerci_amd.synthetic import workload, random_vector, WORKLOADS

    if a.workload is None:
        a.workload = "davidson_34x192x1200_dom40-192" if a.mode == "davidson" else "sigma_16x300x1200_full"
    t = time.perf_counter()
    syn = workload(a.workload, a.seed, l_scale=a.l_scale)
    log(f"[bench] generated {a.workload} in {time.perf_counter() - t:.1f} s")
    ladder_dtype = "f32" if syn.L.dtype == np.float32 else "f64"
    t = time.perf_counter()
    plan = tci.Plan.from_problem(syn, ladder_dtype=ladder_dtype)
    tci.synchronize()
    log(f"[bench] plan: n_csf={plan.n_csf} items={plan.info.n_ladder_items} "
        f"device {plan.info.device_bytes / 2**30:.1f} GiB, built in {time.perf_counter() - t:.1f} s")
    info = plan.info
    cfg = {"workload": a.workload, "ladder_launch": a.ladder_launch, "n_int": syn.n_int, "n_v": syn.n_v, "n_chol": syn.n_chol,
           "pairs": int(len(syn.pairs)), "n_csf": int(plan.n_csf), "L_storage": ladder_dtype,
           "ladder_flops_per_sigma": info.flops_ladder, "sigma_flops": info.flops_sigma}
    out = {"metric": METRIC, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "scaling": "weak",
           "vs_baseline": None,
           "data": f"synthetic (seeded N(0,1) exp(-J/80) x "
                   f"{a.l_scale if a.l_scale is not None else WORKLOADS[a.workload][6]} Cholesky vectors)"}

    def ladder_probe(prec, n):
        """ladder and sigma event times of n eager builds (graph replays time only the whole build)"""
        xp = tci.DeviceArray.from_numpy(random_vector(plan.n_csf, a.seed + 1))
        yp = tci.DeviceArray(plan.n_csf)
        plan.sigma(xp, yp, precision=prec)
        plan.profile(True)
        for _ in range(n):
            plan.sigma(xp, yp, precision=prec)
        r = plan.profile_read(prec)
        plan.profile(False)
        return r

    cfg["sigma_graph"] = a.graph
    cfg["autotune"] = os.environ.get("TCI_AUTOTUNE", "1") != "0"
    if a.mode == "sigma":
        x = tci.DeviceArray.from_numpy(random_vector(plan.n_csf, a.seed))
        y = tci.DeviceArray(plan.n_csf)
        for _ in range(a.warmup):
            plan.sigma(x, y, precision=a.precision, graph=a.graph)
        tci.synchronize()
        plan.profile(not a.graph)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            plan.sigma(x, y, precision=a.precision, graph=a.graph)
        tci.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        pr = ladder_probe(a.precision, a.steps) if a.graph else plan.profile_read(a.precision)
        cfg["schedule"] = plan.tuning().strip().splitlines()
        lad = pr["ladder_ms"] / max(pr["ladder_launches"], 1) / 1e3
        if a.precision != "fp64":  # precision study: this sigma against the FP64 one
            s_mx = y.numpy()
            plan.sigma(x, y, precision="fp64")
            s_64 = y.numpy()
            xv = x.numpy()
            out["sigma_check"] = {"sigma_rel_err": float(np.linalg.norm(s_mx - s_64) / np.linalg.norm(s_64)),
                                  "rayleigh_err_Eh": float(abs(xv @ s_mx - xv @ s_64))}
        out.update({"value": info.flops_sigma / dt / 1e12, "unit": "TFLOP/s", "higher_is_better": True,
                    "ms_per_step": dt * 1e3, "dtype": a.precision, "config": cfg,
                    "roofline": roofline(info, syn, a.precision, lad, measured_traffic(a.workload, a.precision),
                                         f"{a.precision} ladder (per sigma)", fused_ranges(plan.tuning()))})
        if a.copy_probe_mib:
            n = a.copy_probe_mib * (1 << 20) // 8
            src, dst = tci.DeviceArray(n), tci.DeviceArray(n)
            dst.copy_from(src)
            tci.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                dst.copy_from(src)
            tci.synchronize()
            out["copy_probe_bytes"] = 8 * n
            out["copy_probe_gbs"] = 2 * 8 * n * 5 / (time.perf_counter() - t0) / 1e9  # read + write
        print(json.dumps(out), flush=True)
        return 0

    # ---- sigma precision check on one random vector ---------------------------
    xr = random_vector(plan.n_csf, a.seed)
    xd = tci.DeviceArray.from_numpy(xr)
    yd = tci.DeviceArray(plan.n_csf)
    plan.sigma(xd, yd, precision="fp64")
    s64 = yd.numpy()
    plan.sigma(xd, yd, precision=a.precision)
    smx = yd.numpy()
    sig_err = {"sigma_rel_err": float(np.linalg.norm(smx - s64) / np.linalg.norm(s64)),
               "rayleigh_err_Eh": float(abs(xr @ smx - xr @ s64))}
    del xd, yd

    # ---- Davidson: mixed (timed steps) and FP64 reference -------------------
    def solve(prec):
        return plan.davidson(roots=a.roots, energy_tol=1e-6, norm_tol=1e-5, precision=prec,
                             switch_tol=a.switch_tol, graph=a.graph)

    for _ in range(a.warmup):
        solve(a.precision)
    tci.synchronize()
    plan.profile(True)
    t0 = time.perf_counter()
    res_m = None
    for s in range(a.steps):
        e_m, res_m = solve(a.precision)
        log(f"[bench] step {s}: {res_m} E={e_m}")
    tci.synchronize()
    t_m = (time.perf_counter() - t0) / a.steps
    pm = plan.profile_read(a.precision)
    plan.profile(True)
    t0 = time.perf_counter()
    e_64, res_64 = solve("fp64")
    tci.synchronize()
    t_64 = time.perf_counter() - t0
    p64 = plan.profile_read("fp64")
    log(f"[bench] fp64 solve: {res_64} E={e_64}")
    if a.graph:
        sig_ms = {k: v["sigma_ms"] / max(v["sigma_calls"], 1) for k, v in (("fp64", p64), ("mixed", pm))}
        p64 = dict(ladder_probe("fp64", 3), sigma_ms=p64["sigma_ms"], sigma_calls=p64["sigma_calls"])
        if a.precision != "fp64":
            pm = dict(ladder_probe(a.precision, 3), sigma_ms=pm["sigma_ms"], sigma_calls=pm["sigma_calls"])
        log(f"[bench] graph sigma ms {sig_ms}")
    lad = p64["ladder_ms"] / max(p64["ladder_launches"], 1) / 1e3
    sig64 = p64["sigma_ms"] / max(p64["sigma_calls"], 1) / 1e3
    out.update({
        "value": t_m, "unit": "s", "higher_is_better": False, "ms_per_step": t_m * 1e3,
        "dtype": "f64" if a.precision == "fp64" else f"{a.precision}+f64", "config": dict(cfg, roots=a.roots),
        "fp64_seconds": t_64, "speedup_vs_fp64": t_64 / t_m,
        "energies_mixed": [float(v) for v in e_m], "energies_fp64": [float(v) for v in e_64],
        "energy_diff_uEh": float(np.abs(e_m - e_64).max() * 1e6),
        "converged": {"mixed": res_m["converged"], "fp64": res_64["converged"]},
        "iterations": {"mixed": res_m["iterations"], "fp64": res_64["iterations"],
                       "sigma_mixed": res_m["sigma_mixed"], "sigma_fp64_in_mixed": res_m["sigma_fp64"],
                       "sigma_fp64": res_64["sigma_fp64"]},
        "sigma_check": sig_err,
        "sigma_tflops_fp64": info.flops_sigma / sig64 / 1e12,
    })
    out["config"]["schedule"] = plan.tuning().strip().splitlines()
    roof64 = roofline(info, syn, "fp64", lad, measured_traffic(a.workload, "fp64"),
                      "FP64 ladder (dgemm phases 1+2, per sigma)")
    roof64["launches"] = p64["ladder_launches"]
    if a.precision != "fp64" and pm["ladder_launches"]:
        lm = pm["ladder_ms"] / pm["ladder_launches"] / 1e3
        sgm = pm["sigma_ms"] / max(pm["sigma_calls"], 1) / 1e3
        out["sigma_tflops_mixed"] = info.flops_sigma / sgm / 1e12
        fz = fused_ranges(plan.tuning())
        kind = "two-phase xgemm" + "".join(f", fused {lo}-{hi}" for lo, hi in fz)
        out["roofline"] = roofline(info, syn, a.precision, lm, measured_traffic(a.workload, a.precision),
                                   f"{a.precision} ladder ({kind}; per sigma)", fz)
        out["roofline"]["launches"] = pm["ladder_launches"]
        out["roofline_fp64"] = roof64
    else:
        out["roofline"] = roof64
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(syn, plan, res_64["sigma_fp64"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
