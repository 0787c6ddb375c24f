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


def ladder_floor_bytes(syn, precision: str) -> float:
    """Byte floor of one sigma's ladder: every walk's L^J[D,D] block read once
    (4 bytes per element in split form or FP32 storage, 8 in FP64), the
    full-domain walks sharing L^J itself; C and R are negligible beside it."""
    d = np.diff(np.asarray(syn.dom_off)).astype(np.float64)
    full = d == syn.n_v
    e_l = 4.0 if (precision != "fp64" or syn.L.dtype == np.float32) else 8.0
    per_j = e_l * ((d[~full] ** 2).sum() + (syn.n_v ** 2 if full.any() else 0.0))
    return float(per_j * syn.n_chol)


def kernel_table(plan, tci, x_host, precision: str, n: int) -> list:
    """Per-kernel event times of n eager profiled sigma builds of one precision."""
    xp = tci.DeviceArray.from_numpy(x_host)
    yp = tci.DeviceArray(plan.n_csf)
    plan.sigma(xp, yp, precision=precision)
    plan.profile(True)
    for _ in range(n):
        plan.sigma(xp, yp, precision=precision)
    tab = plan.profile_kernels(precision)
    lad = plan.profile_read(precision)
    plan.profile(False)
    for k in tab:
        k["ms_per_sigma"] = k["ms"] / n
        k["launches_per_sigma"] = k["launches"] / n
    return tab, lad


def kernel_traffic(workload: str, precision: str, kernel: str):
    """HBM bytes per launch of one kernel (rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE, each scaled by the copy-kernel calibration), from the
    machine-readable lines of profiles/kernel_traffic.txt, or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "kernel_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.rstrip("\n").split("|")  # kernel names may hold '|', as in "(ab|cd)"
        if len(f) >= 4 and f[0] == workload and f[1] == precision and "|".join(f[2:-1]) == kernel:
            return float(f[-1])
    return None


def kernel_roofline(k: dict, precision: str, traffic) -> dict:
    """Roofline of one kernel: algorithmic flops per launch (2 m n k over the
    valid extents of its products) over its average launch time, against the
    dense MFMA peak of the precision the products run in (the published
    2.5 PFLOP/s bf16 for a split product, whose three MFMAs make the units
    busy 3x that fraction)."""
    t = k["ms"] / max(k["launches"], 1) / 1e3
    ach = k["flops"] / t
    r = {"kernel": k["name"], "bound": "mfma", "achieved": ach / 1e12, "peak": PEAK[precision] / 1e12,
         "unit": "TFLOP/s", "frac": ach / PEAK[precision], "traffic": traffic,
         "avg_launch_ms": t * 1e3, "launches_per_sigma": k["launches_per_sigma"],
         "algorithmic_flops_per_launch": k["flops"]}
    if precision != "fp64":
        r["mfma_busy_frac"] = 3.0 * ach / PEAK[precision]
    if traffic:
        r["traffic_gbs"] = traffic / t / 1e9
    return r


def resident_ladder(tuning: str, precision: str) -> bool:
    """Whether the plan's sigma of this precision takes the ladder from the
    resident (ab|cd) block (its tuning record line ends in "-> resident")."""
    return re.search(r"^%s ladder \(ab\|cd\) resident:.*-> resident$" % re.escape(precision), tuning,
                     re.MULTILINE) is not None


def resident_counts(syn):
    """(n_p, n_m, walks, walks with i != j) of the resident-(ab|cd) product."""
    n_v = syn.n_v
    pairs = np.asarray(syn.pairs).reshape(-1, 2)
    return n_v * (n_v + 1) // 2, n_v * (n_v - 1) // 2, len(pairs), int((pairs[:, 0] != pairs[:, 1]).sum())


def ladder_section(info, syn, precision: str, seconds: float, fused=(), resident: bool = False) -> dict:
    """The whole ladder of one sigma (several launches) against the MFMA roof.
    Cholesky-direct: F_lad = sum_P 4 d_P^3 n_chol algorithmic flops, byte floor
    L^J[D,D] read once, and the chosen schedule's own minimum bytes.  Resident
    (ab|cd): F = 2 n_p^2 n_walks + 2 n_m^2 n_walks(i != j) (the product over
    the packed symmetric and antisymmetric blocks), byte floor V+- read once
    (8 bytes an element in FP64, 4 in split form) plus X and S."""
    if resident:
        n_p, n_m, nw, nwm = resident_counts(syn)
        flops = 2.0 * n_p * n_p * nw + 2.0 * n_m * n_m * nwm
        e = 8.0 if precision == "fp64" else 4.0
        floor = e * (n_p * n_p + n_m * n_m) + e * (n_p * nw + n_m * nwm) + 8.0 * (n_p * nw + n_m * nwm)
        sched = floor
    else:
        flops = info.flops_ladder
        floor = ladder_floor_bytes(syn, precision)
        sched = ladder_alg_bytes(syn, precision, fused)
    return {"algorithm": "resident (ab|cd)" if resident else "Cholesky-direct",
            "ms": seconds * 1e3, "algorithmic_flops": flops, "tflops": flops / seconds / 1e12,
            "frac_mfma": flops / seconds / PEAK[precision],
            "byte_floor": floor, "floor_gbs": floor / seconds / 1e9, "schedule_bytes": sched}


def sigma_flops(info, syn, tuning: str, precision: str) -> float:
    """Executed flops of one sigma build of this precision: the plan's count
    with the ladder term replaced by the algorithm the build runs (the
    resident (ab|cd) product when the tuning record says so)."""
    lad = ladder_section(info, syn, precision, 1.0, (), resident_ladder(tuning, precision))["algorithmic_flops"]
    return info.flops_sigma - info.flops_ladder + lad


def cpu_baseline(syn, x, s64_gpu, n_sigma_fp64: int, budget_s: float = 25.0) -> dict:
    """The reference's CPU sigma path, restated in C with its loop structure
    (oracle/cpu_ladder.c: OpenMP over the outer virtual a, one DGEMM per a
    assembling (ab|cd) from L^J, per-walk DGEMV products,
    three_and_four_seg_mod_1.f90:318-632,3240-3627) and every other sigma term
    from the numpy oracle, timed on this host's cores in FP64.  The other
    terms run whole; the ladder runs whole when it fits the budget, otherwise
    over an evenly strided sample of the outer virtuals, scaled by the sampled
    fraction.  A whole run is checked against the GPU FP64 sigma of the same
    vector.  value = seconds of the FP64 Davidson solve's sigma builds on the
    CPU (its vector work is not counted)."""
    from threadpoolctl import threadpool_info
    from oracle.problem import Problem
    from oracle.sigma import Integrals, sigma as oracle_sigma, ALL_TERMS
    from oracle.cpu_sigma import cpu_ladder, host_cores
    cores = host_cores()
    prob = Problem(syn.n_int, syn.n_v, syn.n_chol, syn.eps, syn.L, syn.pairs, syn.dom_off, syn.dom_idx)
    L64 = np.ascontiguousarray(syn.L, dtype=np.float64)
    t0 = time.perf_counter()
    ints = Integrals(prob)
    t_int = time.perf_counter() - t0
    t0 = time.perf_counter()
    rest = oracle_sigma(prob, x, ints, terms=ALL_TERMS - {"ladder"})
    t_rest = time.perf_counter() - t0
    del ints
    # one outer virtual on one thread, to size the sample
    probe = np.zeros_like(rest)
    t0 = time.perf_counter()
    cpu_ladder(prob, x, probe, a_list=[syn.n_v // 2], L=L64)
    t_one = time.perf_counter() - t0
    n_s = int(min(syn.n_v, max(cores, (budget_s - t_rest) / max(t_one, 1e-6) * cores)))
    a_list = np.unique(np.linspace(0, syn.n_v - 1, n_s).round().astype(np.int64))
    lad = np.zeros_like(rest)
    t0 = time.perf_counter()
    cpu_ladder(prob, x, lad, a_list=a_list, L=L64)
    t_lad = time.perf_counter() - t0
    whole = a_list.size == syn.n_v
    t_sigma = t_rest + t_lad * syn.n_v / a_list.size
    rel = None
    if whole and s64_gpu is not None:
        cs = rest + lad
        rel = float(np.linalg.norm(cs - s64_gpu) / np.linalg.norm(s64_gpu))
    blas = [f"{b.get('internal_api')} {b.get('version')}" for b in threadpool_info() if b.get("user_api") == "blas"]
    return {"value": t_sigma * n_sigma_fp64, "unit": "s", "cores": cores, "kind": "port",
            "sample": (f"one FP64 sigma of this workload on {cores} host cores: the ladder in C with the "
                       f"reference's loop structure (OpenMP over a, DGEMM per a, DGEMV per walk; BLAS "
                       f"{', '.join(blas) or 'unknown'}, one BLAS thread per OpenMP thread) over "
                       f"{a_list.size} of {syn.n_v} outer virtuals ({t_lad:.1f} s"
                       f"{'' if whole else ', scaled by the sampled fraction'}), the other terms in numpy "
                       f"({t_rest:.1f} s); times the {n_sigma_fp64} sigma builds of the FP64 solve"),
            "sigma_seconds": t_sigma, "ladder_seconds_measured": t_lad, "ladder_outer_virtuals": int(a_list.size),
            "rest_seconds": t_rest, "integral_assembly_seconds": t_int,
            "check_rel_err_vs_gpu_fp64": rel}


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
    ap.add_argument("--iteration-graph", dest="iteration_graph", action="store_true", default=True,
                    help="Davidson: replay each iteration's device work (residuals, preconditioner, "
                         "Gram-Schmidt, sigma builds, subspace columns) from a captured HIP graph (default "
                         "with --graph)")
    ap.add_argument("--sigma-graph-only", dest="iteration_graph", action="store_false",
                    help="Davidson: graphs for the sigma builds only, the vector work eager")
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
           "ladder_flops_per_sigma": info.flops_ladder, "sigma_flops": info.flops_sigma,
           "sigma_flops_note": "plan counts with the Cholesky-direct ladder; the sigma_tflops figures use the "
                               "ladder algorithm each precision runs (ladder_section)"}
    out = {"metric": METRIC, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "scaling": "weak",
           "vs_baseline": None,
           "data": f"synthetic (seeded N(0,1) exp(-J/80) x "
                   f"{a.l_scale if a.l_scale is not None else WORKLOADS[a.workload][6]} Cholesky vectors)"}

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
        tab, pr = kernel_table(plan, tci, random_vector(plan.n_csf, a.seed + 1), a.precision, max(a.steps, 3))
        cfg["schedule"] = plan.tuning().strip().splitlines()
        lad = pr["ladder_ms"] / max(pr["ladder_launches"], 1) / 1e3
        if a.precision != "fp64":  # precision study: this sigma against the FP64 one
            s_mx = y.numpy()
            plan.sigma(x, y, precision="fp64")
            s_64 = y.numpy()
            xv = x.numpy()
            out["sigma_check"] = {"sigma_rel_err": float(np.linalg.norm(s_mx - s_64) / np.linalg.norm(s_64)),
                                  "rayleigh_err_Eh": float(abs(xv @ s_mx - xv @ s_64))}
        dom = max(tab, key=lambda k: k["ms_per_sigma"])
        out.update({"value": sigma_flops(info, syn, plan.tuning(), a.precision) / dt / 1e12, "unit": "TFLOP/s",
                    "higher_is_better": True,
                    "ms_per_step": dt * 1e3, "dtype": a.precision, "config": cfg,
                    "roofline": kernel_roofline(dom, a.precision, kernel_traffic(a.workload, a.precision, dom["name"])),
                    "ladder_section": ladder_section(info, syn, a.precision, lad, fused_ranges(plan.tuning()),
                                                     resident_ladder(plan.tuning(), a.precision)),
                    "kernels": [dict(name=k["name"], ms_per_sigma=k["ms_per_sigma"],
                                     launches_per_sigma=k["launches_per_sigma"],
                                     tflops=(k["flops"] * k["launches"] / k["ms"] / 1e9) if k["ms"] > 0 else None)
                                for k in sorted(tab, key=lambda k: -k["ms_per_sigma"])[:8]]})
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
    dav_graph = ("iteration" if a.iteration_graph else True) if a.graph else False
    cfg["davidson_graph"] = dav_graph or "none"
    def solve(prec):
        return plan.davidson(roots=a.roots, energy_tol=1e-6, norm_tol=1e-5, precision=prec,
                             switch_tol=a.switch_tol, graph=dav_graph)

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
    # per-kernel tables from eager profiled builds after the timed region
    # (graph replays time only whole builds)
    xk = random_vector(plan.n_csf, a.seed + 1)
    tab64, lp64 = kernel_table(plan, tci, xk, "fp64", 3)
    p64 = dict(lp64, sigma_ms=p64["sigma_ms"], sigma_calls=p64["sigma_calls"])
    tabm = []
    if a.precision != "fp64":
        tabm, lpm = kernel_table(plan, tci, xk, a.precision, 3)
        pm = dict(lpm, sigma_ms=pm["sigma_ms"], sigma_calls=pm["sigma_calls"])
    lad = p64["ladder_ms"] / max(p64["ladder_launches"], 1) / 1e3
    sig64 = p64["sigma_ms"] / max(p64["sigma_calls"], 1) / 1e3
    out.update({
        "value": t_m, "unit": "s", "higher_is_better": False, "ms_per_step": t_m * 1e3,
        "dtype": "f64" if a.precision == "fp64" else f"{a.precision}+f64", "config": dict(cfg, roots=a.roots),
        "fp64_seconds": t_64, "speedup_vs_fp64": t_64 / t_m,
        "energies_mixed": [float(v) for v in e_m], "energies_fp64": [float(v) for v in e_64],
        "energy_diff_uEh": float(np.abs(e_m - e_64).max() * 1e6),
        "converged": {"mixed": res_m["converged"], "fp64": res_64["converged"]},
        "solver_seconds": {"mixed": {"total": res_m["seconds"], "sigma": res_m["seconds_sigma"]},
                           "fp64": {"total": res_64["seconds"], "sigma": res_64["seconds_sigma"]}},
        "iterations": {"mixed": res_m["iterations"], "fp64": res_64["iterations"],
                       "sigma_mixed": res_m["sigma_mixed"], "sigma_fp64_in_mixed": res_m["sigma_fp64"],
                       "sigma_fp64": res_64["sigma_fp64"]},
        "sigma_check": sig_err,
        "sigma_tflops_fp64": sigma_flops(info, syn, plan.tuning(), "fp64") / sig64 / 1e12,
    })
    out["config"]["schedule"] = plan.tuning().strip().splitlines()
    # the dominant kernel of the timed step: per-sigma kernel time weighted by
    # the mixed solve's sigma counts in each precision
    n_mix, n_ref = res_m["sigma_mixed"], res_m["sigma_fp64"]
    cands = [(k["ms_per_sigma"] * n_mix, a.precision, k) for k in tabm] + \
            [(k["ms_per_sigma"] * n_ref, "fp64", k) for k in tab64]
    _, dprec, dk = max(cands, key=lambda t: t[0])
    out["roofline"] = kernel_roofline(dk, dprec, kernel_traffic(a.workload, dprec, dk["name"]))
    out["roofline"]["share_of_step"] = dk["ms_per_sigma"] * (n_mix if dprec != "fp64" else n_ref) / 1e3 / t_m
    k64 = max(tab64, key=lambda k: k["ms_per_sigma"])
    out["roofline_fp64"] = kernel_roofline(k64, "fp64", kernel_traffic(a.workload, "fp64", k64["name"]))
    out["ladder_section"] = {"fp64": ladder_section(info, syn, "fp64", lad,
                                                    resident=resident_ladder(plan.tuning(), "fp64"))}
    out["kernels"] = {prec: [dict(name=k["name"], ms_per_sigma=k["ms_per_sigma"],
                                  launches_per_sigma=k["launches_per_sigma"],
                                  tflops=(k["flops"] * k["launches"] / k["ms"] / 1e9) if k["ms"] > 0 else None)
                             for k in sorted(t, key=lambda k: -k["ms_per_sigma"])[:8]]
                      for prec, t in (("fp64", tab64), (a.precision, tabm)) if t}
    if a.precision != "fp64" and pm["ladder_launches"]:
        lm = pm["ladder_ms"] / pm["ladder_launches"] / 1e3
        sgm = pm["sigma_ms"] / max(pm["sigma_calls"], 1) / 1e3
        out["sigma_tflops_mixed"] = sigma_flops(info, syn, plan.tuning(), a.precision) / sgm / 1e12
        out["ladder_section"][a.precision] = ladder_section(info, syn, a.precision, lm, fused_ranges(plan.tuning()),
                                                            resident_ladder(plan.tuning(), a.precision))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(syn, xr, s64, res_64["sigma_fp64"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
