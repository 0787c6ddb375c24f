"""Build libtigersigma.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m tigerci_amd.build [--jobs N] [--incremental]

Each .hip source compiles to an object in tigerci_amd/_build/ (in parallel),
then one hipcc link produces tigerci_amd/libtigersigma.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtigersigma.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]
SOURCES = ["plan.hip", "gemm.hip", "dense_gemm.hip", "dense_split.hip", "pairs_full.hip", "ladder.hip", "ladder_mixed.hip", "terms.hip", "davidson.hip",
           "selftest.hip", "tune.hip", "xgemm192_bf16.hip", "xgemm192_f16.hip", "ladder_fused_split.hip",
           "ladder_fused128.hip", "ladder_fused192.hip", "xgemm192d.hip", "noalloc.hip", "ladder_vv.hip", "walks.hip", "xgemm256.hip", "chol_transform.hip", "spin.hip", "records.hip"]


def _deps(path: str, seen: set) -> set:
    """Files a source includes with quotes, transitively."""
    with open(path) as f:
        for line in f:
            t = line.strip()
            if t.startswith("#include") and '"' in t:
                inc = os.path.normpath(os.path.join(os.path.dirname(path), t.split('"')[1]))
                if inc not in seen and os.path.exists(inc):
                    seen.add(inc)
                    _deps(inc, seen)
    return seen


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and os.path.exists(obj):
        path = os.path.join(CSRC, src)
        newest = max(os.path.getmtime(f) for f in _deps(path, {path}))
        if os.path.getmtime(obj) > newest:
            return obj
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    err = "\n".join(x for x in r.stderr.splitlines() if "packed-fp32-ops" not in x)
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every source (or, with ``incremental``, those whose object is
    older than the source or a file it includes) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources newer than their objects or the files they include")
    a = ap.parse_args()
    print(build(a.jobs, a.incremental))
