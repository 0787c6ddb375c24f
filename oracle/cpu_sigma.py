"""CPU sigma with the reference's ladder loop structure (the CPU baseline).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): tests/ and the
cpu_baseline leg of bench.py use it; nothing in tigerci_amd imports it.

``cpu_ladder`` runs oracle/cpu_ladder.c (OpenMP over the outer virtual a, one
DGEMM per a assembling (ab|cd) from L^J, per-walk DGEMV products; the loop of
all_external_3_4_loops_cho_2 / ext34_parallel_load /
ext_3_4_seg_loops_vec_lmo_res_2, three_and_four_seg_mod_1.f90:318-632,
3240-3627) with scipy's BLAS, one BLAS thread per OpenMP thread as the
reference runs sequential BLAS inside its OpenMP loop.  ``cpu_sigma`` adds
every other term from the numpy oracle.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .problem import Problem
from .sigma import Integrals, sigma as oracle_sigma, ALL_TERMS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtci_cpu_ladder.so")
_lib = None


def build() -> str:
    """Compile cpu_ladder.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", HERE, "libtci_cpu_ladder.so"], check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        i64, vp = ctypes.c_int64, ctypes.c_void_p
        lib.tci_cpu_ladder.argtypes = [vp, vp, i64, i64, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, i64,
                                       ctypes.c_int]
        lib.tci_cpu_ladder.restype = ctypes.c_int
        _lib = lib
    return _lib


def _blas_pointer(name: str) -> int:
    """Address of scipy's BLAS routine ``name`` (Fortran interface)."""
    from scipy.linalg import cython_blas
    cap = cython_blas.__pyx_capi__[name]
    api = ctypes.pythonapi
    api.PyCapsule_GetName.restype = ctypes.c_char_p
    api.PyCapsule_GetName.argtypes = [ctypes.py_object]
    api.PyCapsule_GetPointer.restype = ctypes.c_void_p
    api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return api.PyCapsule_GetPointer(cap, api.PyCapsule_GetName(cap))


def host_cores() -> int:
    """Cores this process may run on."""
    return len(os.sched_getaffinity(0))


def cpu_ladder(prob: Problem, x: np.ndarray, sigma: np.ndarray | None = None, a_list=None,
               threads: int | None = None, L: np.ndarray | None = None) -> np.ndarray:
    """sigma += ladder part of (H - E_ref) x over the outer virtuals a_list
    (default all).  FP64: L (default prob.L) is converted once if it is not
    float64; pass the converted array to reuse it."""
    from threadpoolctl import threadpool_limits
    lib = _load()
    lay = prob.layout()
    L = np.ascontiguousarray(prob.L if L is None else L, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(lay.n_csf) if sigma is None else sigma
    pairs = np.ascontiguousarray(prob.pairs, dtype=np.int32)
    dom_off = np.ascontiguousarray(prob.dom_off, dtype=np.int64)
    dom_idx = np.ascontiguousarray(prob.dom_idx, dtype=np.int32)
    aa = np.ascontiguousarray(lay.aa_off, dtype=np.int64)
    ab = np.ascontiguousarray(lay.ab_off, dtype=np.int64)
    al = None if a_list is None else np.ascontiguousarray(a_list, dtype=np.int64)
    with threadpool_limits(limits=1, user_api="blas"):
        rc = lib.tci_cpu_ladder(_blas_pointer("dgemm"), _blas_pointer("dgemv"), prob.n_int, prob.n_v, prob.n_chol,
                                L.ctypes.data, prob.n_pairs, pairs.ctypes.data if pairs.size else None,
                                dom_off.ctypes.data, dom_idx.ctypes.data if dom_idx.size else None,
                                aa.ctypes.data if aa.size else None, ab.ctypes.data if ab.size else None,
                                x.ctypes.data, out.ctypes.data, al.ctypes.data if al is not None else None,
                                0 if al is None else al.size, threads or host_cores())
    if rc != 0:
        raise MemoryError("cpu_ladder: scratch allocation failed")
    return out


def cpu_sigma(prob: Problem, x: np.ndarray, ints: Integrals | None = None, threads: int | None = None) -> np.ndarray:
    """(H - E_ref) x: the ladder from cpu_ladder, every other term from the
    numpy oracle."""
    ints = ints or Integrals(prob)
    out = oracle_sigma(prob, x, ints, terms=ALL_TERMS - {"ladder"})
    return cpu_ladder(prob, x, out, threads=threads)
