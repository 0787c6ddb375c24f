"""tigerci_amd -- MI355X sigma = H.c and Davidson-Liu for local MRSDCI with
Cholesky-decomposed integrals.

Host-side mirror of the reference's sigma interface over the C ABI of
``libtigersigma.so`` (include/tigersigma.h):

  Plan(problem arrays)            set-up that htimesc does on its first call
                                  (htimesc.f90:77-82) plus makeIntegrals
                                  (make_all_integrals.f90:23)
  Plan.sigma(c, out)              htimesc(civec, sigmavec, data) on device
                                  vectors (htimesc.f90:46-147)
  htimesc(civec, sigmavec, plan)  the reference's argument order and
                                  semantics: sigmavec += H civec, host arrays
  Plan.diagonal()                 H_ii (diag_element_mod_2.f90:83)
  Plan.davidson(...)              davidson_liu (davidson.f90:47-239)

The HIP library is required: nothing here falls back to a CPU path.  Build it
with ``python -m tigerci_amd.build``.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtigersigma.so")

TCI_F64, TCI_F32 = 0, 1
PRECISION = {"fp64": 0, "bf16x3": 1, "fp16x3": 2}
TCI_ACCUMULATE = 0x100
TCI_ADD_EREF = 0x200
TCI_SIGMA_GRAPH = 0x400
TCI_DAVIDSON_GRAPH = 0x800


class TigerSigmaError(RuntimeError):
    pass


class _Problem(ctypes.Structure):
    _fields_ = [("n_int", ctypes.c_int64), ("n_v", ctypes.c_int64), ("n_chol", ctypes.c_int64),
                ("eps", ctypes.c_void_p), ("chol", ctypes.c_void_p), ("chol_dtype", ctypes.c_int32),
                ("ladder_dtype", ctypes.c_int32), ("n_pairs", ctypes.c_int64), ("pairs", ctypes.c_void_p),
                ("dom_off", ctypes.c_void_p), ("dom_idx", ctypes.c_void_p), ("h", ctypes.c_void_p)]


class _Info(ctypes.Structure):
    _fields_ = [("n_csf", ctypes.c_int64), ("n_ladder_items", ctypes.c_int64), ("e_ref", ctypes.c_double),
                ("flops_ladder", ctypes.c_double), ("flops_sigma", ctypes.c_double),
                ("bytes_ladder", ctypes.c_double), ("device_bytes", ctypes.c_int64)]


class _DavParams(ctypes.Structure):
    _fields_ = [("roots", ctypes.c_int32), ("maxit", ctypes.c_int32), ("energy_tol", ctypes.c_double),
                ("norm_tol", ctypes.c_double), ("precision", ctypes.c_int32), ("switch_tol", ctypes.c_double),
                ("max_vectors", ctypes.c_int32), ("verbose", ctypes.c_int32)]


class _DavResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("sigma_fp64", ctypes.c_int32), ("sigma_mixed", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("seconds", ctypes.c_double), ("seconds_sigma", ctypes.c_double),
                ("collapses", ctypes.c_int32)]


_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libtigersigma.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TigerSigmaError(f"{p} is missing: build it with `python -m tigerci_amd.build`")
    lib = ctypes.CDLL(p)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.tci_plan_create.argtypes = [ctypes.POINTER(_Problem), ctypes.POINTER(vp)]
    lib.tci_plan_destroy.argtypes = [vp]
    lib.tci_plan_destroy.restype = None
    lib.tci_plan_get_info.argtypes = [vp, ctypes.POINTER(_Info)]
    lib.tci_sigma_apply.argtypes = [vp, vp, vp, i64, i32, vp]
    lib.tci_htimesc.argtypes = [vp, vp, vp, i64]
    lib.tci_htimesc_noalloc.argtypes = [vp, vp, vp, i64]
    lib.tci_diagonal.argtypes = [vp, vp, i64, vp]
    lib.tci_davidson_run.argtypes = [vp, ctypes.POINTER(_DavParams), vp, vp, ctypes.POINTER(_DavResult)]
    lib.tci_selftest_mfma.argtypes = []
    lib.tci_device_alloc.argtypes = [i64, ctypes.POINTER(vp)]
    lib.tci_device_free.argtypes = [vp]
    lib.tci_memcpy_h2d.argtypes = [vp, vp, i64]
    lib.tci_memcpy_d2h.argtypes = [vp, vp, i64]
    lib.tci_memcpy_d2d.argtypes = [vp, vp, i64]
    lib.tci_device_synchronize.argtypes = []
    lib.tci_profile_enable.argtypes = [vp, i32]
    lib.tci_profile_read.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
    lib.tci_plan_tuning.argtypes = [vp, ctypes.c_char_p, i64]
    lib.tci_profile_kernels.argtypes = [vp, i32, ctypes.c_char_p, i64]
    lib.tci_sigma_apply_batch.argtypes = [vp, vp, vp, i64, i32, i32, vp]
    lib.tci_davidson_gen_run.argtypes = [vp, ctypes.POINTER(_DavParams), vp, vp, vp, ctypes.POINTER(_DavResult)]
    lib.tci_acpf_gvalues.argtypes = [i32, i32, ctypes.c_double, vp]
    lib.tci_internal_walks.argtypes = [i32, i32, i32, i32, i32, vp, i32, i32, i64, vp, vp, ctypes.POINTER(i64)]
    lib.tci_allowed_virtuals.argtypes = [i32, i32, i32, vp, i64, vp, vp, vp, vp, vp, vp, vp, i64,
                                         ctypes.POINTER(i64)]
    lib.tci_cholesky_ao_to_mo.argtypes = [i32, i32, i64, i64, vp, vp, vp, vp, vp]
    lib.tci_spin_tables.argtypes = [i32, i32, vp, vp, vp, vp]
    lib.tci_contracted_cycles.argtypes = [i32, i32, i64, vp, vp, vp]
    lib.tci_ladder_records.argtypes = [i32, i32, i32, i64, vp, vp, vp, vp, vp, i64, vp, vp]
    lib.tci_last_error.restype = ctypes.c_char_p
    for f in ("tci_plan_create", "tci_plan_get_info", "tci_sigma_apply", "tci_htimesc", "tci_htimesc_noalloc",
              "tci_diagonal",
              "tci_davidson_run", "tci_selftest_mfma", "tci_device_alloc", "tci_device_free", "tci_memcpy_h2d",
              "tci_memcpy_d2h", "tci_memcpy_d2d", "tci_device_synchronize", "tci_profile_enable",
              "tci_profile_read", "tci_profile_kernels", "tci_plan_tuning", "tci_davidson_gen_run", "tci_acpf_gvalues",
              "tci_sigma_apply_batch", "tci_internal_walks", "tci_allowed_virtuals", "tci_cholesky_ao_to_mo",
              "tci_spin_tables", "tci_contracted_cycles", "tci_ladder_records"):
        getattr(lib, f).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = _lib.tci_last_error().decode(errors="replace") if _lib is not None else ""
        raise TigerSigmaError(f"tigersigma error {rc}: {msg}")


def selftest_mfma() -> None:
    lib = load_library()
    _check(lib.tci_selftest_mfma())


ACPF, ACPF2, MRAQCC, ACPF_CUSTOM, ACPF_TEST = 1, 2, 3, 4, 9


def acpf_gvalues(functional: int, n_corr_elec: int, custom_g: float = 0.0) -> np.ndarray:
    """g1..g5 of an ACPF-type functional (get_gvalues, gtimesc.f90)."""
    g = np.zeros(5)
    _check(load_library().tci_acpf_gvalues(functional, n_corr_elec, custom_g, g.ctypes.data))
    return g


def synchronize() -> None:
    _check(load_library().tci_device_synchronize())


class DeviceArray:
    """A float64 vector in device memory, allocated through the library's own
    HIP runtime (so no second GPU runtime enters the process)."""

    def __init__(self, n: int):
        self._lib = load_library()
        self.n = int(n)
        p = ctypes.c_void_p()
        _check(self._lib.tci_device_alloc(8 * self.n, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.n:
            raise TigerSigmaError("size mismatch")
        _check(self._lib.tci_memcpy_h2d(self.ptr, a.ctypes.data, 8 * self.n))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n)
        _check(self._lib.tci_memcpy_d2h(out.ctypes.data, self.ptr, 8 * self.n))
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        _check(self._lib.tci_memcpy_d2d(self.ptr, other.ptr, 8 * self.n))

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self._lib.tci_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(t) -> int:
    """Device address of a DeviceArray, of an object exposing data_ptr(), or an int."""
    if isinstance(t, int):
        return t
    if isinstance(t, DeviceArray):
        return t.ptr
    return t.data_ptr()


def _need(t, n: int, what: str) -> None:
    """A DeviceArray passed as a vector must hold at least n doubles; the
    library reads and writes n of them (raw addresses are the caller's
    responsibility)."""
    if isinstance(t, DeviceArray) and t.n < n:
        raise TigerSigmaError(f"{what} holds {t.n} doubles, needs {n}")


def _stream(stream):
    if stream is None:
        return None
    return stream if isinstance(stream, int) else stream.cuda_stream


@dataclass
class PlanInfo:
    n_csf: int
    n_ladder_items: int
    e_ref: float
    flops_ladder: float
    flops_sigma: float
    bytes_ladder: float
    device_bytes: int


class Plan:
    """Device plan of one CI problem (see the module docstring)."""

    def __init__(self, n_int: int, n_v: int, n_chol: int, eps, chol, pairs, dom_off, dom_idx,
                 ladder_dtype: str = "f64", h=None):
        """``h``: optional one-electron integrals, an (n_orb, n_orb) symmetric
        matrix or its packed lower triangle (the reference's pp order); when
        given, the Fock matrix is built from it and ``eps`` may be None."""
        self._lib = load_library()
        n_orb = n_int + n_v
        hp = None
        if h is not None:
            h = np.asarray(h, dtype=np.float64)
            if h.shape == (n_orb, n_orb):
                r, c = np.tril_indices(n_orb)
                hp = np.zeros(n_orb * (n_orb + 1) // 2)
                hp[r * (r + 1) // 2 + c] = h[r, c]
            elif h.size == n_orb * (n_orb + 1) // 2:
                hp = np.ascontiguousarray(h.ravel())
            else:
                raise TigerSigmaError("h must be (n_orb, n_orb) or packed n_orb(n_orb+1)/2")
        if eps is None:
            if hp is None:
                raise TigerSigmaError("eps or h is required")
            eps = np.zeros(n_orb)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        chol = np.ascontiguousarray(chol)
        if chol.dtype not in (np.float64, np.float32):
            chol = chol.astype(np.float64)
        if chol.size != n_orb * (n_orb + 1) // 2 * n_chol:
            raise TigerSigmaError("chol must hold n_orb(n_orb+1)/2 records of n_chol values")
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        dom_off = np.ascontiguousarray(dom_off, dtype=np.int64)
        dom_idx = np.ascontiguousarray(dom_idx, dtype=np.int32)
        self._keep = (eps, chol, pairs, dom_off, dom_idx, hp)
        pr = _Problem(n_int, n_v, n_chol, eps.ctypes.data, chol.ctypes.data,
                      TCI_F64 if chol.dtype == np.float64 else TCI_F32,
                      TCI_F64 if ladder_dtype == "f64" else TCI_F32,
                      pairs.shape[0], pairs.ctypes.data if pairs.size else None,
                      dom_off.ctypes.data, dom_idx.ctypes.data if dom_idx.size else None,
                      hp.ctypes.data if hp is not None else None)
        h = ctypes.c_void_p()
        _check(self._lib.tci_plan_create(ctypes.byref(pr), ctypes.byref(h)))
        self._h = h
        self._keep = None
        inf = _Info()
        _check(self._lib.tci_plan_get_info(self._h, ctypes.byref(inf)))
        self.info = PlanInfo(inf.n_csf, inf.n_ladder_items, inf.e_ref, inf.flops_ladder, inf.flops_sigma,
                             inf.bytes_ladder, inf.device_bytes)
        self.n_csf = inf.n_csf

    @classmethod
    def from_problem(cls, prob, ladder_dtype: str = "f64") -> "Plan":
        """Build from any object with the attributes n_int, n_v, n_chol, eps, L,
        pairs, dom_off, dom_idx (L: records x n_chol) and optionally h."""
        return cls(prob.n_int, prob.n_v, prob.n_chol, prob.eps, prob.L, prob.pairs, prob.dom_off,
                   prob.dom_idx, ladder_dtype=ladder_dtype, h=getattr(prob, "h", None))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tci_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sigma(self, c, out, precision: str = "fp64", accumulate: bool = False, add_eref: bool = False,
              stream=None, graph: bool = False) -> None:
        """out = (H - E_ref) c on device (DeviceArray or device addresses).
        ``graph``: replay the build from a HIP graph captured on first use."""
        flags = PRECISION[precision] | (TCI_ACCUMULATE if accumulate else 0) | (TCI_ADD_EREF if add_eref else 0)
        flags |= TCI_SIGMA_GRAPH if graph else 0
        _need(c, self.n_csf, "c")
        _need(out, self.n_csf, "out")
        _check(self._lib.tci_sigma_apply(self._h, _ptr(c), _ptr(out), self.n_csf, flags, _stream(stream)))

    def sigma_batch(self, c, out, nvec: int, precision: str = "fp64", accumulate: bool = False,
                    add_eref: bool = False, stream=None) -> None:
        """nvec sigma builds at once; c and out hold nvec vectors back to back."""
        flags = PRECISION[precision] | (TCI_ACCUMULATE if accumulate else 0) | (TCI_ADD_EREF if add_eref else 0)
        _need(c, nvec * self.n_csf, "c")
        _need(out, nvec * self.n_csf, "out")
        _check(self._lib.tci_sigma_apply_batch(self._h, _ptr(c), _ptr(out), self.n_csf, nvec, flags,
                                               _stream(stream)))

    def diagonal(self, out=None, stream=None):
        """Exact H_ii - E_ref into ``out`` (device); returns numpy when out is None."""
        if out is None:
            d = DeviceArray(self.n_csf)
            _check(self._lib.tci_diagonal(self._h, d.ptr, self.n_csf, None))
            return d.numpy()
        _need(out, self.n_csf, "out")
        _check(self._lib.tci_diagonal(self._h, _ptr(out), self.n_csf, _stream(stream)))
        return None

    def sigma_numpy(self, x: np.ndarray, precision: str = "fp64") -> np.ndarray:
        """Convenience: host vector in, host sigma out (device work in between)."""
        c = DeviceArray.from_numpy(x)
        y = DeviceArray(self.n_csf)
        self.sigma(c, y, precision=precision)
        return y.numpy()

    def tuning(self) -> str:
        """Launch schedules picked by the plan's autotuning (empty when off)."""
        buf = ctypes.create_string_buffer(1 << 16)
        _check(self._lib.tci_plan_tuning(self._h, buf, len(buf)))
        return buf.value.decode()

    def profile(self, on: bool = True) -> None:
        _check(self._lib.tci_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, precision: str = "fp64") -> dict:
        lm, sm = ctypes.c_double(), ctypes.c_double()
        ln, sn = ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.tci_profile_read(self._h, PRECISION[precision], ctypes.byref(lm), ctypes.byref(ln),
                                          ctypes.byref(sm), ctypes.byref(sn)))
        return dict(ladder_ms=lm.value, ladder_launches=ln.value, sigma_ms=sm.value, sigma_calls=sn.value)

    def profile_kernels(self, precision: str = "fp64") -> list:
        """Per-kernel event times of the profiled eager builds of one precision:
        [{name, launches, ms (total), flops (algorithmic, per launch)}]."""
        buf = ctypes.create_string_buffer(1 << 16)
        _check(self._lib.tci_profile_kernels(self._h, PRECISION[precision], buf, len(buf)))
        out = []
        for line in buf.value.decode().splitlines():
            name, n, ms, fl = line.rsplit("|", 3)
            out.append(dict(name=name, launches=int(n), ms=float(ms), flops=float(fl)))
        return out

    def davidson(self, roots: int = 1, energy_tol: float = 1e-6, norm_tol: float = 1e-6,
                 precision: str = "fp64", switch_tol: float | None = None, maxit: int = 0,
                 max_vectors: int = 0, vectors=None, verbose: bool = False, graph: bool | str = False,
                 g_values=None):
        """Lowest roots; returns (total energies, result dict).  ``vectors`` is an
        optional DeviceArray of roots*n_csf doubles for the eigenvectors;
        ``graph=True`` replays every sigma build from a captured HIP graph,
        ``graph="iteration"`` every iteration's device work;
        ``g_values`` (five g values, see ``acpf_gvalues``) solves the
        generalized ACPF-type problem (H - E_ref) c = E G c instead."""
        if switch_tol is None:
            switch_tol = energy_tol
        if graph not in (False, True, "iteration"):
            raise ValueError("graph must be False, True or 'iteration'")
        prec = PRECISION[precision] | (TCI_DAVIDSON_GRAPH if graph == "iteration" else
                                       TCI_SIGMA_GRAPH if graph else 0)
        p = _DavParams(roots, maxit, energy_tol, norm_tol, prec, switch_tol, max_vectors,
                       1 if verbose else 0)
        e = np.zeros(roots)
        r = _DavResult()
        if vectors is not None:
            _need(vectors, roots * self.n_csf, "vectors")
        vp = _ptr(vectors) if vectors is not None else None
        if g_values is None:
            _check(self._lib.tci_davidson_run(self._h, ctypes.byref(p), e.ctypes.data, vp, ctypes.byref(r)))
        else:
            g = np.ascontiguousarray(g_values, dtype=np.float64)
            if g.shape != (5,):
                raise ValueError("g_values must hold five values")
            _check(self._lib.tci_davidson_gen_run(self._h, ctypes.byref(p), g.ctypes.data, e.ctypes.data, vp,
                                                  ctypes.byref(r)))
        return e, dict(iterations=r.iterations, sigma_fp64=r.sigma_fp64, sigma_mixed=r.sigma_mixed,
                       converged=bool(r.converged), seconds=r.seconds, seconds_sigma=r.seconds_sigma,
                       collapses=r.collapses)

    def htimesc_host(self, civec: np.ndarray, sigmavec: np.ndarray) -> None:
        if civec.dtype != np.float64 or sigmavec.dtype != np.float64 or not sigmavec.flags.c_contiguous:
            raise TigerSigmaError("htimesc needs contiguous float64 arrays")
        civec = np.ascontiguousarray(civec)
        _check(self._lib.tci_htimesc(self._h, civec.ctypes.data, sigmavec.ctypes.data, self.n_csf))


def htimesc(civec: np.ndarray, sigmavec: np.ndarray, data: Plan) -> None:
    """sigmavec += H civec (bare Hamiltonian), the argument order of
    ``subroutine htimesc(civec, sigmavec, data)`` (htimesc.f90:46): civec is
    read-only, sigmavec accumulates (the caller zeroes it, davidson.f90:126)."""
    data.htimesc_host(civec, sigmavec)


def internal_walks(n_int: int, n_v: int, n_elec: int, references, n_ext: int, n_frozen: int = 0,
                   spin_mult: int = 1):
    """(occupations, kept) of every walk of the internal graph with
    n_elec - n_ext electrons, in lexical index order (tci_internal_walks).
    ``references``: internal occupations, one row per reference."""
    lib = load_library()
    refs = np.ascontiguousarray(np.asarray(references, dtype=np.int8).reshape(-1, n_int))
    n = ctypes.c_int64(0)
    _check(lib.tci_internal_walks(n_int, n_v, n_elec, spin_mult, refs.shape[0], refs.ctypes.data, n_frozen, n_ext,
                                  0, None, None, ctypes.byref(n)))
    occ = np.zeros((max(n.value, 1), n_int), np.int8)
    kept = np.zeros(max(n.value, 1), np.uint8)
    if n.value:
        _check(lib.tci_internal_walks(n_int, n_v, n_elec, spin_mult, refs.shape[0], refs.ctypes.data, n_frozen,
                                      n_ext, n.value, occ.ctypes.data, kept.ctypes.data, ctypes.byref(n)))
    return occ[:n.value], kept[:n.value].astype(bool)


def allowed_virtuals(references, walks, wp, tov):
    """(status, list_of_walk, lists) of tci_allowed_virtuals for internal walk
    occupations ``walks`` (n_walks x n_int), wp (n_int x n_int) and tov
    (n_int x n_int x n_v) interaction tables."""
    lib = load_library()
    walks = np.ascontiguousarray(np.asarray(walks, dtype=np.int8))
    n_walks, n_int = walks.shape
    refs = np.ascontiguousarray(np.asarray(references, dtype=np.int8).reshape(-1, n_int))
    wp = np.ascontiguousarray(np.asarray(wp, dtype=np.uint8))
    tov = np.ascontiguousarray(np.asarray(tov, dtype=np.uint8))
    n_v = tov.shape[2]
    status = np.zeros(n_walks, np.int32)
    low = np.zeros(n_walks, np.int64)
    off = np.zeros(n_walks + 1, np.int64)
    cap = n_walks * n_v
    virt = np.zeros(max(cap, 1), np.int32)
    nl = ctypes.c_int64(0)
    _check(lib.tci_allowed_virtuals(n_int, n_v, refs.shape[0], refs.ctypes.data, n_walks, walks.ctypes.data,
                                    wp.ctypes.data, tov.ctypes.data, status.ctypes.data, low.ctypes.data,
                                    off.ctypes.data, virt.ctypes.data, cap, ctypes.byref(nl)))
    lists = [virt[off[k]:off[k + 1]].tolist() for k in range(nl.value)]
    return status, low, lists


def cholesky_ao_to_mo(ao_pairs, ao_vec, mo_coef, norms: bool = True):
    """(records, norms) of tci_cholesky_ao_to_mo: ao_pairs (n_pairs x 2, 0-based
    AO indices), ao_vec (n_chol x n_pairs), mo_coef (n_ao x n_mo, C[mu][p])."""
    lib = load_library()
    pairs = np.ascontiguousarray(np.asarray(ao_pairs, dtype=np.int32).reshape(-1, 2))
    vec = np.ascontiguousarray(np.asarray(ao_vec, dtype=np.float64))
    C = np.asarray(mo_coef, dtype=np.float64)
    n_ao, n_mo = C.shape
    cf = np.ascontiguousarray(C.T)  # column-major C: C[mu + n_ao p]
    n_chol = vec.shape[0]
    n_rec = n_mo * (n_mo + 1) // 2
    rec = np.zeros((n_rec, n_chol))
    nrm = np.zeros(n_rec) if norms else None
    _check(lib.tci_cholesky_ao_to_mo(n_ao, n_mo, n_chol, pairs.shape[0], pairs.ctypes.data, vec.ctypes.data,
                                     cf.ctypes.data, rec.ctypes.data, nrm.ctypes.data if norms else None))
    return rec, nrm


def _fsn(n: int, mult: int = 1) -> int:
    if n < 0 or (n + mult - 1) % 2 or mult - 1 > n:
        return 0
    s2 = mult - 1
    return mult * math.factorial(n) // (math.factorial((n + s2) // 2 + 1) * math.factorial((n - s2) // 2))


def spin_tables(n_open: int, spin_mult: int = 1):
    """dict of tci_spin_tables: sft ((n_open-1) x nf), spin_matrix
    ((n_open+1) x nf), cycles and transpositions ({(l, k): nf x nf} for l < k,
    {(i, j): nf x nf} for i > j, 1-based shells as in the reference)."""
    lib = load_library()
    nf = _fsn(n_open, spin_mult)
    nc = n_open * (n_open - 1) // 2
    sft = np.zeros((max(n_open - 1, 0), nf), dtype=np.int32)
    sm = np.zeros((n_open + 1, nf), dtype=np.int32)
    cyc = np.zeros((nc, nf, nf))
    tr = np.zeros((nc, nf, nf))
    _check(lib.tci_spin_tables(n_open, spin_mult, sft.ctypes.data, sm.ctypes.data, cyc.ctypes.data, tr.ctypes.data))
    lab = lambda a, b: (b - 1) * (b - 2) // 2 + a - 1
    cycles = {(l, k): cyc[lab(l, k)] for k in range(2, n_open + 1) for l in range(1, k)}
    transp = {(i, j): tr[lab(j, i)] for i in range(2, n_open + 1) for j in range(1, i)}
    return {"nf": nf, "sft": sft, "spin_matrix": sm, "cycles": cycles, "transpositions": transp}


def contracted_cycles(n_open: int, records, spin_mult: int = 1):
    """List of the matrices of tci_contracted_cycles for records (i, j, k, l,
    d1, d2, left, right), 1-based shells; each fsn(left) x fsn(right)."""
    lib = load_library()
    rec = np.ascontiguousarray(np.asarray(records, dtype=np.int32).reshape(-1, 8))
    shapes = [(_fsn(int(r[6]), spin_mult), _fsn(int(r[7]), spin_mult)) for r in rec]
    off = np.zeros(len(rec), dtype=np.int64)
    tot = 0
    for q, (a, b) in enumerate(shapes):
        off[q] = tot
        tot += a * b
    out = np.zeros(max(tot, 1))
    _check(lib.tci_contracted_cycles(n_open, spin_mult, len(rec), rec.ctypes.data, off.ctypes.data, out.ctypes.data))
    return [out[o:o + a * b].reshape(a, b) for o, (a, b) in zip(off, shapes)]


def ladder_records(n_int: int, n_v: int, head_weight, num_singles, lists, ignorable=None, spin_mult: int = 1):
    """n_records x 7 int32 array of tci_ladder_records; lists: per walk the
    sorted 0-based virtuals; ignorable: n_v x n_v flags [c][a] or None."""
    lib = load_library()
    hw = np.ascontiguousarray(np.asarray(head_weight, dtype=np.int64))
    ns = np.ascontiguousarray(np.asarray(num_singles, dtype=np.int32))
    off = np.zeros(len(lists) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    lv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int32) for x in lists]) if off[-1] else
                              np.zeros(1, dtype=np.int32))
    ig = None if ignorable is None else np.ascontiguousarray(np.asarray(ignorable, dtype=np.uint8))
    n = ctypes.c_int64(0)
    args = (n_int, n_v, spin_mult, len(lists), hw.ctypes.data, ns.ctypes.data, off.ctypes.data, lv.ctypes.data,
            None if ig is None else ig.ctypes.data)
    _check(lib.tci_ladder_records(*args, 0, None, ctypes.byref(n)))
    rec = np.zeros((max(n.value, 1), 7), dtype=np.int32)
    if n.value:
        _check(lib.tci_ladder_records(*args, n.value, rec.ctypes.data, ctypes.byref(n)))
    return rec[:n.value]
