"""The Fortran binding (tigerci_amd/fortran/tigersigma.f90) compiles with the
ROCm flang under the reference's integer kind (-fdefault-integer-8), links
against the C-ABI library and round-trips a validation error without a GPU;
its bind(c) types have the sizes of the C structs."""
import ctypes
import os
import shutil
import subprocess

import pytest

import tigerci_amd as tci

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")
def test_fortran_binding_links_and_reports(tmp_path):
    lib = tci.load_library()
    libdir = os.path.dirname(lib._name)
    mod = os.path.join(ROOT, "tigerci_amd", "fortran", "tigersigma.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "abi_probe.f90")
    exe = str(tmp_path / "abi_probe")
    subprocess.run([FLANG, "-fdefault-integer-8", "-J", str(tmp_path), "-c", mod, "-o", str(tmp_path / "m.o")],
                   check=True, cwd=tmp_path)
    subprocess.run([FLANG, "-fdefault-integer-8", "-I", str(tmp_path), prog, str(tmp_path / "m.o"),
                    f"-L{libdir}", "-ltigersigma", f"-Wl,-rpath,{libdir}", "-o", exe], check=True, cwd=tmp_path)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    assert out[0].split() == ["status", "1"]
    assert "must be positive" in out[1]
    sizes = [int(v) for v in out[2].split()[1:]]
    assert sizes == [ctypes.sizeof(tci._Problem), ctypes.sizeof(tci._Info), ctypes.sizeof(tci._DavParams),
                     ctypes.sizeof(tci._DavResult)]
    # ACPF-2, 14 correlated electrons: the g values Acetylene_SuperMol_MRACPF2.output:745-749 prints
    acpf = out[3].split()
    assert acpf[0] == "acpf2" and acpf[1] == "0"
    assert [float(v) for v in acpf[2:]] == [1.0, 1.0, 0.27473, 0.27473, 0.14286]
