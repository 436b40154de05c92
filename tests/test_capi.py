"""The C-ABI library loads, exports every symbol include/flint_amd.h declares, and rejects bad
arguments on host paths (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from flint_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            out |= set(re.findall(r"\b(flint_\w+)\s*\(", txt))
    return out


def test_library_exports_declared_symbols():
    L = native.lib()
    syms = declared_symbols()
    assert len(syms) >= 8
    for s in sorted(syms):
        assert hasattr(L, s), s


def test_load_missing_folder_is_io_error(tmp_path):
    L = native.lib()
    out = C.c_void_p()
    rc = L.flint_mech_load(str(tmp_path).encode(), b"", C.byref(out))
    assert rc == 3 and not out
    assert b"phase.txt" in L.flint_last_error()


def test_create_rejects_bad_sizes():
    L = native.lib()
    out = C.c_void_p()
    W = np.ones(5)
    t = np.zeros(10)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = L.flint_mech_create(b"WD", 0, p(W), 1, 2, 3, p(t), p(t), 0, None, 0, None, 1, 2, p(t), p(t), 0, None,
                             C.byref(out))
    assert rc == 1
    rc = L.flint_mech_create(b"no-such-mechanism", 5, p(W), 2, 1, 3, p(t), p(t), 0, None, 0, None, 1, 2, p(t), p(t), 0,
                             None, C.byref(out))
    assert rc == 1  # Tk_max <= Tk_min


def test_fortran_drop_in_reports_missing_folder(tmp_path):
    """The Fortran module's read_idealgas_thermo returns FLINT's nonzero ios with the library's
    message for a folder without phase.txt (no GPU call is made)."""
    import subprocess
    exe = os.path.join(ROOT, "flint_amd", "fortran", "_build", "drop_in_check")
    if not os.path.exists(exe):
        pytest.skip("Fortran module not built (no flang)")
    r = subprocess.run([exe, str(tmp_path / "nothing"), "x", "y"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "phase.txt" in r.stdout
