"""The Fortran side of the boundary: lsbflow_amd/fortran/lsb_capi.f90 (iso_c_binding
interfaces of every entry point) compiles with amdflang, links against
liblsbflow_amd.so, and a Fortran program calling the host-side setup routines
through it gets the oracle's results (no GPU calls)."""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd import _lib
from tests.test_oracle_linstab import lintest_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

pytestmark = pytest.mark.skipif(FC is None, reason="amdflang not available")


def test_interface_module_declares_every_entry_point():
    src = open(os.path.join(ROOT, "lsbflow_amd", "fortran", "lsb_capi.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(lsb_[a-z0-9_]+)"\)', src))
    hdr = open(os.path.join(ROOT, "include", "lsbflow.h")).read()
    declared = set(re.findall(r"\b(lsb_[a-z0-9_]+)\s*\(", hdr))
    assert bound == declared


def test_fortran_program_through_the_binding(tmp_path):
    libdir = os.path.dirname(_lib.LIB_PATH)
    mod = os.path.join(ROOT, "lsbflow_amd", "fortran", "lsb_capi.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "binding_check.f90")
    exe = str(tmp_path / "binding_check")
    subprocess.run([FC, "-c", mod, "-o", str(tmp_path / "lsb_capi.o"), "-module-dir", str(tmp_path)],
                   check=True, cwd=tmp_path)
    subprocess.run([FC, prog, str(tmp_path / "lsb_capi.o"), "-I", str(tmp_path), "-o", exe,
                    "-L", libdir, "-llsbflow_amd", f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)

    o = Oracle(OracleConfig(nx=8, ny=4, nz=40, mapping=3, zlen=2.0))
    nzp, n1, n2 = o.nzp, 3, 2
    y = np.random.default_rng(5).standard_normal((nzp, n2, n1))     # Fortran y(n1, n2, nzp)
    ls = lintest_case(1805.0, 0.5)                                  # an Orr-Sommerfeld case (nzp = 65)
    guess = ls.alpha * 0.4
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as fh:
        np.array([nzp, n1, n2], dtype=np.int32).tofile(fh)
        np.array([o.xl], dtype=np.float64).tofile(fh)
        o.gf.astype(np.float64).tofile(fh)
        y.astype(np.float64).tofile(fh)
        np.array([ls.nzp], dtype=np.int32).tofile(fh)
        for a in (ls.d.T, ls.d2.T, ls.ubar, ls.ubarp, ls.ubarpp, ls.zpts):   # column-major matrices
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
        np.array([1e-3, 1e-10, ls.alpha, guess, 0.0], dtype=np.float64).tofile(fh)
    out = tmp_path / "out.bin"
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    st = np.frombuffer(raw[:20], dtype=np.int32)
    assert list(st[:3]) == [0, 0, 0]
    assert st[3] == -1 and "nx" in r.stdout             # LSB_E_ARG with the library's message
    assert st[4] == 18 * 4 + 17 * 8                       # sizeof(lsb_config) seen from Fortran
    nzm = nzp - 2
    sizes = [nzp * nzp] * 3 + [n1 * n2 * nzp, nzm, nzm * nzm, nzm * nzm]
    nb = 20 + 8 * sum(sizes)
    vals = np.frombuffer(raw[20:nb], dtype=np.float64)
    parts = np.split(vals, np.cumsum(sizes)[:-1])
    # the Orr-Sommerfeld mode through complex Fortran arguments equals the C / Python call's
    assert np.frombuffer(raw[nb:nb + 4], dtype=np.int32)[0] == 0
    osv = np.frombuffer(raw[nb + 4:], dtype=np.complex128)
    assert osv.size == 1 + 3 * ls.nzp
    lu, lv, lw, lom = L.solve_orrsomm(ls.d, ls.d2, ls.ubar, ls.ubarp, ls.ubarpp, ls.zpts, 1e-3, 1e-10, ls.alpha,
                                      0.0, guess, 0)
    assert osv[0] == lom
    assert np.array_equal(osv[1:1 + ls.nzp], lu) and np.array_equal(osv[1 + 2 * ls.nzp:], lw)
    assert not np.any(osv[1 + ls.nzp:1 + 2 * ls.nzp])
    dor, d, d2 = (p.reshape(nzp, nzp, order="F") for p in parts[:3])
    assert np.array_equal(dor, o.dor)
    assert np.max(np.abs(d2 - o.d2)) < 1e-14 * np.max(np.abs(o.d2))
    dydx = parts[3].reshape(nzp, n2, n1)
    ref = o.dct_diff(y, 1)
    assert np.max(np.abs(dydx - ref)) < 1e-12 * np.max(np.abs(ref))
    lam = parts[4]
    e = parts[5].reshape(nzm, nzm, order="F")
    ei = parts[6].reshape(nzm, nzm, order="F")
    D = o.d2[1:o.nz, 1:o.nz]
    assert np.max(np.abs(e @ np.diag(lam) @ ei - D)) < 1e-13 * np.max(np.abs(D))
    assert np.max(np.abs(np.sort(lam) - np.sort(o.diag)) / np.abs(np.sort(o.diag))) < 1e-10


def test_drop_in_modules_compile_and_link(tmp_path):
    """The drop-in modules (lsbflow_amd/fortran/lsb_shims.f90) and a caller that uses them
    as the reference's driver does compile with amdflang and link against the library
    (the run is the GPU test tests/test_gpu_fortran_shims.py)."""
    from tests.test_gpu_fortran_shims import build_shim_program
    exe = build_shim_program(tmp_path)
    assert os.path.exists(exe)
    exe2 = build_shim_program(tmp_path, "shim_check2")
    assert os.path.exists(exe2)
    assert os.path.exists(build_shim_program(tmp_path, "step_check"))
    for mod in ("modhorfft", "derivatives", "helmholtz", "sgs_models", "filters", "dealias"):
        assert os.path.exists(tmp_path / f"{mod}.mod"), mod
