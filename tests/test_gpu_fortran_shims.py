"""Fortran callers on the device through the drop-in modules (lsbflow_amd/fortran/
lsb_shims.f90: modhorfft, derivatives, helmholtz, sgs_models, filters, dealias with the
reference's names and arguments): tests/fortran/shim_check.f90 creates a context, then
calls horfft, dfdz and solve_helmholtz exactly as the reference's driver does (the
Helmholtz operators taken from the context as the driver's mats module holds them) and
the inverse horfft; tests/fortran/shim_check2.f90 calls dfdx, dfdy, divuv,
solve_nhom_helmholtz, sigma_model, filter_q_3pt, setup_dealiasing, dealiasing_2d_fftw and
norm.  Each result against the oracle on the same inputs.  tests/fortran/step_check.f90
drives the device-resident step (tier B) from Fortran: its state after the steps equals,
bit for bit, the same steps driven through the Python binding (which the step tests hold
to the oracle)."""

import os
import shutil
import subprocess

import ctypes as C

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd.core import Config, setup_case

from oracle import Oracle, OracleConfig
from lsbflow_amd import _lib
from tests.conftest import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(FC is None, reason="amdflang not available")]


def build_shim_program(tmp_path, program="shim_check"):
    """Compile the interface module, the drop-in modules and the caller; link the library."""
    libdir = os.path.dirname(_lib.LIB_PATH)
    fdir = os.path.join(ROOT, "lsbflow_amd", "fortran")
    objs = []
    for name in ("lsb_capi", "lsb_shims"):
        obj = str(tmp_path / f"{name}.o")
        subprocess.run([FC, "-c", os.path.join(fdir, f"{name}.f90"), "-o", obj, "-I", str(tmp_path),
                        "-module-dir", str(tmp_path)], check=True, cwd=tmp_path)
        objs.append(obj)
    exe = str(tmp_path / program)
    subprocess.run([FC, os.path.join(ROOT, "tests", "fortran", program + ".f90"), *objs, "-I", str(tmp_path),
                    "-o", exe, "-L", libdir, "-llsbflow_amd", f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    return exe


def test_fortran_driver_calls_through_the_drop_in_modules(tmp_path):
    exe = build_shim_program(tmp_path)
    nx, ny, nz = 32, 16, 40
    o = Oracle(OracleConfig(nx=nx, ny=ny, nz=nz, mapping=3, zlen=2.0, xlen=2 * np.pi, ylen=np.pi, dealiasflag=2,
                            xnu=1e-3, dt=1e-3))
    nzp, nxpp = o.nzp, o.nxpp
    rng = np.random.default_rng(31)
    ph = np.zeros((nzp, ny, nxpp))
    ph[..., :nx] = rng.standard_normal((nzp, ny, nx))
    q = np.zeros((nzp, ny, nxpp))
    q[..., :nx] = rng.standard_normal((nzp, ny, nx))
    rhs = o.horfft_fwd(q)
    sigma = 1.0 / (1e-3 * 1e-3)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        np.array([sigma], dtype=np.float64).tofile(fh)
        ph.tofile(fh)
        rhs.tofile(fh)
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.float64)
    ns, npf = nzp * ny * nx, nzp * ny * nxpp
    sp, dz, sol, back = (got[:ns].reshape(nzp, ny, nx), got[ns:2 * ns].reshape(nzp, ny, nx),
                         got[2 * ns:3 * ns].reshape(nzp, ny, nx), got[3 * ns:3 * ns + npf].reshape(nzp, ny, nxpp))
    sp_ref = o.horfft_fwd(ph)
    errs = {"horfft -1": rel_l2(sp, sp_ref), "dfdz": rel_l2(dz, o.dfdz(sp_ref)),
            "solve_helmholtz": rel_l2(sol, o.solve_helmholtz(rhs, sigma)),
            "horfft +1": rel_l2(back[..., :nx], o.horfft_inv(sp_ref)[..., :nx])}
    print("Fortran caller through the drop-in modules vs oracle:", {k: f"{v:.2e}" for k, v in errs.items()},
          flush=True)
    for k, e in errs.items():
        assert e < 1e-12, (k, e)
    assert np.all(back[..., nx:] == 0.0)


def test_fortran_driver_sgs_filters_dealias_through_the_drop_in_modules(tmp_path):
    exe = build_shim_program(tmp_path, "shim_check2")
    nx, ny, nz = 32, 16, 40
    o = Oracle(OracleConfig(nx=nx, ny=ny, nz=nz, mapping=3, zlen=2.0, xlen=2 * np.pi, ylen=np.pi, dealiasflag=1,
                            dealias_ref_compat=1, sgsflag=1, sgsmodel=2, xnu=1e-3, dt=1e-3, a=1.0, b=2.0))
    nzp, nxpp = o.nzp, o.nxpp
    rng = np.random.default_rng(37)
    spec = [rng.standard_normal((nzp, ny, nx)) for _ in range(4)]
    uf, vf, rhs, bc = spec
    g = rng.standard_normal((3, 3, nzp, ny, nxpp))                     # g[l][m] = du_l/dx_m
    g[..., nx:] = 0.0
    pa, pb = (np.zeros((nzp, ny, nxpp)) for _ in range(2))
    pa[..., :nx] = rng.standard_normal((nzp, ny, nx))
    pb[..., :nx] = rng.standard_normal((nzp, ny, nx))
    sigma = 250.0
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        np.array([sigma], dtype=np.float64).tofile(fh)
        for a in (uf, vf, rhs, bc):
            a.tofile(fh)
        np.ascontiguousarray(np.swapaxes(g, 0, 1)).tofile(fh)          # duidxj(:, :, :, l, m)
        pa.tofile(fh)
        pb.tofile(fh)
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    nlg = np.frombuffer(raw[:8], dtype=np.int32)
    got = np.frombuffer(raw[8:], dtype=np.float64)
    ns, npf = nzp * ny * nx, nzp * ny * nxpp
    sizes = [ns, ns, ns, ns, 9 * npf, npf, npf, ns, ns, ns, ns]
    parts, o0 = [], 0
    for n in sizes:
        parts.append(got[o0:o0 + n])
        o0 += n
    assert o0 == got.size
    dx, dy, dv, sol = (parts[i].reshape(nzp, ny, nx) for i in range(4))
    tau = np.swapaxes(parts[4].reshape(3, 3, nzp, ny, nxpp), 0, 1)
    pf, pab = parts[5].reshape(nzp, ny, nxpp), parts[6].reshape(nzp, ny, nxpp)
    normed = [parts[7 + i].reshape(nzp, ny, nx) for i in range(4)]
    assert tuple(int(v) for v in nlg) == tuple(o.dealias_sizes())
    assert np.array_equal(dx, o.dfdx(uf)) and np.array_equal(dy, o.dfdy(uf))
    refn = [uf.copy(), vf.copy(), o.dfdx(uf), o.dfdy(uf)]
    o.norm(*refn)
    for a, b in zip(normed, refn):
        assert np.array_equal(a, b)
    errs = {"divuv": rel_l2(dv, o.divuv(uf, vf)),
            "solve_nhom_helmholtz": rel_l2(sol, o.solve_nhom_helmholtz(rhs, bc, sigma)),
            "sigma_model": rel_l2(tau[..., :nx], o.sigma_model(g)[..., :nx]),
            "filter_q_3pt": rel_l2(pf[..., :nx], o.filter_q_3pt(pa, 1.0, 2.0, 1)[..., :nx]),
            "dealiasing_2d_fftw": rel_l2(pab[..., :nx], o.dealiasing_2d_fftw(pa, pb, True)[..., :nx])}
    print("Fortran caller (SGS, filters, dealiasing) through the drop-in modules vs oracle:",
          {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    bars = {"divuv": 1e-15, "solve_nhom_helmholtz": 1e-11, "sigma_model": 1e-11, "filter_q_3pt": 1e-14,
            "dealiasing_2d_fftw": 1e-12}
    for k, e in errs.items():
        assert e < bars[k], (k, e)


@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_device_resident_steps(tmp_path, precision):
    exe = build_shim_program(tmp_path, "step_check")
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, xnu=2e-4, samp=0.3, sgsflag=1, sgsmodel=2)
    nsteps = 6
    cfg = Config.from_case(case, precision=precision).to_c()
    ub, ubp, ubpp = L.cases.base_flow(case)
    u, v, w = L.cases.initial_condition(case)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(bytes(cfg))
        np.array([nsteps], dtype=np.int32).tofile(fh)
        for a in (ub, ubp, ubpp, u, v, w):
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    istep = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    t = float(np.frombuffer(raw[4:12], dtype=np.float64)[0])
    got = np.frombuffer(raw[12:], dtype=np.float64)
    c = setup_case(case, precision=precision)
    try:
        c.state_set(u, v, w)
        c.timestep(nsteps)
        ref = c.state_get()
    finally:
        c.close()
    n = u.size
    keys = [("uf", 1), ("vf", 1), ("wf", 1), ("pf", 1), ("pnf", 1), ("unif", 3), ("unmif", 3), ("nlnif", 3),
            ("nlnmif", 3)]
    o0 = 0
    for k, m in keys:
        a = got[o0:o0 + m * n].reshape(ref[k].shape)
        o0 += m * n
        assert np.array_equal(a, ref[k]), k
    assert o0 == got.size
    assert istep == nsteps and abs(t - nsteps * case.dt) < 1e-12
    print(f"Fortran tier-B driver, precision {precision}: {nsteps} steps bit-identical to the Python-driven run",
          flush=True)
