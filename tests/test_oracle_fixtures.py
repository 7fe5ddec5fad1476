"""Pin the oracle (and the native host setup) to the data the reference ships in run/.

Fixtures in tests/golden/reference_run/ were copied by make_reference_fixtures.py
from run/grid.dat, run/paral.dat and run/output.txt of the reference.  They come
from run/input.dat: 224 x 24 x 32, mapping 2, Blasius base flow.
"""

import json
import os
import struct

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_run")


def fortran_records(path):
    data = open(path, "rb").read()
    off, out = 0, []
    while off < len(data):
        n = struct.unpack("<i", data[off:off + 4])[0]
        out.append(data[off + 4: off + 4 + n])
        assert struct.unpack("<i", data[off + 4 + n: off + 8 + n])[0] == n
        off += 8 + n
    return out


@pytest.fixture(scope="module")
def run_case():
    cfg = OracleConfig(nx=224, ny=24, nz=32, xlen=333.3333, ylen=20.0, zlen=33.3333, x0=8.3333, u0=1.0,
                       xnu=0.001, mapping=2, blfac=8.0, bctype=0, sgsflag=0, dealiasflag=2)
    o = Oracle(cfg)
    o.setup_blasius()
    return o


@pytest.fixture(scope="module")
def meta():
    return json.load(open(os.path.join(GOLD, "output_txt.json")))


def test_grid_bit_exact(run_case):
    """setup_grid (init.f90:130-273, mapping 2) reproduces run/grid.dat bit for bit."""
    r = fortran_records(os.path.join(GOLD, "grid.dat"))
    assert struct.unpack("<3i", r[0]) == (224, 24, 33)
    x0, xl, xlen, ylen, zlen = struct.unpack("<5d", r[1])
    o = run_case
    assert (x0, xl, xlen, ylen, zlen) == (8.3333, o.xl, 333.3333, 20.0, 33.3333)
    for arr, rec in ((o.xpts, r[2]), (o.ypts, r[3]), (o.zpts, r[4]), (o.gf, r[5])):
        assert np.array_equal(arr, np.frombuffer(rec, "<f8"))


def test_native_grid_bit_exact():
    """The library's own setup_grid gives the same bits (host-side, no GPU call)."""
    import ctypes as C
    from lsbflow_amd import _lib
    cfg = L.Config(nx=224, ny=24, nz=32, mapping=2, xlen=333.3333, ylen=20.0, zlen=33.3333, x0=8.3333,
                   xnu=0.001, blfac=8.0, bctype=0)
    r = fortran_records(os.path.join(GOLD, "grid.dat"))
    zp, gf = np.frombuffer(r[4], "<f8"), np.frombuffer(r[5], "<f8")
    # build_cheb_zder_matrix on the fixture grid and compare with the oracle's d
    o = Oracle(OracleConfig(nx=16, ny=8, nz=32, mapping=2, xlen=333.3333, zlen=33.3333, x0=8.3333, xnu=0.001))
    dor, d, d2 = L.build_cheb_zder_matrix(33, 1.0, gf)
    assert np.array_equal(dor, o.dor)
    assert np.array_equal(d, o.d)
    assert np.array_equal(o.zpts, zp)
    del cfg, C, _lib


def test_eigenvalues_output_txt(run_case, meta):
    """diagonalize_matrix(d2(2:nz,2:nz)) eigenvalues vs the E14.7 print in run/output.txt."""
    ref = np.sort(np.array(meta["eigenvalues"]))
    o = run_case
    assert len(ref) == o.nzm
    mine = np.sort(o.diag)
    assert np.max(np.abs(mine - ref) / np.abs(ref)) < 5e-7
    lam, e, ei = L.diagonalize_matrix(o.d2[1:o.nz, 1:o.nz])
    assert np.max(np.abs(np.sort(lam) - ref) / np.abs(ref)) < 5e-7
    assert np.max(np.abs(np.sort(lam) - mine) / np.abs(mine)) < 1e-10


def test_paral_dat_blasius(run_case, meta):
    """blasfl + dpdz/dct_diff (init.f90:404-424) vs run/paral.dat (Newton count vs output.txt:88)."""
    r = fortran_records(os.path.join(GOLD, "paral.dat"))
    assert struct.unpack("<i", r[0])[0] == 33
    names = ["ubar", "ubarp", "ubarpp", "wbar", "wbarp", "wbarpp"]
    tol = {"ubar": 1e-13, "ubarp": 2e-12, "ubarpp": 5e-10, "wbar": 1e-15, "wbarp": 1e-14, "wbarpp": 1e-13}
    o = run_case
    for i, nm in enumerate(names):
        ref = np.frombuffer(r[i + 1], "<f8")
        assert np.max(np.abs(getattr(o, nm) - ref)) < tol[nm], nm
    assert o.blasfl_iters == meta["blasfl_iterations"]
    # dpdz through the library's host dct_diff gives the same derivatives
    ubarp = L.dct_diff(o.ubar, o.xl, o.gf, 1)
    assert np.max(np.abs(ubarp - np.frombuffer(r[2], "<f8"))) < 2e-12


def test_initial_dt(run_case, meta):
    """dt = cfl*dx/max|u+ubar| of setup_ic (init.f90:451) vs output.txt."""
    o = run_case
    dt = 0.1 * o.dx / max(np.max(np.abs(o.ubar)), 1.0)
    assert abs(dt - meta["initial_dt"]) < 1e-14 * meta["initial_dt"]


def test_logged_run_divergence(run_case, meta):
    """10 steps of the logged run: the divergence stays at the reference's round-off level
    (output.txt prints 5.7e-34 at step 10).  The logged Max|u+ubar| comes from an older
    revision without the base-flow diffusion term and is not compared."""
    o = run_case
    o.cfg.dt = meta["initial_dt"]
    z = np.zeros((o.nzp, o.ny, o.nx))
    st = o.new_state(z, z, z)
    for _ in range(10):
        o.timestep(st)
    d = o.divcheck(st["uf"], st["vf"], st["wf"])
    assert d < 1e-30
    assert meta["divergence_avg_error"][0] < 1e-30
