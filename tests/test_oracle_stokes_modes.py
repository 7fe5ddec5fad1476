"""Pin the oracle's integrated step (navierstokes.f90:1157-1496, both boundary paths) to
exact decaying Stokes modes of the channel, started from their exact history levels
(nstart = 1), on a 16 x 8 horizontal grid with nu = 0.05 over t = 2:

* shear mode (no pressure, no nonlinear term): the amplitude equals the scalar BDF3
  recurrence to 1e-12 and the error against exp(-lam t) falls at third order as dt
  halves, for bctype 0 and for bctype 3 with each ceiling condition (top_bc 0, 1, 2);
* pressure mode with no-slip walls under the Zang-Hussaini boundary values at both walls
  (bctype 3, top_bc 1): second order, set by the pressure extrapolation 2 p^n - p^(n-1)
  of the boundary values (navierstokes.f90:1321-1333);
* the same mode under bctype 0, whose homogeneous solves drop those boundary values:
  first order (the splitting error of the projection);
* pressure mode with the free-flow ceiling W = W''' = 0 (top_bc 0, the D^2(u, v) -> 0
  condition of navierstokes.f90:1335-1347): first order towards that mode.

The modes and the recurrence come from tests/stokes_modes.py (numpy/scipy only).
"""

import math

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
from tests import stokes_modes as SM

NU, T = 0.05, 2.0


def oracle(dt, bctype, top_bc, nz):
    return Oracle(OracleConfig(nx=16, ny=8, nz=nz, mapping=3, zlen=2.0, xlen=2 * math.pi, ylen=math.pi, xnu=NU,
                               bctype=bctype, top_bc=top_bc, samp=0.0, dt=dt, nstart=1, dct_exact=1))


def run_shear(dt, bctype, top_bc):
    o = oracle(dt, bctype, top_bc, 32)
    v0, lam = SM.shear_mode(o.zpts, o.ny, o.nx, NU)
    zr = np.zeros_like(v0)
    st = o.new_state(zr, v0, zr)
    (h1,), (h2,) = SM.exact_history([v0], lam, dt)
    st["unif"][1], st["unmif"][1] = h1, h2
    n = int(round(T / dt))
    for _ in range(n):
        o.timestep(st)
    ex = v0 * math.exp(-lam * n * dt)
    rec = v0 * SM.bdf3_amplitude(lam, dt, n)
    others = max(float(np.abs(st[k]).max()) for k in ("uf", "wf", "pf"))
    return rel(st["vf"], ex), rel(st["vf"], rec), others


def run_pressure(dt, bctype, top_bc, kind):
    o = oracle(dt, bctype, top_bc, 48)
    u0, w0, p0, lam = SM.pressure_mode(o.zpts, o.ny, o.nx, NU, kind)
    zr = np.zeros_like(u0)
    st = o.new_state(u0, zr, w0)
    (hu1, hw1, hp1), (hu2, hw2, _) = SM.exact_history([u0, w0, p0], lam, dt)
    st["unif"][0], st["unif"][2], st["unmif"][0], st["unmif"][2] = hu1, hw1, hu2, hw2
    st["pf"], st["pnf"] = p0.copy(), hp1
    n = int(round(T / dt))
    for _ in range(n):
        o.timestep(st)
    f = math.exp(-lam * n * dt)
    return rel(np.stack([st["uf"], st["wf"]]), np.stack([u0 * f, w0 * f]))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_modes_are_spectral_layout_of_the_fields():
    """The helper's spectral arrays are the oracle's transform of the physical fields."""
    o = oracle(0.1, 0, 0, 32)
    v0, _ = SM.shear_mode(o.zpts, o.ny, o.nx, NU)
    x = np.arange(o.nx) * (o.cfg.xlen / o.nx)
    ph = np.zeros((o.nzp, o.ny, o.nxpp))
    ph[:, :, : o.nx] = np.sin(x)[None, None, :] * np.sin(math.pi * o.zpts / 2)[:, None, None]
    assert np.max(np.abs(o.horfft_fwd(ph) - v0)) < 1e-15
    for kind in ("noslip", "freeflow"):
        u0, w0, p0, lam = SM.pressure_mode(o.zpts, o.ny, o.nx, NU, kind, amp=1.0)
        # continuity: i kx u + dw/dz = 0, and the boundary values of the kind
        div = o.dfdx(u0) + o.dfdz(w0)
        assert np.max(np.abs(div)) < 1e-9
        assert np.max(np.abs(w0[[0, -1]])) < 1e-12 and np.max(np.abs(u0[-1])) < 1e-12
        if kind == "noslip":
            assert np.max(np.abs(u0[0])) < 1e-12
        else:
            assert np.max(np.abs(o.d2fdz2(u0)[0])) < 1e-6 * np.max(np.abs(o.d2fdz2(u0)))


@pytest.mark.parametrize("bctype,top_bc", [(0, 0), (3, 0), (3, 1), (3, 2)])
def test_shear_mode_third_order(bctype, top_bc):
    errs = []
    for dt in (0.2, 0.1, 0.05):
        e, erec, others = run_shear(dt, bctype, top_bc)
        assert erec < 1e-12, (dt, erec)
        assert others == 0.0
        errs.append(e)
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"shear mode bctype {bctype} top_bc {top_bc}: errors {errs}, orders {orders}", flush=True)
    assert all(2.9 < q < 3.1 for q in orders), orders


@pytest.mark.parametrize("bctype,top_bc,kind,order,lo,hi", [
    (3, 1, "noslip", 2, 1.9, 2.2),
    (0, 0, "noslip", 1, 0.9, 1.1),
    (3, 0, "freeflow", 1, 0.95, 1.1),
])
def test_pressure_mode_order(bctype, top_bc, kind, order, lo, hi):
    errs = [run_pressure(dt, bctype, top_bc, kind) for dt in (0.1, 0.05, 0.025)]
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"{kind} pressure mode bctype {bctype} top_bc {top_bc}: errors {errs}, orders {orders}", flush=True)
    assert all(lo < q < hi for q in orders), orders
    assert errs[-1] < (2e-5 if order == 2 else 5e-3), errs
