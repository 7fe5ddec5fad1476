"""The device-resident step on the exact decaying Stokes modes of the channel (FP64 and,
for the shear mode, the mixed form), started from their exact history levels
(nstart = 1): the same closed-form bars as the oracle's (third order and the scalar BDF3
recurrence for the shear mode on every boundary path; second order for the no-slip
pressure mode under the Zang-Hussaini walls, first order under bctype 0 and towards the
free-flow-ceiling mode), and the device's error equal to the oracle's (1e-6 relative) at
the finest dt.
The modes come from tests/stokes_modes.py (numpy/scipy only)."""

import math

import numpy as np
import pytest

from lsbflow_amd.core import Config, Context
from tests import stokes_modes as SM
from tests.test_oracle_stokes_modes import NU, T, rel, run_pressure

pytestmark = pytest.mark.gpu


def device(dt, bctype, top_bc, nz, precision=0, nx=16, ny=8):
    cfg = Config(nx=nx, ny=ny, nz=nz, mapping=3, zlen=2.0, xlen=2 * math.pi, ylen=math.pi, xnu=NU, bctype=bctype,
                 top_bc=top_bc, samp=0.0, dt=dt, nstart=1, dct_exact=1, precision=precision)
    c = Context(cfg)
    z = np.zeros(nz + 1)
    c.set_paral(z, z, z)
    return c


def run_shear_dev(dt, bctype, top_bc, precision=0, nx=16, ny=8):
    c = device(dt, bctype, top_bc, 32, precision, nx, ny)
    try:
        v0, lam = SM.shear_mode(c.grid()["zpts"], c.ny, c.nx, NU)
        zr = np.zeros_like(v0)
        (h1,), (h2,) = SM.exact_history([v0], lam, dt)
        un, unm = np.zeros((3,) + v0.shape), np.zeros((3,) + v0.shape)
        un[1], unm[1] = h1, h2
        c.state_set(zr, v0, zr, unif=un, unmif=unm)
        n = int(round(T / dt))
        c.timestep(n)
        g = c.state_get()
    finally:
        c.close()
    ex = v0 * math.exp(-lam * n * dt)
    rec = v0 * SM.bdf3_amplitude(lam, dt, n)
    return rel(g["vf"], ex), rel(g["vf"], rec), max(float(np.abs(g[k]).max()) for k in ("uf", "wf", "pf"))


def run_pressure_dev(dt, bctype, top_bc, kind):
    c = device(dt, bctype, top_bc, 48)
    try:
        u0, w0, p0, lam = SM.pressure_mode(c.grid()["zpts"], c.ny, c.nx, NU, kind)
        zr = np.zeros_like(u0)
        (hu1, hw1, hp1), (hu2, hw2, _) = SM.exact_history([u0, w0, p0], lam, dt)
        un, unm = np.zeros((3,) + u0.shape), np.zeros((3,) + u0.shape)
        un[0], un[2], unm[0], unm[2] = hu1, hw1, hu2, hw2
        c.state_set(u0, zr, w0, pf=p0, pnf=hp1, unif=un, unmif=unm)
        n = int(round(T / dt))
        c.timestep(n)
        g = c.state_get()
    finally:
        c.close()
    f = math.exp(-lam * n * dt)
    return rel(np.stack([g["uf"], g["wf"]]), np.stack([u0 * f, w0 * f])), g


@pytest.mark.parametrize("bctype,top_bc", [(0, 0), (3, 0), (3, 1), (3, 2)])
def test_shear_mode_third_order_on_device(bctype, top_bc):
    errs = []
    for dt in (0.2, 0.1, 0.05):
        e, erec, others = run_shear_dev(dt, bctype, top_bc)
        assert erec < 1e-12, (dt, erec)
        assert others < 1e-300
        errs.append(e)
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"device shear mode bctype {bctype} top_bc {top_bc}: errors {errs}, orders {orders}", flush=True)
    assert all(2.9 < q < 3.1 for q in orders), orders


def test_shear_mode_mixed_form():
    """The mixed form's FP32 nonlinear chain acts on an exactly zero nonlinear term here, so
    it advances the mode exactly as the FP64 form (recurrence to 1e-12); 32 x 16 horizontal
    grid, a length of the register-resident FP32 transforms."""
    e, erec, _ = run_shear_dev(0.1, 3, 2, precision=1, nx=32, ny=16)
    assert erec < 1e-12, erec


@pytest.mark.parametrize("bctype,top_bc,kind,lo,hi", [
    (3, 1, "noslip", 1.9, 2.2),
    (0, 0, "noslip", 0.9, 1.1),
    (3, 0, "freeflow", 0.95, 1.1),
])
def test_pressure_mode_order_on_device(bctype, top_bc, kind, lo, hi):
    errs = []
    for dt in (0.1, 0.05, 0.025):
        e, g = run_pressure_dev(dt, bctype, top_bc, kind)
        errs.append(e)
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"device {kind} pressure mode bctype {bctype} top_bc {top_bc}: errors {errs}, orders {orders}", flush=True)
    assert all(lo < q < hi for q in orders), orders
    # the device's error equals the oracle's at the finest dt
    eo = run_pressure(0.025, bctype, top_bc, kind)
    assert abs(errs[-1] - eo) < 1e-6 * eo, (errs[-1], eo)
