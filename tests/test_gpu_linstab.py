"""Linear-stability acceptance on the device (orrsomm.f90, main.f90:63-65): the most
unstable Orr-Sommerfeld mode of the Blasius layer at Re_d = 1805, alpha* = 0.5 (the first
case of run/energy.dat), injected at kx = alpha with energy 1e-10, integrated 25 steps of
dt = cfl dx/max ubar (cfl = 0.1; the 0.0360036 of run/output_lintest.txt) by the
device-resident step with lintest after each step.

Bars: the device's energies and growth rates against the oracle integrating the same
state (1e-6 relative on the averaged growth rate <dE/dt>); <dE/dt> against the mode's
omega_i within 0.1% (the reference reports 0.099% for this case); and against the
reference's own <dE/dt> (-0.0115823, run/energy.dat) within 1e-3 relative, the room left by
the sweep's unrecorded dealiasing flag and spanwise size."""

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd.core import Config, Context
from tests.test_oracle_linstab import FIX, lintest_case

pytestmark = pytest.mark.gpu


def test_orr_sommerfeld_growth_rate_on_device():
    row = FIX["energy_dat"][0]
    assert (row["re_d"], row["alphastar"]) == (1805.0, 0.5)
    o = lintest_case(1805.0, 0.5)
    al = o.alpha
    u, v, w, om = L.solve_orrsomm(o.d, o.d2, o.ubar, o.ubarp, o.ubarpp, o.zpts, 1e-3, 1e-10, al, 0.0, al * 0.4, 0)
    dt = 0.1 * o.dx / max(np.max(np.abs(o.ubar)), 1.0)
    assert abs(dt - 0.0360036) < 1e-6
    o.cfg.dt = dt
    z = np.zeros((o.nzp, o.ny, o.nx))
    uf, vf, wf = z.copy(), z.copy(), z.copy()
    uf[:, 0, 2], uf[:, 0, 3] = u.real, u.imag     # orrsomm.f90 / init.f90:622-627: mode kx = alpha
    wf[:, 0, 2], wf[:, 0, 3] = w.real, w.imag
    c = o.cfg
    cfg = Config(nx=c.nx, ny=c.ny, nz=c.nz, mapping=2, bctype=0, dealiasflag=2, xlen=c.xlen, ylen=c.ylen,
                 zlen=c.zlen, x0=c.x0, u0=c.u0, xnu=c.xnu, blfac=c.blfac, dt=dt)
    n = 25
    ctx = Context(cfg)
    try:
        ctx.set_paral(o.ubar, o.ubarp, o.ubarpp, o.wbar, o.wbarpp)
        ctx.state_set(uf, vf, wf)
        e1, e2, g1, g2 = (np.zeros(n) for _ in range(4))
        for _ in range(n):
            ctx.timestep(1)
            ctx.state_lintest(e1, e2, g1, g2)
    finally:
        ctx.close()
    st = o.new_state(uf, vf, wf)
    oe1, oe2, og1, og2 = (np.zeros(n) for _ in range(4))
    for i in range(1, n + 1):
        o.timestep(st)
        o.lintest(st["uf"], st["vf"], st["wf"], i, dt, oe1, oe2, og1, og2)
    assert np.max(np.abs(e2 - oe2) / oe2) < 1e-9
    avg = np.sum(g2) / (n - 1)            # write_linstab_results, orrsomm.f90:77
    oavg = np.sum(og2) / (n - 1)
    assert abs(avg - oavg) < 1e-6 * abs(oavg), (avg, oavg)
    assert abs(avg - om.imag) < 1e-3 * abs(om.imag), (avg, om)
    assert abs(avg - row["dedt"]) < 1e-3 * abs(row["dedt"]), (avg, row)
