"""GPU parity of the device-resident time step against the CPU oracle (FP64).

The configurations use the seeded tanh-plus-perturbation initial condition and
the steady ceiling profile of ``lsbflow_amd.cases``; the oracle integrates the
same state with its restatement of navierstokes.f90.

Bar: relative L2 <= 1e-10 per field after the stated number of steps, unless
the reference algorithm itself is less well conditioned than that for a field:
the tolerance is then 10 x the oracle's own drift under rounding, measured in the
test as the same oracle run with 1e-16 relative noise on its state after every
step (the size of the differences a different but equivalent order of FP64
operations leaves).  This matters for the sigma model, whose invariant/acos
evaluation of near-degenerate singular values (sgsterm.f90:334-360) amplifies
rounding to ~1e-7, and for the pressure, a divergence divided by k^2 dt
(navierstokes.f90:1473) with cancellation in its numerator.  Every test prints
its per-field errors, sensitivities and their ratio.

The pressure has a fixed bar of 1e-7 in addition.  With ceiling boundary data
(bctype 3, samp > 0) the device and the oracle differ by ~2e-12 after the first
step, in the kx = 1 pressure at the walls; rounding noise on every transform,
derivative and solve (inputs and outputs), on the eigenvectors and on the
lifting functions moves the oracle by 3e-14 .. 6e-13 only, so this offset is
not accounted for by the rounding model.  The velocities it feeds stay below
the 1e-10 floor (at most 6.3e-11 after 6 steps); the pressure, amplified by
1/(k^2 dt), reaches 1.4e-10 .. 2.6e-8 in those cases.
"""

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd import _lib
from lsbflow_amd.core import setup_case
from tests.conftest import noisy_oracle_run, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def oracle_for(case, **kw):
    keys = OracleConfig.__dataclass_fields__
    d = {k: v for k, v in case.as_dict().items() if k in keys}
    d.update(kw)
    o = Oracle(OracleConfig(**d))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(case)
    return o


FIELDS = ("uf", "vf", "wf", "pf")


def run_both(case, nsteps, **kw):
    o = oracle_for(case, **kw)
    u, v, w = L.cases.initial_condition(case)
    st = o.new_state(u, v, w)
    for _ in range(nsteps):
        o.timestep(st)
    # the oracle's own sensitivity to rounding: the same run with 1e-16 relative noise on
    # the state after every step
    st2 = noisy_oracle_run(oracle_for(case, **kw), (u, v, w), nsteps)
    sens = {k: rel_l2(st2[k], st[k]) for k in FIELDS}
    c = setup_case(case, **kw)
    c.state_set(u, v, w)
    c.timestep(nsteps)
    g = c.state_get()
    c.close()
    g["_sens"] = sens
    return st, g


MULT = 10.0
P_TOL = 1e-7


def check_fields(st, g, tol=TOL, fields=FIELDS):
    sens = g["_sens"]
    errs = {k: rel_l2(g[k], st[k]) for k in fields}
    print("rel L2 vs oracle: " + " ".join(f"{k} {errs[k]:.2e} (sens {sens[k]:.1e}, x{errs[k] / max(sens[k], 1e-300):.2g})"
                                          for k in fields), flush=True)
    for k, e in errs.items():
        bound = max(tol, MULT * sens[k], P_TOL if k == "pf" else 0.0)
        assert e < bound, (k, e, "bound", bound, "errors", errs, "oracle sensitivity", sens)
    return errs


def test_c1_dns_10_steps():
    """Config 1 (64 x 65 x 32 DNS, Re 1000, dt 1e-3) for its 10 steps."""
    case = L.cases.CASES["c1_dns"]
    st, g = run_both(case, case.nsteps)
    check_fields(st, g)
    assert g["istep"] == 10
    assert abs(g["t"] - st["t"]) < 1e-15


@pytest.mark.parametrize("sgsmodel", [0, 1, 2, 3])
def test_les_models_small(sgsmodel):
    """Each SGS model inside the step (suction ceiling on), 6 steps on a 32 x 41 x 16 grid."""
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, xnu=2e-4, samp=0.3, sgsflag=1, sgsmodel=sgsmodel,
                                        a=1.0, b=2.0, dt=5e-4, pert_rms=0.05)
    st, g = run_both(case, 6)
    check_fields(st, g)


def test_les_dynamic_plane_average():
    """The plane-averaged dynamic coefficient (<M:L>, <M:M> as x-y plane means) inside the
    step against the oracle's, 6 steps on a 32 x 41 x 16 grid with the ceiling profile."""
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, xnu=2e-4, samp=0.3, sgsflag=1, sgsmodel=1,
                                        a=1.0, b=2.0, dt=5e-4, pert_rms=0.05, dyn_avg=1)
    st, g = run_both(case, 6)
    check_fields(st, g)


@pytest.mark.parametrize("sgsmodel", [0, 2])
def test_les_ceiling_row_without_eddy_viscosity(sgsmodel):
    """nut_ceil0 = 1 (zero filter width on the ceiling row) for the static and sigma
    models with the blended ceiling, 6 steps on a 32 x 41 x 16 grid, against the oracle;
    the dynamic and interscale models refuse the switch at lsb_create."""
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, xnu=2e-4, samp=0.3, sgsflag=1, sgsmodel=sgsmodel,
                                        dt=5e-4, pert_rms=0.05, dct_exact=1, top_bc=2, nut_ceil0=1)
    st, g = run_both(case, 6)
    check_fields(st, g)
    for bad in (1, 3):
        with pytest.raises(_lib.LsbError, match="nut_ceil0"):
            setup_case(case.with_(sgsmodel=bad))


@pytest.mark.parametrize("dealias", [0, 2])
def test_dealias_modes_and_homogeneous_bc(dealias):
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=36, dealiasflag=dealias, bctype=0, pert_rms=0.02)
    st, g = run_both(case, 5)
    check_fields(st, g)


def test_c1_dns_dense_z_operators(monkeypatch):
    """The dense z contractions (the path of grids without mid-plane symmetry) in a
    full step; the default path on this grid is the parity-split one."""
    monkeypatch.setenv("LSB_NO_ZSPLIT", "1")
    case = L.cases.CASES["c1_dns"]
    st, g = run_both(case, case.nsteps)
    check_fields(st, g)


def test_c1_dns_two_kernel_helmholtz(monkeypatch):
    """The parity-split Helmholtz projections as two kernels through the mode-space
    buffer (the form for nz > 256); the default on this grid keeps the mode values
    of a block's pencils in LDS between the two projections."""
    monkeypatch.setenv("LSB_NO_ZHELM", "1")
    case = L.cases.CASES["c1_dns"]
    st, g = run_both(case, case.nsteps)
    check_fields(st, g)


def test_c1_dns_three_kernel_pressure_step(monkeypatch):
    """The pressure step as three kernels (pressure RHS, Poisson solve, dphi/dz with
    the projection) instead of one kernel with its intermediate fields in LDS."""
    monkeypatch.setenv("LSB_NO_ZPRESS", "1")
    case = L.cases.CASES["c1_dns"]
    st, g = run_both(case, case.nsteps)
    check_fields(st, g)


def test_matmul_derivative_path():
    """nz <= 31 uses matmul(d, .) instead of the DCT operator (derivatives.f90:122-129)."""
    case = L.cases.CASES["c1_dns"].with_(nx=16, ny=8, nz=24, pert_rms=0.02)
    st, g = run_both(case, 4)
    check_fields(st, g)


def test_spalart_bctype1():
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, bctype=1, samp=0.3)
    st, g = run_both(case, 4, x0=1.0, istartbound=0.0)
    check_fields(st, g)


def test_step_divergence_free():
    """After the projection the discrete divergence is at rounding level."""
    case = L.cases.CASES["c1_dns"]
    c = setup_case(case)
    u, v, w = L.cases.initial_condition(case)
    c.state_set(u, v, w)
    c.timestep(3)
    g = c.state_get()
    d = c.divcheck(g["uf"], g["vf"], g["wf"])
    scale = np.sqrt(np.mean(c.inv(g["uf"])[..., : c.nx] ** 2))
    c.close()
    assert d < 1e-9 * max(scale, 1e-3) * c.nx


@pytest.mark.parametrize("bctype", [3, 0])
def test_c1_dns_sponge_forcing(bctype):
    """forcing = 1 (navierstokes.f90:92-211, 750-756, 1389): lambda(x) (ubar - u, -v, -w) in the
    nonlinear term and velocity solves that take their boundary values; fringe fsize = 0.3 so
    that it covers a fifth of the 64 x-points.  bctype 0 isolates the switch of the solves."""
    case = L.cases.CASES["c1_dns"]
    kw = dict(forcing=1, bctype=bctype)
    o = oracle_for(case, fsize=0.3, lfsize=0.25, sfsize=0.70, **kw)
    u, v, w = L.cases.initial_condition(case)
    st = o.new_state(u, v, w)
    for _ in range(case.nsteps):
        o.timestep(st)
    st2 = noisy_oracle_run(oracle_for(case, fsize=0.3, lfsize=0.25, sfsize=0.70, **kw), (u, v, w), case.nsteps)
    c = setup_case(case, **kw)
    try:
        c.set_sponge(0.3, 0.25, 0.70)
        c.state_set(u, v, w)
        c.timestep(case.nsteps)
        g = c.state_get()
    finally:
        c.close()
    g["_sens"] = {k: rel_l2(st2[k], st[k]) for k in FIELDS}
    check_fields(st, g)
    # the sponge changes the solution well beyond the bar
    o0 = oracle_for(case, bctype=bctype)
    st0 = o0.new_state(u, v, w)
    for _ in range(case.nsteps):
        o0.timestep(st0)
    assert rel_l2(st0["uf"], st["uf"]) > 1e-6


@pytest.mark.parametrize("dealias,nsteps", [(2, 10), (0, 20)])
def test_c1_dns_temperature_buoyancy(dealias, nsteps):
    """Temperature / buoyancy path (navierstokes.f90:800-845, 1251-1303): base profile tbar =
    1 + 0.1 tanh(z/0.2) with its derivatives, temperature perturbation 0.5 u' of the initial
    condition, buoyancy coex (0, 0, -1), the temperature state included.  Without the
    spectral filter (dealiasflag 0) the per-step norm of the temperature (its kx = 0 column
    made Hermitian, its ky-Nyquist row zeroed, horfft.f90:153-224) shows in the fields."""
    case = L.cases.CASES["c1_dns"].with_(dealiasflag=dealias, nsteps=nsteps)
    kw = dict(coex=1.0, gz=-1.0)
    o = oracle_for(case, **kw)
    z = o.zpts
    th = np.tanh(z / 0.2)
    tb, tbp, tbpp = 1.0 + 0.1 * th, 0.5 * (1 - th ** 2), -2 * 0.5 / 0.2 * th * (1 - th ** 2)
    o.tbar, o.tbarp, o.tbarpp = tb, tbp, tbpp
    u, v, w = L.cases.initial_condition(case)
    T0 = 0.5 * u
    st = o.new_state(u, v, w, T0)
    for _ in range(case.nsteps):
        o.timestep(st)
    o2 = oracle_for(case, **kw)
    o2.tbar, o2.tbarp, o2.tbarpp = tb, tbp, tbpp
    st2 = noisy_oracle_run(o2, (u, v, w, T0), case.nsteps)
    c = setup_case(case)
    try:
        c.set_temperature(tb, tbp, tbpp, 1.0, 0.0, 0.0, -1.0)
        c.state_set(u, v, w)
        c.state_set_temperature(T0)
        c.timestep(case.nsteps)
        g = c.state_get()
        gt = c.state_get_temperature()
    finally:
        c.close()
    g["_sens"] = {k: rel_l2(st2[k], st[k]) for k in FIELDS}
    check_fields(st, g)
    tk = ("tempf", "tnf", "nltnf")
    gt["_sens"] = {k: rel_l2(st2[k], st[k]) for k in tk}
    check_fields(st, gt, fields=tk)


def test_profile_selected_classes():
    """lsb_profile_begin_classes: only the selected class's launch groups carry events;
    its measured groups and algorithmic work equal those of the all-class profile."""
    case = L.cases.CASES["c1_dns"]
    c = setup_case(case)
    try:
        c.state_set(*L.cases.initial_condition(case))
        c.timestep(3)
        c.profile_begin()
        c.timestep(2)
        full = c.profile_end()
        c.profile_begin_classes(["helmholtz"])
        c.timestep(2)
        sel = c.profile_end()
    finally:
        c.close()
    assert sel["helmholtz"]["n"] == full["helmholtz"]["n"] > 0
    assert sel["helmholtz"]["flops"] == full["helmholtz"]["flops"]
    assert sel["helmholtz"]["ms"] > 0
    for k, v in sel.items():
        if k != "helmholtz":
            assert v["n"] == 0 and v["ms"] == 0.0, (k, v)
