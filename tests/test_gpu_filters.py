"""GPU parity of the TNS / mode-cut filtering (filters.f90:9-161, 617-654) and of the time
step with filtering on (navierstokes.f90:31-39) against the oracle.

Bars: the filters alone 1e-13 relative L2 per field (FP64 transforms and 3-pt passes that
round differently from numpy's), the energy integrals 1e-12 times their conditioning,
identical decisions; the filtered time step 1e-10 like the other step tests (the
pressure: 1000 x the oracle's own sensitivity when larger).
"""

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd.core import setup_case
from tests.conftest import noisy_oracle_run, rel_l2

pytestmark = pytest.mark.gpu

CASE = L.cases.CASES["c1_dns"]


def oracle_for(case, **kw):
    keys = OracleConfig.__dataclass_fields__
    d = {k: v for k, v in case.as_dict().items() if k in keys}
    d.update(kw)
    o = Oracle(OracleConfig(**d))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(case)
    return o


def noise(o, rng, amp):
    ph = np.zeros((o.nzp, o.ny, o.nx + 2))
    ph[:, :, : o.nx] = amp * rng.standard_normal((o.nzp, o.ny, o.nx))
    ph[0] = 0.0
    ph[-1] = 0.0
    return o.horfft_fwd(ph)


def mode(o, mx, my):
    x = 2 * np.pi * np.arange(o.nx) / o.nx
    y = 2 * np.pi * np.arange(o.ny) / o.ny
    ph = np.zeros((o.nzp, o.ny, o.nx + 2))
    ph[:, :, : o.nx] = ((0.3 + 0.7 * o.zpts)[:, None, None] * np.cos(my * y)[None, :, None]
                        * np.cos(mx * x)[None, None, :])
    return o.horfft_fwd(ph)


def test_filterall3pt_matches_oracle(rng):
    o = oracle_for(CASE)
    f = [noise(o, rng, 1.0) for _ in range(5)]
    c = setup_case(CASE)
    try:
        got = c.filterall3pt(*f)
    finally:
        c.close()
    want = o.filterall3pt(*f)
    for g, w in zip(got, want):
        assert rel_l2(g, w) < 1e-13


@pytest.mark.parametrize("kind", ["mode_low", "mode_high", "noise"])
def test_tns_matches_oracle(rng, kind):
    o = oracle_for(CASE)
    if kind == "noise":
        f = [noise(o, rng, 1.0) for _ in range(4)]
    else:
        mx, my = (3, 2) if kind == "mode_low" else (24, 12)
        z = np.zeros((o.nzp, o.ny, o.nx))
        f = [mode(o, mx, my), z, z, z]
    c = setup_case(CASE)
    try:
        got = c.tns_automatic_filtering(f[0], f[1], f[2], None, f[3])
    finally:
        c.close()
    want = o.tns_automatic_filtering(f[0], f[1], f[2], None, f[3])
    assert got[7] == want[7]
    # I = z-integral of the energy removed, (q - F q)^2: its rounding error relative to I is
    # about eps * sqrt(Iq / I), Iq the same integral of q^2 (a low mode loses ~1e-10 of itself)
    q = np.stack([o.horfft_inv(x)[:, :, : o.nx] for x in f[:3]])
    e_q = np.sum(q ** 2, axis=(0, 2, 3)) / (o.nx * o.ny)
    Iq = sum(0.5 * (e_q[k] + e_q[k + 1]) * (o.zpts[k] - o.zpts[k + 1]) for k in range(o.nz))
    for idx in (5, 6):
        tol = 1e-12 + 1e-14 * np.sqrt(Iq / want[idx])
        assert abs(got[idx] - want[idx]) < tol * want[idx], (idx, got[idx], want[idx], tol)
    for k in (0, 1, 2, 4):
        assert rel_l2(got[k], want[k]) < 1e-13 if np.any(want[k]) else not np.any(got[k])
    if kind == "mode_low":
        assert not got[7] and np.array_equal(got[0], f[0])
    if kind == "mode_high":
        assert got[7]


@pytest.mark.parametrize("variant", ["filterall3pt", "tns_applied", "tns_not_applied"])
def test_timestep_with_filtering(variant):
    """ifilt = 1, modfilt = 2: the filter runs before steps 2, 4, ... (navierstokes.f90:31-39).
    TNS (sgsflag 1, sgsmodel 4, no SGS stress): with small-scale noise in the initial
    condition the energy ratio (0.011-0.027) exceeds the limit and the filter applies;
    on the plain initial condition (5e-4) it does not."""
    kw = dict(ifilt=1, modfilt=2)
    ckw = {}
    if variant != "filterall3pt":
        kw.update(sgsflag=1, sgsmodel=4)
        ckw.update(sgsflag=1, sgsmodel=4)
    o = oracle_for(CASE, **kw)
    u, v, w = L.cases.initial_condition(CASE)
    if variant in ("filterall3pt", "tns_applied"):
        rng = np.random.default_rng(7)
        u, v, w = (f + noise(o, rng, 1e-4) for f in (u, v, w))
    nsteps = 6
    st = o.new_state(u, v, w)
    for _ in range(nsteps):
        o.timestep(st)
    c = setup_case(CASE, **ckw)
    try:
        c.set_filtering(1, 2)
        c.state_set(u, v, w)
        c.timestep(nsteps)
        g = c.state_get()
    finally:
        c.close()
    # the step tests' bar: 1e-10, or 10 x the oracle's own drift under rounding noise where
    # that is larger, and 1e-7 for the pressure (a divergence over k^2 dt)
    st2 = noisy_oracle_run(oracle_for(CASE, **kw), (u, v, w), nsteps)
    for k in ("uf", "vf", "wf", "pf"):
        e, s = rel_l2(g[k], st[k]), rel_l2(st2[k], st[k])
        bound = max(1e-10, 10.0 * s, 1e-7 if k == "pf" else 0.0)
        print(f"{variant} {k} {e:.2e} (sens {s:.1e})", flush=True)
        assert e < bound, (k, e, bound)
    # the filter changed the solution (or, for tns_not_applied, did not)
    o0 = oracle_for(CASE, **{k: v for k, v in kw.items() if k not in ("ifilt", "modfilt")})
    st0 = o0.new_state(u, v, w)
    for _ in range(nsteps):
        o0.timestep(st0)
    d = rel_l2(st0["uf"], st["uf"])
    assert (d == 0.0) if variant == "tns_not_applied" else (d > 1e-9)
