"""GPU parity of the snapshot / restart records and the per-step diagnostics.

Bars: the record layouts (integer/byte work) are bit-exact against the oracle's
packing of the same FP64 values; values that pass through the 2-D transform are
within one float32 spacing (snapshot) or 1e-13 relative (maxcalc, wall shear) of
the oracle, whose transforms round differently in the last FP64 bit.  A restart
read back into a fresh context and stepped on matches the oracle stepped from
the same record to the step tests' 1e-10.
"""

import math
import os

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd._lib import LsbError
from lsbflow_amd.core import setup_case
from tests.conftest import rel_l2, rounding_noise, wrap_noisy_ops

pytestmark = pytest.mark.gpu

CASE = "c1_dns"


def oracle_for(case):
    keys = OracleConfig.__dataclass_fields__
    d = {k: v for k, v in case.as_dict().items() if k in keys}
    o = Oracle(OracleConfig(**d))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(case)
    return o


def random_spec(o, rng, amp=1.0):
    ph = np.zeros((o.nzp, o.ny, o.nx + 2))
    ph[:, :, : o.nx] = amp * rng.standard_normal((o.nzp, o.ny, o.nx))
    return o.horfft_fwd(ph)


def within_f32_spacing(a, b, k=1):
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(np.float32)).astype(np.float64)
    return bool(np.all(np.abs(a - b) <= k * sp))


@pytest.fixture(scope="module")
def case():
    return L.cases.CASES[CASE]


# ------------------------------------------------------------------ snapshots
@pytest.mark.parametrize("name", [0, 2])
def test_snapshot_record_matches_oracle(case, rng, name):
    o = oracle_for(case)
    o.tbar = 300.0 + rng.standard_normal(o.nzp)
    f = [random_spec(o, rng) for _ in range(5)]
    c = setup_case(case)
    try:
        rec = c.snapshot_record(*f, name=name, isum=1234, t=2.5, dt=1e-3, tbar=o.tbar)
        want = o.snapshot_record(*f, name=name, isum=1234, t=2.5, dt=1e-3)
        assert np.array_equal(rec[:7], want[:7])
        assert within_f32_spacing(rec[7:], want[7:])
        # the packing itself, bit for bit, on the library's own inverse transforms
        phys = []
        for fl in f:
            ph = np.zeros((o.nzp, o.ny, o.nx + 2))
            c.horfft(ph, np.ascontiguousarray(fl), 1)
            phys.append(ph)
        assert np.array_equal(rec[7:], o.snapshot_pack(phys, name))
        # tempf = NULL: temp column = tbar (name 0) or 0 (name 2)
        rec0 = c.snapshot_record(f[0], f[1], f[2], None, f[4], name=name, isum=1, t=0.0, dt=0.0, tbar=o.tbar)
        phys[3] = None
        assert np.array_equal(rec0[7:], o.snapshot_pack(phys, name))
    finally:
        c.close()


def test_state_snapshot_file_and_reader(case, tmp_path):
    o = oracle_for(case)
    c = setup_case(case)
    try:
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        c.timestep(2)
        g = c.state_get()
        path = str(tmp_path / "q00002.dat")
        rec = c.state_snapshot(0, path)
        assert np.array_equal(np.fromfile(path, dtype="<f4"), rec)
        want = c.snapshot_record(g["uf"], g["vf"], g["wf"], None, g["pf"], name=0, isum=g["istep"], t=g["t"],
                                 dt=case.dt)
        assert np.array_equal(rec, want)
        # an existing file is kept as prev_<name> (mpi_io.f90:400-404)
        c.timestep(1)
        rec3 = c.state_snapshot(0, path)
        assert np.array_equal(np.fromfile(str(tmp_path / "prev_q00002.dat"), dtype="<f4"), rec)
        assert np.array_equal(np.fromfile(path, dtype="<f4"), rec3)
        q, isum, t, dt = c.read_snapshot(path)
        h, qo = Oracle.read_snapshot(rec3, o.nx, o.ny, o.nzp)
        assert isum == 3 and t == float(np.float32(g["t"] + case.dt)) and dt == float(np.float32(case.dt))
        assert np.array_equal(q, qo)
    finally:
        c.close()


# -------------------------------------------------------------------- restart
@pytest.mark.parametrize("precision", [0, 1])
def test_restart_write_read_continue(case, tmp_path, precision):
    o = oracle_for(case)
    c = setup_case(case, precision=precision)
    c2 = setup_case(case, precision=precision)
    try:
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        c.timestep(3)
        g = c.state_get()
        path = str(tmp_path / "start.dat")
        rec = c.write_restart(path)
        assert np.array_equal(np.fromfile(path, dtype="<f4"), rec)
        st = o.new_state(g["uf"], g["vf"], g["wf"])
        for k in ("unif", "unmif", "nlnif", "nlnmif"):
            st[k] = g[k]
        st.update(istep=g["istep"], t=g["t"], dt=case.dt, dt2=case.dt, dt3=case.dt)
        assert np.array_equal(rec, o.restart_record(st))
        # read into a fresh context: the float32 values, pf = pnf = 0
        c2.read_restart(path)
        g2 = c2.state_get()
        back = o.read_restart(rec)
        for k in ("uf", "vf", "wf", "unif", "unmif", "nlnif", "nlnmif"):
            assert np.array_equal(g2[k], back[k]), k
        assert not np.any(g2["pf"]) and not np.any(g2["pnf"])
        assert g2["istep"] == 3 and g2["t"] == back["t"]
        if precision == 1:
            return  # the mixed form's continuation is covered by the mixed-form tests
        # continue both from the record: the restarted run starts again with the nstart
        # (here 0: Euler, then BDF2) start-up steps, as the reference's main loop does
        o2 = oracle_for(case)
        st2 = o2.new_state(back["uf"], back["vf"], back["wf"])
        st2.update(back)
        for _ in range(2):
            o2.timestep(st2)
        c2.timestep(2)
        g3 = c2.state_get()
        for k in ("uf", "vf", "wf"):
            assert rel_l2(g3[k], st2[k]) < 1e-10, k
        assert g3["istep"] == 5
        # state_set with a step count on a fresh context recomputes the BDF3/EX3 coefficients
        c4 = setup_case(case)
        try:
            c4.state_set(back["uf"], back["vf"], back["wf"], unif=back["unif"], unmif=back["unmif"],
                         nlnif=back["nlnif"], nlnmif=back["nlnmif"], istep=7, t=back["t"])
            c4.timestep(1)
            g4 = c4.state_get()
            o4 = oracle_for(case)
            st4 = o4.new_state(back["uf"], back["vf"], back["wf"])
            for k in ("unif", "unmif", "nlnif", "nlnmif"):
                st4[k] = back[k]
            st4.update(istep=7, t=back["t"])
            st4["abfac"] = o4.abcoef(o4.cfg.nstart, 8)
            st4["bdf"], st4["bdfv"] = o4.bdfcoef(o4.cfg.nstart, 8, case.dt, case.dt, case.dt)
            o4.timestep(st4)
            for k in ("uf", "vf", "wf"):
                assert rel_l2(g4[k], st4[k]) < 1e-10, k
            assert g4["istep"] == 8
        finally:
            c4.close()
        # a record of other dimensions is refused
        c3 = L.core.Context(L.core.Config.from_case(case, nz=32))
        try:
            with pytest.raises(Exception):
                c3.read_restart(path)
        finally:
            c3.close()
    finally:
        c.close()
        c2.close()


# ---------------------------------------------------------------- diagnostics
@pytest.mark.parametrize("t,fsize", [(0.0, 0.0), (5.0, 0.5)])
def test_maxcalc_matches_oracle(case, rng, t, fsize):
    o = oracle_for(case)
    o.cfg.istartbound = 1.0
    f = [random_spec(o, rng, 0.3) for _ in range(4)]
    tb = 300.0 + rng.standard_normal(o.nzp)
    o.tbar = tb
    c = setup_case(case, istartbound=1.0)
    try:
        for dt in (1e-3, 0.02):
            s, d1, d2, d3, ch = c.maxcalc(*f, dtmax=1.0, dt=dt, dt2=dt, dt3=0.5, t=t, fsize=fsize, tbar=tb)
            so, e1, e2, e3, cho = o.maxcalc(*f, 1.0, dt, dt, 0.5, t, fsize)
            assert np.max(np.abs(s - so) / np.maximum(np.abs(so), 1e-300)) < 1e-13, (s, so)
            assert ch == cho and d1 == e1 and d2 == e2 and d3 == e3 == 0.5
        # no temperature field and no tbar: tmax = tmin = 0
        s3, *_ = c.maxcalc(f[0], f[1], f[2], dtmax=1.0, dt=1e-3, dt2=1e-3, dt3=1e-3, t=t, fsize=fsize)
        o.tbar = np.zeros(o.nzp)
        so3, *_ = o.maxcalc(f[0], f[1], f[2], None, 1.0, 1e-3, 1e-3, 1e-3, t, fsize)
        assert s3[6] == 0.0 and s3[7] == 0.0
        assert np.max(np.abs(s3[:6] - so3[:6]) / np.abs(so3[:6])) < 1e-13
    finally:
        c.close()


def test_state_maxcalc_and_nancheck(case, rng):
    o = oracle_for(case)
    c = setup_case(case)
    try:
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        c.timestep(1)
        g = c.state_get()
        s, ch = c.state_maxcalc(dtmax=1.0)
        so, e1, e2, e3, cho = o.maxcalc(g["uf"], g["vf"], g["wf"], None, 1.0, case.dt, case.dt, case.dt, g["t"])
        assert np.max(np.abs(s[:6] - so[:6]) / np.abs(so[:6])) < 1e-13
        assert np.max(np.abs(s[8:] - so[8:]) / np.abs(so[8:])) < 1e-13
        assert ch == cho
        rec = c.write_restart(None)
        assert rec[2] == np.float32(e1) and rec[3] == np.float32(e2)
        assert not c.state_nancheck()
        assert not c.nancheck(g["uf"], g["vf"], g["wf"])
        bad = g["vf"].copy()
        bad[3, 4, 5] = np.inf
        assert not c.nancheck(g["uf"], bad, g["wf"])
        bad[7, 1, 2] = np.nan
        assert c.nancheck(g["uf"], bad, g["wf"])
        assert o.nancheck(g["uf"], bad, g["wf"])
    finally:
        c.close()


def test_wall_shear_matches_oracle(case, rng):
    o = oracle_for(case)
    c = setup_case(case)
    try:
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        c.timestep(2)
        g = c.state_get()
        cf, xs, xr = c.wall_shear()
        cfo, xso, xro = o.wall_shear(g["uf"])
        assert np.max(np.abs(cf - cfo)) < 1e-12 * np.max(np.abs(cfo))
        for a, b in ((xs, xso), (xr, xro)):
            assert (math.isnan(a) and math.isnan(b)) or abs(a - b) < 1e-12 * max(1.0, abs(b))
        # a field with a separated region: one streamwise mode on row nz
        f = np.zeros((o.nzp, o.ny, o.nx))
        f[o.nz - 1, 0, 0] = 0.3 - o.ubar[o.nz - 1] + o.ubar[o.nz]
        f[o.nz - 1, 0, 2] = 0.5
        cf1, xs1, xr1 = c.wall_shear(f)
        cfo1, xso1, xro1 = o.wall_shear(f)
        assert np.max(np.abs(cf1 - cfo1)) < 1e-12 * np.max(np.abs(cfo1))
        assert not math.isnan(xs1) and not math.isnan(xr1)
        assert abs(xs1 - xso1) < 1e-12 and abs(xr1 - xro1) < 1e-12
    finally:
        c.close()


def test_adaptive_dt_in_timestep(case):
    """dtadapt = 1, iomod = 5: after step 5 maxcalc finds every CFL below the lower limits
    (dt = 1e-3 on this grid) and grows dt by 1.25; the next steps run with the
    variable-step BDF/EX coefficients, against the oracle given the same dt sequence."""
    o = oracle_for(case)
    c = setup_case(case)
    try:
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        c.set_adaptive_dt(1, 5, 0.0)
        c.timestep(5)
        rec = c.write_restart(None)
        st = o.new_state(u, v, w)
        for _ in range(5):
            o.timestep(st)
        _, e1, e2, e3, cho = o.maxcalc(st["uf"], st["vf"], st["wf"], None, 1.0, st["dt"], st["dt2"], st["dt3"],
                                       st["t"])
        assert cho and abs(e1 - case.dt * 1.25) < 1e-18
        assert rec[2] == np.float32(e1) and rec[3] == np.float32(e2) and rec[4] == np.float32(e3)
        st["dt"], st["dt2"], st["dt3"] = e1, e2, e3
        for _ in range(3):
            o.timestep(st)
        c.timestep(3)
        g = c.state_get()
        for k in ("uf", "vf", "wf"):
            assert rel_l2(g[k], st[k]) < 1e-10, k
        assert abs(g["t"] - st["t"]) < 1e-15
    finally:
        c.close()


# ------------------------------------------------------- linear-stability energy
def test_ener_and_lintest_match_oracle(case, rng):
    o = oracle_for(case)
    c = setup_case(case)
    try:
        f = [random_spec(o, rng, 0.1) for _ in range(3)]
        for itot in (0, 1):
            e = c.ener(*f, itot=itot)
            eo = o.ener(*f, itot)
            assert abs(e - eo) < 1e-13 * abs(eo)
        # three steps of the configuration with lintest after each (main.f90:63-65)
        u, v, w = L.cases.initial_condition(case)
        c.state_set(u, v, w)
        n = 3
        arrs = [np.zeros(n) for _ in range(4)]
        tier_a = [np.zeros(n) for _ in range(4)]
        oa = [np.zeros(n) for _ in range(4)]
        for istep in range(1, n + 1):
            c.timestep(1)
            c.state_lintest(*arrs)
            g = c.state_get()
            c.lintest(g["uf"], g["vf"], g["wf"], istep, case.dt, *tier_a)
            o.lintest(g["uf"], g["vf"], g["wf"], istep, case.dt, *oa)
        for a, b in zip(arrs, tier_a):
            assert np.array_equal(a, b)
        for k in (0, 1):  # energies
            assert np.max(np.abs(arrs[k] - oa[k]) / np.abs(oa[k])) < 1e-13
        # growth rates: differences of energies that change by ~1e-6 per step
        for k in (2, 3):
            assert arrs[k][0] == 0.0
            assert np.max(np.abs(arrs[k][1:] - oa[k][1:])) < 1e-6 * np.max(np.abs(oa[k][1:])) + 1e-12
    finally:
        c.close()


# --------------------------------------------------------- on-the-fly averaging
@pytest.mark.parametrize("temperature", [False, True])
def test_time_average_and_qavg_snapshot(case, temperature):
    """otfavg = 1, istartavg = 3, iomod = 4 (main.f90:69-76, 98-106): the device averages
    after 9 steps against the oracle's, the qavg record of lsb_state_snapshot_average
    equal bit for bit to the record of the downloaded averages (name 1: u + ubar,
    temp + tbar), the averages zero after it, 4 more steps against the oracle again,
    and zero after a restart read.  With the temperature path on (FP64, buoyancy), the
    temperature average too.  Bar: the step tests' 1e-10, or 10 x the oracle's own drift
    under rounding noise where that is larger (1e-7 for the pressure)."""
    def make_oracle():
        o = oracle_for(case)
        if temperature:
            o.cfg.coex, o.cfg.gz = 1.0, -1.0
            th = np.tanh(o.zpts / 0.2)
            o.tbar, o.tbarp, o.tbarpp = 1.0 + 0.1 * th, 0.5 * (1 - th ** 2), -2 * 0.5 / 0.2 * th * (1 - th ** 2)
        return o

    o = make_oracle()
    rng_noise = np.random.default_rng(99)
    o_noisy = wrap_noisy_ops(make_oracle(), rng_noise)
    u, v, w = L.cases.initial_condition(case)
    T0 = 0.5 * u if temperature else None

    def oracle_run(fields, nsteps, avg, st=None, noisy=False):
        oo = o_noisy if noisy else o
        st = oo.new_state(*fields) if st is None else st
        for _ in range(nsteps):
            oo.timestep(st)
            if noisy:
                rounding_noise(st, rng_noise)
            oo.time_average(st, avg, 3, 4)
        return st

    def bound_for(k, avg, avg2):
        return max(1e-10, 10.0 * rel_l2(avg2[k], avg[k]), 1e-7 if k == "pf" else 0.0)

    zeros = lambda: {k: np.zeros((o.nzp, o.ny, o.nx)) for k in Oracle.AVG_FIELDS}  # noqa: E731
    ic = [u, v, w] + ([T0] if temperature else [])
    avg = zeros()
    st = oracle_run(ic, 9, avg)
    avg2 = zeros()
    st2 = oracle_run(ic, 9, avg2, noisy=True)
    keys = Oracle.AVG_FIELDS if temperature else ("uf", "vf", "wf", "pf")
    c = setup_case(case)
    try:
        if temperature:
            c.set_temperature(o.tbar, o.tbarp, o.tbarpp, 1.0, 0.0, 0.0, -1.0)
        c.set_time_average(1, 3, 4)
        c.state_set(u, v, w)
        if temperature:
            c.state_set_temperature(T0)
        c.timestep(9)
        ga = dict(zip(Oracle.AVG_FIELDS, c.state_get_time_average()))
        for k in keys:
            bound = bound_for(k, avg, avg2)
            assert rel_l2(ga[k], avg[k]) < bound, (k, rel_l2(ga[k], avg[k]), bound)
        if not temperature:
            assert not np.any(ga["tempf"])
        rec = c.state_snapshot_average()
        want = c.snapshot_record(*[ga[k] for k in Oracle.AVG_FIELDS], name=1, isum=9, t=st["t"], dt=case.dt,
                                 tbar=o.tbar if temperature else None)
        assert np.array_equal(rec, want)
        assert all(not np.any(x) for x in c.state_get_time_average())
        avg, avg2 = zeros(), zeros()
        oracle_run(None, 4, avg, st)
        oracle_run(None, 4, avg2, st2, noisy=True)
        c.timestep(4)
        ga = dict(zip(Oracle.AVG_FIELDS, c.state_get_time_average()))
        for k in keys:
            bound = bound_for(k, avg, avg2)
            assert rel_l2(ga[k], avg[k]) < bound, (k, rel_l2(ga[k], avg[k]), bound)
        c.read_restart(None, c.write_restart(None))
        assert all(not np.any(x) for x in c.state_get_time_average())
        c.set_time_average(0, 0, 1)
        with pytest.raises(LsbError):
            c.state_snapshot_average()
    finally:
        c.close()
