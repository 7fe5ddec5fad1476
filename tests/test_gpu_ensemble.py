"""Ensembles of contexts stepped from captured HIP graphs (lsb_ensemble_*).

A member stepped inside an ensemble runs the same kernels in the same order on its own
stream as when it is stepped alone, so after any number of ensemble steps its state is
bit-identical to the same member stepped eagerly.  Checked over more steps than the 60 of
the history-rotation period (every phase graph launched at least once, some twice), with
the ensemble's launches interleaved with eager downloads of the members' state; the
smallest member of the uneven ensemble is also held to the oracle at the step bar.
"""

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd import _lib
from lsbflow_amd.core import Ensemble, setup_case
from oracle import Oracle, OracleConfig
from tests.conftest import noisy_oracle_run, rel_l2
from tests.test_gpu_step import check_fields

pytestmark = pytest.mark.gpu

FIELDS = ("uf", "vf", "wf", "pf", "pnf", "unif", "unmif", "nlnif", "nlnmif")


def members_small():
    base = L.cases.CASES["c2_dyn_les"]
    return [base.with_(name="e0", nx=32, ny=16, nz=32, sgsmodel=0, xnu=2e-5, dt=2.5e-4),
            base.with_(name="e1", nx=48, ny=16, nz=40, sgsmodel=2, xnu=2e-5, dt=2.5e-4),
            base.with_(name="e2", nx=64, ny=32, nz=24, sgsmodel=1, xnu=2e-5, dt=2.5e-4)]


def eager(case, precision, nsteps):
    c = setup_case(case, precision=precision)
    try:
        c.state_set(*L.cases.initial_condition(case))
        c.timestep(nsteps)
        return c.state_get()
    finally:
        c.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_bit_identical_to_eager(precision):
    cases = members_small()  # e1's nx/2 = 24 runs the generic-length transforms
    ref = [eager(cs, precision, 3 + 70) for cs in cases]
    ctxs = [setup_case(cs, precision=precision) for cs in cases]
    try:
        for c, cs in zip(ctxs, cases):
            c.state_set(*L.cases.initial_condition(cs))
            c.timestep(3)
        ens = Ensemble(ctxs)
        try:
            info = ens.info()
            assert info["members"] == 3 and info["graphs"] == 60
            # a member inside an ensemble refuses eager steps
            with pytest.raises(_lib.LsbError, match="ensemble"):
                ctxs[0].timestep(1)
            ens.timestep(30)
            ens.synchronize()
            mid = [c.state_get() for c in ctxs]   # eager work between the launches
            assert all(np.all(np.isfinite(m["uf"])) for m in mid)
            ens.timestep(40)
            ens.synchronize()
            assert ens.info()["phase"] == 70 % 60
            for c, r, cs in zip(ctxs, ref, cases):
                g = c.state_get()
                for k in FIELDS:
                    assert np.array_equal(g[k], r[k]), (cs.name, k)
                assert g["istep"] == r["istep"] == 73 and g["t"] == r["t"]
        finally:
            ens.close()
    finally:
        for c in ctxs:
            c.close()
    print(f"precision {precision}: 3 members x 70 graph steps bit-identical to eager", flush=True)


def test_ensemble_refuses_startup_steps():
    cs = members_small()[0]
    c = setup_case(cs)
    try:
        c.state_set(*L.cases.initial_condition(cs))
        c.timestep(1)
        with pytest.raises(_lib.LsbError, match="start-up"):
            Ensemble([c])
    finally:
        c.close()


def test_smallest_uneven_member_against_oracle():
    """The smallest member of the uneven ensemble (FP64) through the graph ensemble beside
    a larger one, 3 start-up + 7 ensemble steps, against the oracle."""
    mem = L.cases.ensemble_members()
    small = min(mem, key=lambda m: m.points)
    other = mem[1]
    keys = OracleConfig.__dataclass_fields__
    o = Oracle(OracleConfig(**{k: v for k, v in small.as_dict().items() if k in keys}))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(small)
    u, v, w = L.cases.initial_condition(small)

    def run(fields):
        s = o.new_state(*fields)
        for _ in range(10):
            o.timestep(s)
        return s

    st = run((u, v, w))
    o2 = Oracle(OracleConfig(**{k: v for k, v in small.as_dict().items() if k in keys}))
    o2.ubar, o2.ubarp, o2.ubarpp = L.cases.base_flow(small)
    st2 = noisy_oracle_run(o2, (u, v, w), 10)
    ctxs = [setup_case(small), setup_case(other)]
    try:
        ctxs[0].state_set(u, v, w)
        ctxs[1].state_set(*L.cases.initial_condition(other))
        for c in ctxs:
            c.timestep(3)
        with Ensemble(ctxs) as ens:
            ens.timestep(7)
            ens.synchronize()
            g = ctxs[0].state_get()
    finally:
        for c in ctxs:
            c.close()
    g["_sens"] = {k: rel_l2(st2[k], st[k]) for k in ("uf", "vf", "wf", "pf")}
    print(f"{small.name} {small.nx}x{small.nzp}x{small.ny}:", flush=True)
    check_fields(st, g)
