"""The LES workloads at their full sizes on the device, with the configured ceiling.

512 x 193 x 256 sigma LES and 256 x 129 x 128 dynamic LES (nx x nzp x ny), the
seeded tanh-plus-perturbation start and the ceiling suction/blowing profile of
``lsbflow_amd.cases`` (amplitude 0.3), exact Chebyshev operator and the
inflow-weighted ceiling (``dct_exact = 1``, ``top_bc = 2``).  The oracle cannot integrate these
sizes in a test's time, so the full-size runs are held to size-independent
properties (finite, bounded, divergence-free to rounding after the projection);
the same switches are held to the oracle at the step bar on reduced grids with
the full wall-normal resolution.
"""

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd.core import setup_case
from oracle import Oracle, OracleConfig
from tests.conftest import rel_l2
from tests.test_gpu_step import check_fields, run_both

pytestmark = pytest.mark.gpu

FIELDS = ("uf", "vf", "wf", "pf")


def _oracle(case):
    keys = OracleConfig.__dataclass_fields__
    o = Oracle(OracleConfig(**{k: v for k, v in case.as_dict().items() if k in keys}))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(case)
    return o


def _divergence(ctx, g):
    """max |ikx u + iky v + dw/dz| over the field, relative to max |w|."""
    d = ctx.divcheck(g["uf"], g["vf"], g["wf"])
    return d / max(float(np.abs(g["wf"]).max()), 1e-300)


@pytest.mark.parametrize("name,nsteps", [("c4_sigma_les", 100), ("c2_dyn_les", 100)])
def test_full_size_configured_ceiling_stays_bounded(name, nsteps):
    case = L.cases.CASES[name]
    u, v, w = L.cases.initial_condition(case)
    ctx = setup_case(case)
    ctx.state_set(u, v, w)
    peaks = []
    for chunk in range(nsteps // 25):
        ctx.timestep(25)
        g = ctx.state_get()
        m = {k: float(np.abs(g[k]).max()) for k in FIELDS}
        peaks.append(m)
        print(f"{name} step {25 * (chunk + 1)}: " + " ".join(f"max|{k}|={v:.3e}" for k, v in m.items()), flush=True)
        assert all(np.isfinite(x) for x in m.values()), m
    rel_div = _divergence(ctx, g)
    ctx.close()
    print(f"{name}: divcheck / max|w| = {rel_div:.3e}", flush=True)
    # the perturbation and the ceiling flow stay O(samp): no growth past the first chunk
    for m in peaks:
        assert m["uf"] < 1.0 and m["wf"] < 1.0, peaks
    assert rel_div < 1e-8


@pytest.mark.parametrize("sgsmodel", [1, 2])
def test_reduced_les_parity_exact_operator_blended_ceiling(sgsmodel):
    """C2/C4 switches against the oracle, full wall-normal resolution of C2 (nzp 129),
    reduced horizontal grid, 10 steps; the step bar with the oracle's own sensitivity
    (the clipped dynamic coefficient and the sigma model's acos amplify rounding)."""
    case = L.cases.CASES["c2_dyn_les"].with_(nx=32, ny=16, sgsmodel=sgsmodel)
    st, g = run_both(case, 10)
    check_fields(st, g)


@pytest.mark.parametrize("dct_exact,top_bc", [(1, 0), (0, 1), (1, 1), (1, 2), (0, 2)])
def test_switches_each_against_oracle(dct_exact, top_bc):
    """Each switch alone and combined, DNS start with the ceiling profile on, 8 steps."""
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40, samp=0.3, xnu=2e-4, dt=5e-4, pert_rms=0.05,
                                        dct_exact=dct_exact, top_bc=top_bc)
    st, g = run_both(case, 8)
    check_fields(st, g)
