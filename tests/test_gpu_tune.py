"""The launch-shape autotuner (lsb_autotune) and the launch variants it chooses between.

Every variant of a tunable family evaluates the same expressions in the same order, so a
step with any of them is bit-identical to a step with any other; the tuner only changes
the time.  Checked with the 3-point filter inside the dynamic model's step (FP64 form
and mixed form) on a grid whose nx spans several LDS tiles and a grid narrower than one,
and with the kx-column tiles of the y passes (8, 16, 32 columns) in both forms.
"""

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd.core import setup_case

pytestmark = pytest.mark.gpu

N_VARIANTS = 24


def run(case, precision, variant, nsteps=2):
    c = setup_case(case, precision=precision)
    try:
        if variant is not None:
            c.set_tune_table({"filter3_f64": {"variant": variant}, "filter3_f32": {"variant": variant}})
        c.state_set(*L.cases.initial_condition(case))
        c.timestep(nsteps)
        return c.state_get()
    finally:
        c.close()


@pytest.mark.parametrize("nx,ny,precision", [(128, 64, 0), (128, 64, 1), (32, 16, 0)])
def test_filter_variants_bit_identical(nx, ny, precision):
    case = L.cases.CASES["c2_dyn_les"].with_(nx=nx, ny=ny, nz=48, pert_rms=1e-2)
    ref = run(case, precision, None)
    for v in range(N_VARIANTS):
        g = run(case, precision, v)
        for k in ("uf", "vf", "wf", "pf"):
            assert np.array_equal(g[k], ref[k]), (v, k)
    print(f"{nx}x{ny} precision {precision}: {N_VARIANTS} filter variants bit-identical", flush=True)


Y_FAMILIES = ("yfwd_f64", "yfwd_f32", "yinv3_f64", "yinv3_f32", "yrhs_f32")


@pytest.mark.parametrize("precision", [0, 1])
def test_y_pass_tiles_bit_identical(precision):
    """Sigma LES on 128 x 49 x 64 (lh = 64: more than one tile of every width), 3 steps."""
    case = L.cases.CASES["c1_dns"].with_(nx=128, ny=64, nz=48, sgsflag=1, sgsmodel=2, xnu=2e-5, dt=2e-4,
                                         pert_rms=1e-2)
    ic = L.cases.initial_condition(case)

    def steps(cols):
        c = setup_case(case, precision=precision)
        try:
            if cols is not None:
                c.set_tune_table({f: {"variant": cols} for f in Y_FAMILIES})
            c.state_set(*ic)
            c.timestep(3)
            return c.state_get()
        finally:
            c.close()

    ref = steps(None)
    for cols in (8, 16, 32):
        g = steps(cols)
        for k in ("uf", "vf", "wf", "pf", "nlnif"):
            assert np.array_equal(g[k], ref[k]), (cols, k)
    print(f"precision {precision}: y-pass tiles of 8, 16, 32 columns bit-identical", flush=True)


def test_autotune_table():
    case = L.cases.CASES["c2_dyn_les"].with_(nx=128, ny=64, nz=64)
    c = setup_case(case)
    try:
        t = c.autotune()
        print("autotune:", t, flush=True)
        for fam in ("filter3_f64", "filter3_f32"):
            assert 0 <= t[fam]["variant"] < N_VARIANTS and t[fam]["ms"] > 0.0, t
        for fam in Y_FAMILIES:
            allowed = (8, 16) if fam.endswith("f64") else (8, 16, 32)
            assert t[fam]["variant"] in allowed and t[fam]["ms"] > 0.0, t
    finally:
        c.close()
