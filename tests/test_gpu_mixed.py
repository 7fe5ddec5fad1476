"""The mixed-precision form (lsb_config.precision = LSB_MIXED) against the FP64 form.

The mixed form keeps the state, the histories, the right-hand sides and the
Helmholtz/Poisson solves in FP64 and runs the nonlinear term's chain (inverse
transforms, products and SGS model, forward transforms, its two Chebyshev
contractions) in FP32; the dynamic model's work fields are FP32 as well, and the
interscale model's chain stays FP64 (its N history and solve correction FP32).
Acceptance: relative L2 <= 1e-5 per field (u, v, w, p and the wall skin friction)
against the FP64 form from the same initial condition after 100 steps of the
512 x 193 x 256 sigma and dynamic-Smagorinsky LES workloads with their configured
ceiling.  The small cases hold the velocities to 1e-5 and state the pressure with a
looser bound: it is recovered as a divergence divided by k^2 dt
(navierstokes.f90:1466-1476), which amplifies the FP32 rounding of the nonlinear
term on these coarse, strongly perturbed grids.
"""

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd.core import setup_case
from tests.conftest import rel_l2
from tests.test_gpu_step import oracle_for

pytestmark = pytest.mark.gpu

VEL = ("uf", "vf", "wf")
VEL_BAR = 1e-5
P_BAR = 1e-3


def run_pair(case, nsteps):
    u, v, w = L.cases.initial_condition(case)
    out, cf = {}, {}
    for prec in (0, 1):
        c = setup_case(case, precision=prec)
        c.state_set(u, v, w)
        c.timestep(nsteps)
        out[prec] = c.state_get()
        cf[prec] = c.wall_shear()[0]
        c.close()
    errs = {k: rel_l2(out[1][k], out[0][k]) for k in VEL + ("pf",)}
    errs["wall_shear"] = rel_l2(cf[1], cf[0])
    print(case.name, nsteps, "steps, mixed vs FP64:", {k: f"{e:.3e}" for k, e in errs.items()}, flush=True)
    for p in (0, 1):
        for k in VEL + ("pf",):
            assert np.all(np.isfinite(out[p][k])), (p, k)
    return errs


def check(errs):
    assert max(errs[k] for k in VEL) < VEL_BAR, errs
    assert errs["pf"] < P_BAR, errs


@pytest.mark.parametrize("zpath", ["split", "dense"])
def test_mixed_dns_small(zpath, monkeypatch):
    """Both z-contraction paths of the FP32 chain (parity-split on this symmetric
    grid; LSB_NO_ZSPLIT=1 the dense one)."""
    if zpath == "dense":
        monkeypatch.setenv("LSB_NO_ZSPLIT", "1")
    case = L.cases.CASES["c1_dns"]
    check(run_pair(case, case.nsteps))


@pytest.mark.parametrize("sgsmodel", [0, 2])
def test_mixed_les_models(sgsmodel):
    case = L.cases.CASES["c1_dns"].with_(name=f"les{sgsmodel}", nx=128, ny=64, nz=64, xnu=2e-5, bctype=0,
                                         sgsflag=1, sgsmodel=sgsmodel, dt=5e-4, pert_rms=1e-2)
    check(run_pair(case, 20))


@pytest.mark.parametrize("nx,ny,sgsmodel", [(96, 48, 0), (192, 32, 2), (64, 48, 2)])
def test_mixed_generic_lengths(nx, ny, sgsmodel):
    """Lengths outside the register-resident set (nx/2 = 48 or 96, ny = 48: factors of 3):
    the mixed form's transforms run as FP32 Stockham passes, d/dx and d/dy taken in FP64
    before the rounding, and the right-hand side as its own pass over the FP32 N pieces."""
    case = L.cases.CASES["c1_dns"].with_(name=f"gen{nx}x{ny}", nx=nx, ny=ny, nz=64, xnu=2e-5, bctype=0,
                                         sgsflag=1, sgsmodel=sgsmodel, dt=5e-4, pert_rms=1e-2)
    check(run_pair(case, 20))


def test_mixed_dynamic_generic_lengths():
    """The mixed dynamic model with its configured ceiling at nx = 96 (nx/2 = 48)."""
    case = L.cases.CASES["c2_dyn_les"].with_(name="dyn_gen", nx=96, ny=64, nz=64, pert_rms=1e-2)
    check(run_pair(case, 20))


@pytest.mark.parametrize("bctype", [3])
def test_mixed_nonhomogeneous_velocity_solves(bctype, monkeypatch):
    """The velocity solves with boundary data (the ceiling profile) in the
    FP32-correction form: sigma = 1/(nu dt bdfv) = 1.8e8 is over 16 x max|lambda|, so the
    mixed form's lifting terms go through the FP64 leading term and the FP32 correction.
    Six steps (the reference's non-homogeneous path grows at these viscosities, so the
    comparison stays short) against the FP64 form."""
    case = L.cases.CASES["c1_dns"].with_(name=f"nhom{bctype}", nx=128, ny=64, nz=64, xnu=2e-5, bctype=bctype,
                                         sgsflag=1, sgsmodel=0, dt=5e-4, pert_rms=1e-2, samp=0.3)
    errs = run_pair(case, 6)
    assert max(errs[k] for k in VEL) < 1e-6, errs
    # against the same mixed form with FP64 velocity solves: the FP32 correction adds
    # nothing visible to the FP32 nonlinear chain's rounding
    monkeypatch.setenv("LSB_NO_ZHELM32", "1")
    errs64 = run_pair(case, 6)
    for k in VEL:
        assert errs[k] < 1.5 * errs64[k] + 1e-9, (k, errs, errs64)


@pytest.mark.parametrize("name", ["c4_sigma_les", "c4_dyn_les"])
def test_mixed_c4_100_steps(name):
    """The acceptance runs: 512 x 193 x 256 LES (sigma, dynamic Smagorinsky) with the
    configured ceiling, 100 steps, mixed vs FP64, every field within 1e-5."""
    case = L.cases.CASES[name]
    errs = run_pair(case, case.nsteps)
    for k, e in errs.items():
        assert e < 1e-5, errs


@pytest.mark.parametrize("dyn_avg", [0, 1])
def test_mixed_dynamic_small(dyn_avg):
    """The mixed dynamic model (FP32 work fields, base flow split off) with the local and
    the plane-averaged coefficient, 128 x 65 x 64, configured ceiling, 20 steps."""
    case = L.cases.CASES["c2_dyn_les"].with_(name=f"dyn{dyn_avg}", nx=128, ny=64, nz=64, dyn_avg=dyn_avg,
                                             pert_rms=1e-2)
    check(run_pair(case, 20))


@pytest.mark.parametrize("nx,ny,nz", [(128, 64, 64), (96, 30, 64), (512, 32, 64), (512, 256, 128)])
def test_mixed_dynamic_fused_average_bit_identical(nx, ny, nz, monkeypatch):
    """The local (1, 1) average of M:L and M:M and stage 3 in one pass give the same bits as
    the averages through HBM followed by stage 3 (LSB_NO_DYNFUSE=1): a register-resident
    grid, a generic one whose tiles are partial in x and y (nx = 96, ny = 30), a grid of
    four x tiles (these three with 4 planes per block and 65 planes, so the last block
    holds one) and a 512 x 129 x 256 grid with 16 planes per block, as at C4."""
    case = L.cases.CASES["c2_dyn_les"].with_(name=f"dyna{nx}x{ny}", nx=nx, ny=ny, nz=nz, dyn_avg=0, pert_rms=1e-2)
    u, v, w = L.cases.initial_condition(case)
    out = {}
    for sep in ("0", "1"):
        monkeypatch.setenv("LSB_NO_DYNFUSE", sep)
        c = setup_case(case, precision=1)
        try:
            c.state_set(u, v, w)
            c.timestep(10)
            out[sep] = c.state_get()
        finally:
            c.close()
    for k in VEL + ("pf", "nlnif", "nlnmif"):
        assert np.all(np.isfinite(out["0"][k])), k
        assert np.array_equal(out["0"][k], out["1"][k]), k
    print("fused (1, 1) average + stage 3 bit-identical:", case.name, flush=True)


@pytest.mark.parametrize("nx,ny,forcing", [(128, 64, 0), (96, 30, 1)])
def test_mixed_dynamic_stage1_pairs_bit_identical(nx, ny, forcing, monkeypatch):
    """Stages 1 and 2 of the mixed dynamic model with two x points per thread (8-byte loads
    and stores) give the same bits as one point per thread (LSB_NO_DYNX2=1), with and
    without the sponge term."""
    case = L.cases.CASES["c2_dyn_les"].with_(name=f"dyn1x{nx}x{ny}", nx=nx, ny=ny, nz=64, dyn_avg=0, pert_rms=1e-2)
    u, v, w = L.cases.initial_condition(case)
    out = {}
    for single in ("0", "1"):
        monkeypatch.setenv("LSB_NO_DYNX2", single)
        c = setup_case(case, precision=1, forcing=forcing)
        try:
            if forcing:
                c.set_sponge(0.3, 0.25, 0.70)
            c.state_set(u, v, w)
            c.timestep(10)
            out[single] = c.state_get()
        finally:
            c.close()
    for k in VEL + ("pf", "nlnif", "nlnmif"):
        assert np.all(np.isfinite(out["0"][k])), k
        assert np.array_equal(out["0"][k], out["1"][k]), k
    print("stages 1 and 2, two points per thread, bit-identical:", case.name, "forcing", forcing, flush=True)


@pytest.mark.parametrize("dealias", [2, 0])
def test_mixed_temperature_buoyancy(dealias):
    """The temperature path in the mixed form: its chain (grad T, the products with the
    velocity, the temperature solve) runs in FP64 from the FP64 velocity spectra, the
    buoyancy enters the FP64 right-hand side.  C1 with a tanh temperature profile,
    perturbation 0.5 u' and buoyancy (0, 0, -1), 10 steps, mixed vs FP64."""
    case = L.cases.CASES["c1_dns"].with_(dealiasflag=dealias)
    u, v, w = L.cases.initial_condition(case)
    z = oracle_for(case).zpts
    th = np.tanh(z / 0.2)
    tb, tbp, tbpp = 1.0 + 0.1 * th, 0.5 * (1 - th ** 2), -2 * 0.5 / 0.2 * th * (1 - th ** 2)
    out = {}
    for prec in (0, 1):
        c = setup_case(case, precision=prec)
        try:
            c.set_temperature(tb, tbp, tbpp, 1.0, 0.0, 0.0, -1.0)
            c.state_set(u, v, w)
            c.state_set_temperature(0.5 * u)
            c.timestep(10)
            out[prec] = dict(c.state_get(), tempf=c.state_get_temperature()["tempf"])
        finally:
            c.close()
    errs = {k: rel_l2(out[1][k], out[0][k]) for k in VEL + ("tempf", "pf")}
    print("temperature, dealiasflag", dealias, "mixed vs FP64:", {k: f"{e:.3e}" for k, e in errs.items()},
          flush=True)
    for k in VEL + ("tempf",):
        assert errs[k] < VEL_BAR, errs
    assert errs["pf"] < P_BAR, errs


@pytest.mark.parametrize("nx,ny", [(128, 64), (96, 48)])
def test_mixed_interscale_model(nx, ny):
    """The interscale model in the mixed form: its nonlinear chain (products of the
    filtered velocities, the non-symmetric tau) runs in FP64, the N history is kept in
    FP32 and the velocity solves take the FP32 correction, on a register-resident grid
    and on a generic one (nx/2 = 48, ny = 48)."""
    case = L.cases.CASES["c1_dns"].with_(name=f"isc{nx}x{ny}", nx=nx, ny=ny, nz=64, xnu=2e-5, bctype=0,
                                         sgsflag=1, sgsmodel=3, dt=5e-4, pert_rms=1e-2)
    check(run_pair(case, 20))


def test_mixed_state_roundtrip_history():
    """The mixed form keeps the N history (nlnif, nlnmif) in FP32 on the device:
    state_set/state_get still exchange FP64 arrays, the history coming back rounded
    to FP32 and every other field unchanged."""
    case = L.cases.CASES["c1_dns"]
    rng = np.random.default_rng(7)
    u, v, w = L.cases.initial_condition(case)
    c = setup_case(case, precision=1)
    try:
        s = (3,) + u.shape
        hist = {k: rng.standard_normal(s) for k in ("unif", "unmif", "nlnif", "nlnmif")}
        p, pn = rng.standard_normal(u.shape), rng.standard_normal(u.shape)
        c.state_set(u, v, w, pf=p, pnf=pn, istep=5, t=0.25, **hist)
        g = c.state_get()
        for k in ("uf", "vf", "wf"):
            assert np.array_equal(g[k], {"uf": u, "vf": v, "wf": w}[k])
        assert np.array_equal(g["pf"], p) and np.array_equal(g["pnf"], pn)
        for k in ("unif", "unmif"):
            assert np.array_equal(g[k], hist[k])
        for k in ("nlnif", "nlnmif"):
            assert np.array_equal(g[k], hist[k].astype(np.float32).astype(np.float64))
    finally:
        c.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_inverse_x_pass_staged_store_bit_identical(precision, monkeypatch):
    """The inverse x pass assembling a block's rows in LDS and storing them as one aligned
    run gives the same bits as its direct row stores (LSB_NO_XSTAGE=1), in both forms, on
    a 512-point x grid with the sigma model (every physical field goes through it)."""
    case = L.cases.CASES["c1_dns"].with_(nx=512, ny=32, nz=32, sgsflag=1, sgsmodel=2, xnu=2e-5, dt=2e-4)
    out = {}
    for off in ("0", "1"):
        monkeypatch.setenv("LSB_NO_XSTAGE", off)
        c = setup_case(case, precision=precision)
        c.state_set(*L.cases.initial_condition(case))
        c.timestep(3)
        out[off] = c.state_get()
        c.close()
    for k in ("uf", "vf", "wf", "pf"):
        assert np.array_equal(out["0"][k], out["1"][k]), k
