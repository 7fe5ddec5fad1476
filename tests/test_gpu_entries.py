"""Exported entry points that the step does not exercise, each against the oracle through
the C-ABI: the P = 1 data flips (horfft.f90:261-454, the stride reorder between x-rows
of nxpp and spectral rows of nx) and the device-pointer forms of horfft, dfdz and
solve_helmholtz (horfft.f90:61-150, derivatives.f90:96-156, helmholtz.f90:130-208) on
hipMalloc'd fields, several fields per call."""

import ctypes as C

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd import _lib
from lsbflow_amd.core import setup_case
from tests.conftest import rel_l2
from tests.test_gpu_helmholtz_c3 import Hip
from tests.test_gpu_step import oracle_for

pytestmark = pytest.mark.gpu

NF = 3


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module")
def setup():
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40)
    ctx = setup_case(case)
    o = oracle_for(case)
    yield case, ctx, o, Hip()
    ctx.close()


def test_dataflip_round_trip(setup):
    case, ctx, o, _ = setup
    lib = _lib.lib()
    nzp, ny, nx, nxpp = ctx.nzp, ctx.ny, ctx.nx, o.nxpp
    rng = np.random.default_rng(3)
    x = rng.standard_normal((nzp, ny, nxpp))
    xf = np.full((nzp, ny, nx), np.nan)
    assert lib.lsb_dataflip_xtoy(ctx.h, dp(x), dp(xf)) == 0
    assert np.array_equal(xf, x[..., :nx])
    back = np.full((nzp, ny, nxpp), 7.0)
    assert lib.lsb_dataflip_ytox(ctx.h, dp(back), dp(xf)) == 0
    assert np.array_equal(back[..., :nx], x[..., :nx])
    assert np.all(back[..., nx:] == 7.0)          # the x-row pads are not written


def test_horfft_dev_both_directions(setup):
    case, ctx, o, hip = setup
    lib = _lib.lib()
    nzp, ny, nx, nxpp = ctx.nzp, ctx.ny, ctx.nx, o.nxpp
    rng = np.random.default_rng(4)
    phys = np.zeros((NF, nzp, ny, nxpp))
    phys[..., :nx] = rng.standard_normal((NF, nzp, ny, nx))
    ref = np.stack([o.horfft_fwd(phys[f]) for f in range(NF)])
    dphys, dspec = hip.alloc(phys.nbytes), hip.alloc(ref.nbytes)
    try:
        hip.put(dphys, phys)
        assert lib.lsb_horfft_dev(ctx.h, dphys, dspec, -1, NF, None) == 0, lib.lsb_last_error()
        assert lib.lsb_synchronize(ctx.h) == 0
        spec = hip.get(dspec, ref.shape)
        assert rel_l2(spec, ref) < 1e-13
        back_ref = np.stack([o.horfft_inv(ref[f]) for f in range(NF)])
        assert lib.lsb_horfft_dev(ctx.h, dphys, dspec, 1, NF, None) == 0, lib.lsb_last_error()
        assert lib.lsb_synchronize(ctx.h) == 0
        back = hip.get(dphys, phys.shape)
        assert rel_l2(back[..., :nx], back_ref[..., :nx]) < 1e-13
    finally:
        hip.lib.hipFree(dphys)
        hip.lib.hipFree(dspec)


def test_dfdz_and_helmholtz_dev(setup):
    case, ctx, o, hip = setup
    lib = _lib.lib()
    u, v, w = L.cases.initial_condition(case.with_(pert_rms=0.05))
    f = np.stack([u, v, w])
    dref = np.stack([o.dfdz(f[i]) for i in range(NF)])
    sigma = 1.0 / (case.xnu * case.dt)
    href = np.stack([o.solve_helmholtz(f[i], sigma) for i in range(NF)])
    din, dout = hip.alloc(f.nbytes), hip.alloc(f.nbytes)
    try:
        hip.put(din, f)
        assert lib.lsb_dfdz_dev(ctx.h, din, dout, NF, None) == 0, lib.lsb_last_error()
        assert lib.lsb_synchronize(ctx.h) == 0
        assert rel_l2(hip.get(dout, f.shape), dref) < 1e-12
        assert lib.lsb_solve_helmholtz_dev(ctx.h, din, sigma, NF, None) == 0, lib.lsb_last_error()
        assert lib.lsb_synchronize(ctx.h) == 0
        got = hip.get(din, f.shape)
        assert rel_l2(got, href) < 1e-12
        assert np.all(got[:, 0] == 0.0) and np.all(got[:, -1] == 0.0)
    finally:
        hip.lib.hipFree(din)
        hip.lib.hipFree(dout)
