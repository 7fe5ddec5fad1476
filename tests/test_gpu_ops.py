"""GPU parity of every tier-A entry point against the CPU oracle (FP64).

Each test builds the same configuration on both sides, feeds seeded inputs
through the C-ABI and compares with the oracle's restatement of the reference
routine.  FP64 tolerance: relative L2 <= 1e-12 for single operators (rounding
only; the reference's own FFTW/LAPACK differ from the oracle's by the same
order).
"""

import math

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd.core import Config, Context
from tests.conftest import rel_l2, wrap_noisy_ops

pytestmark = pytest.mark.gpu

TOL = 1e-12


def make(nx=32, ny=16, nz=40, **kw):
    kw.setdefault("mapping", 3)
    ocfg = OracleConfig(nx=nx, ny=ny, nz=nz, **kw)
    o = Oracle(ocfg)
    ccfg = Config(nx=nx, ny=ny, nz=nz, **{k: v for k, v in kw.items() if k in Config.__dataclass_fields__})
    c = Context(ccfg)
    return o, c


@pytest.fixture(params=["split", "split2", "dense"])
def zpath(request, monkeypatch):
    """On a grid symmetric about the mid-plane the z operators run parity-split (two
    half-size products) and the two Helmholtz projections run as one kernel with the
    mode values in LDS; LSB_NO_ZHELM=1 keeps them as two kernels through HBM (the
    form for nz > 256), LSB_NO_ZSPLIT=1 the dense contraction that any other grid
    uses.  All must match the oracle."""
    monkeypatch.delenv("LSB_NO_ZSPLIT", raising=False)
    monkeypatch.delenv("LSB_NO_ZHELM", raising=False)
    if request.param == "dense":
        monkeypatch.setenv("LSB_NO_ZSPLIT", "1")
    elif request.param == "split2":
        monkeypatch.setenv("LSB_NO_ZHELM", "1")
    return request.param


def spec_field(o, rng, n=None):
    sh = (o.nzp, o.ny, o.nx) if n is None else (n, o.nzp, o.ny, o.nx)
    return rng.standard_normal(sh)


def phys_field(o, rng, n=None):
    sh = (o.nzp, o.ny, o.nxpp) if n is None else (n, o.nzp, o.ny, o.nxpp)
    f = rng.standard_normal(sh)
    f[..., o.nx:] = 0.0
    return f


@pytest.mark.parametrize("nx,ny,nz", [(32, 16, 40), (24, 20, 16), (40, 12, 33), (64, 32, 64)])
def test_horfft_forward_inverse(nx, ny, nz, rng):
    o, c = make(nx, ny, nz)
    a = phys_field(o, rng, 3)
    f = c.fwd(a)
    assert rel_l2(f, o.horfft_fwd(a)) < TOL
    s = spec_field(o, rng, 2)
    p = c.inv(s)
    ref = o.horfft_inv(s)
    assert rel_l2(p[..., :nx], ref[..., :nx]) < TOL
    assert np.all(p[..., nx:] == 0.0)


def test_horfft_roundtrip_and_modes(rng):
    o, c = make(32, 16, 40)
    # single Fourier mode cos(kx x + ky y): coefficient 1/2 at (m, j) and 0 elsewhere
    x = np.arange(o.nx) * 2 * math.pi / o.nx
    y = np.arange(o.ny) * 2 * math.pi / o.ny
    a = np.zeros((o.nzp, o.ny, o.nxpp))
    a[:, :, : o.nx] = np.cos(3 * x[None, :] + 2 * y[:, None])[None]
    f = c.fwd(a).view(np.complex128)
    assert abs(f[0, 2, 3] - 0.5) < 1e-14
    f[:, 2, 3] = 0
    assert np.max(np.abs(f)) < 1e-14
    # round trip of a band-limited field
    a = phys_field(o, rng)
    s = c.fwd(a)
    o.norm(s, s.copy(), s.copy(), s.copy())
    b = c.inv(s)
    assert rel_l2(c.fwd(b), s) < TOL


def test_norm(rng):
    o, c = make(32, 16, 40)
    fs = [spec_field(o, rng) for _ in range(4)]
    gs = [f.copy() for f in fs]
    c.norm(*fs)
    o.norm(*gs)
    for a, b in zip(fs, gs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nz", [40, 16])
def test_spectral_derivatives(nz, rng, zpath):
    o, c = make(32, 16, nz)
    u, v, w = (spec_field(o, rng) for _ in range(3))
    assert np.array_equal(c.dfdx(u), o.dfdx(u))
    assert np.array_equal(c.dfdy(u), o.dfdy(u))
    assert rel_l2(c.divuv(u, v), o.divuv(u, v)) < 1e-15
    assert rel_l2(c.dfdz(u), o.dfdz(u)) < TOL
    assert rel_l2(c.d2fdz2(u), o.d2fdz2(u)) < TOL
    div_ref = o.dfdz(w) + o.divuv(u, v)
    assert rel_l2(c.div(u, v, w), div_ref) < TOL
    lap_ref = o.d2fdz2(u) - (o.wavxx[None, None, :] + o.wavyy[None, :, None]) * u
    assert rel_l2(c.lapl(u), lap_ref) < TOL
    assert abs(c.divcheck(u, v, w) - o.divcheck(u, v, w)) < 1e-10 * o.divcheck(u, v, w)


def test_dpdz(rng, zpath):
    o, c = make(32, 16, 40)
    u = rng.standard_normal((o.nzp, 3, 5))
    for p in (1, 2):
        assert rel_l2(c.dpdz(u, p), o.dpdz(u, p)) < TOL


@pytest.mark.parametrize("sigma", [0.0, 37.5, 1.0e4])
def test_solve_helmholtz(sigma, rng, zpath):
    o, c = make(32, 16, 40)
    r = spec_field(o, rng)
    got = c.solve_helmholtz(r, sigma)
    ref = o.solve_helmholtz(r, sigma)
    assert rel_l2(got, ref) < 1e-11
    # residual of (D2 - k^2 - sigma) u = rhs on the interior rows (the oracle's d2)
    k2 = (o.wavxx[None, :] + o.wavyy[:, None])
    res = np.einsum("ij,jkl->ikl", o.d2, got) - (k2[None] + sigma) * got - r
    assert np.max(np.abs(res[1:-1])) < 1e-7 * max(1.0, np.max(np.abs(r)))


@pytest.mark.parametrize("sigma", [0.0, 250.0])
def test_solve_nhom_helmholtz(sigma, rng, zpath):
    o, c = make(32, 16, 40)
    r = spec_field(o, rng)
    bc = spec_field(o, rng)
    got = c.solve_nhom_helmholtz(r, bc, sigma)
    ref = o.solve_nhom_helmholtz(r, bc, sigma)
    assert rel_l2(got, ref) < 1e-11
    # Dirichlet data reproduced on the boundary planes
    assert np.max(np.abs(got[0] - bc[0])) < 1e-12 * np.max(np.abs(bc))
    assert np.max(np.abs(got[-1] - bc[-1])) < 1e-12 * np.max(np.abs(bc))


def test_helmholtz_matches_with_driver_matrices(rng, zpath):
    """lsb_set_mats with the oracle's (LAPACK) eigenbasis gives the same solution."""
    o, c = make(32, 16, 40)
    c.set_mats(diag=o.diag, smat=o.smat, simat=o.simat)
    r = spec_field(o, rng)
    assert rel_l2(c.solve_helmholtz(r, 12.0), o.solve_helmholtz(r, 12.0)) < 1e-13


def _grad_fields(o, rng, scale=1.0):
    g = rng.standard_normal((3, 3, o.nzp, o.ny, o.nxpp)) * scale
    g[..., o.nx:] = 0.0
    return g


def test_sigma_model(rng):
    o, c = make(32, 16, 40, sgsflag=1, sgsmodel=2)
    g = _grad_fields(o, rng)
    got = c.sigma_model(g)
    ref = o.sigma_model(g)
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < 1e-11


def test_static_smagorinsky(rng):
    o, c = make(32, 16, 40, sgsflag=1, sgsmodel=0, xnu=1e-3)
    g = _grad_fields(o, rng)
    got = c.static_smagorinsky(0.1, g)
    ref = o.static_smagorinsky(0.1, g)
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < 1e-12


def test_dynamic_smagorinsky(rng):
    o, c = make(32, 16, 40, sgsflag=1, sgsmodel=1, a=1.0, b=2.0)
    ui = phys_field(o, rng, 3)
    g = _grad_fields(o, rng)
    uu = np.einsum("l...,m...->lm...", ui, ui)
    got = c.dynamic_smagorinsky(ui, uu, g)
    ref = o.dynamic_smagorinsky(ui, uu, g)
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < 1e-11


def test_interscale_model(rng):
    o, c = make(32, 16, 40, sgsflag=1, sgsmodel=3, a=1.0, b=6.0)
    ui = phys_field(o, rng, 3)
    got = c.interscale_model(ui)
    ref = o.interscale_model(ui)
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < 1e-12


@pytest.mark.parametrize("ab", [(1.0, 2.0), (1.0, 1.0)])
def test_filter_q_3pt(ab, rng):
    o, c = make(32, 16, 40, a=1.0, b=2.0)
    q = phys_field(o, rng, 4)
    got = c.filter_q_3pt(q, ab[0], ab[1], 1)
    ref = o.filter_q_3pt(q, ab[0], ab[1], 1)
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < 1e-14


def test_spectral_filter(rng):
    o, c = make(32, 16, 40)
    u = spec_field(o, rng, 3)
    got = c.filterqspec(u)
    ref = o.filterqspec(u)
    assert rel_l2(got, ref) < 1e-14
    got1 = c.specfilter_exp_xy(u[0], 0.5, 2.0, 10.0, 1)
    ref1 = o.specfilter_exp_xy(u[0], 0.5, 2.0, 10.0, 1)
    assert rel_l2(got1, ref1) < 1e-14


@pytest.mark.parametrize("ref_compat", [1, 0])
def test_dealiasing(ref_compat, rng):
    o, c = make(32, 16, 40, dealiasflag=1, dealias_ref_compat=ref_compat)
    assert c.setup_dealiasing() == o.dealias_sizes()
    a = phys_field(o, rng)
    b = phys_field(o, rng)
    got = c.dealiasing_2d_fftw(a, b)
    ref = o.dealiasing_2d_fftw(a, b, bool(ref_compat))
    assert rel_l2(got[..., : o.nx], ref[..., : o.nx]) < TOL


@pytest.mark.parametrize("ref_compat", [1, 0])
def test_dealiasf(ref_compat, rng):
    """dealiasing.f90:449-618, spectral in and out."""
    o, c = make(32, 16, 40, dealiasflag=1, dealias_ref_compat=ref_compat)
    af, bf = spec_field(o, rng), spec_field(o, rng)
    got = c.dealiasf_2d_fftw(af, bf)
    ref = o.dealiasf_2d_fftw(af, bf, bool(ref_compat))
    assert rel_l2(got, ref) < TOL
    c.close()


def test_test_fft(rng):
    """horfft.f90:7-58: a kx-Nyquist field (dropped by the transform) gives its mean |u| = 1
    (to the transforms' rounding, 1e-12); a random resolved field comes back to rounding,
    as in the oracle."""
    o, c = make(32, 16, 40)
    alt = np.zeros((o.nzp, o.ny, o.nxpp))
    alt[..., : o.nx] = (-1.0) ** np.arange(o.nx)
    assert abs(c.test_fft(alt) - 1.0) < 1e-12
    f = o.horfft_inv(spec_field(o, rng) * (np.arange(o.ny)[None, :, None] != o.ny // 2))
    e, eo = c.test_fft(f), o.test_fft(f)
    assert e < 1e-15 * np.max(np.abs(f)) * 10 and eo < 1e-15 * np.max(np.abs(f)) * 10
    c.close()


@pytest.mark.parametrize("grid,mapping,per,rev", [(3, 3, 0, 0), (3, 3, 0, 1), (3, 3, 1, 0), (2, 2, 1, 0),
                                                 (3, 1, 0, 0), (2, 1, 1, 0)])
def test_zinteg(grid, mapping, per, rev, rng):
    """derivatives.f90:9-70 on (n1, n2) = (nxpp, ny) pencils; mapping is the routine's
    argument (1: plane 1 copies plane 2), the grid's own mapping 3 or 2."""
    o, c = make(32, 16, 40, mapping=grid, zlen=2.0 if grid == 3 else 20.0)
    tmp = phys_field(o, rng)
    got = c.zinteg(tmp, mapping, per, rev)
    ref = o.zinteg(tmp, mapping, per, rev)
    assert rel_l2(got, ref) < 1e-14
    assert np.array_equal(got == 0, ref == 0)
    c.close()


@pytest.mark.parametrize("sgs", [(0, 0), (1, 2), (1, 0), (1, 1), (1, 3)])
def test_nlterm_skewsym(sgs, rng):
    case = L.cases.CASES["c1_dns"].with_(nx=32, ny=16, nz=40)
    o, c = make(32, 16, 40, sgsflag=sgs[0], sgsmodel=sgs[1], a=1.0, b=2.0)
    ub, ubp, ubpp = L.cases.base_flow(case)
    o.ubar, o.ubarp, o.ubarpp = ub, ubp, ubpp
    c.set_paral(ub, ubp, ubpp)
    u, v, w = L.cases.initial_condition(case.with_(pert_rms=0.05))
    uif = np.stack([u, v, w])
    got = c.nlterm_skewsym(uif)
    ref, _ = o.nlterm_skewsym(uif, np.zeros_like(u), sgs[0])
    # the oracle's response to 1e-16 relative noise on its input and on the output of every
    # transform and derivative bounds the attainable agreement (the sigma model's
    # invariant/acos evaluation amplifies it on this nearly parallel flow)
    o2 = Oracle(OracleConfig(nx=32, ny=16, nz=40, mapping=3, sgsflag=sgs[0], sgsmodel=sgs[1], a=1.0, b=2.0))
    o2.ubar, o2.ubarp, o2.ubarpp = ub, ubp, ubpp
    wrap_noisy_ops(o2, rng)
    pert = uif * (1.0 + 1e-16 * rng.standard_normal(uif.shape))
    ref2, _ = o2.nlterm_skewsym(pert, np.zeros_like(u), sgs[0])
    sens = rel_l2(ref2, ref)
    err = rel_l2(got, ref)
    print(f"nlterm sgs {sgs}: rel L2 {err:.2e} (sens {sens:.1e})", flush=True)
    assert err < max(1e-11, 10.0 * sens), (err, sens)
