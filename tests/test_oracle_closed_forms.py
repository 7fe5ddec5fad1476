"""Pin the oracle's restated routines with closed forms, brute force and invariants.

Each check is chosen so that a plausible restatement mistake (wrong scale, sign,
index, transposed operand, dropped term) fails it.  CPU only.
"""

import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from oracle import Oracle, OracleConfig


def small(**kw):
    base = dict(nx=16, ny=8, nz=40, mapping=3, zlen=2.0, xlen=2 * math.pi, ylen=math.pi)
    base.update(kw)
    return Oracle(OracleConfig(**base))


def phys(o, f):
    x = o.xpts - o.xpts[0]
    x = np.arange(o.nx) * (2 * math.pi / o.nx) * (o.cfg.xlen / (2 * math.pi))
    y = np.arange(o.ny) * (o.cfg.ylen / o.ny)
    a = np.zeros((o.nzp, o.ny, o.nxpp))
    a[:, :, : o.nx] = f(x[None, None, :], y[None, :, None], o.zpts[:, None, None])
    return a


# ----------------------------------------------------------------- horfft
def test_horfft_single_mode_and_scale():
    o = small()
    a = phys(o, lambda x, y, z: np.cos(3 * x + 4 * y) + 0 * z)          # ky = 4*beta, beta = 2
    f = o.horfft_fwd(a).view(np.complex128)
    assert abs(f[0, 2, 3] - 0.5) < 1e-15                               # ky index 2 (beta = 2)
    f2 = f.copy()
    f2[:, 2, 3] = 0
    assert np.max(np.abs(f2)) < 1e-15
    c = o.horfft_fwd(phys(o, lambda x, y, z: 2.5 + 0 * x + 0 * y + 0 * z)).view(np.complex128)
    assert np.allclose(c[:, 0, 0], 2.5, atol=1e-15, rtol=0)


def test_horfft_nyquist_dropped_and_negative_ky():
    o = small()
    alt = phys(o, lambda x, y, z: np.cos(math.pi * np.round(x * o.nx / (2 * math.pi))) + 0 * y + 0 * z)
    assert np.max(np.abs(o.horfft_fwd(alt))) < 1e-15                   # kx-Nyquist is not stored
    s = phys(o, lambda x, y, z: np.sin(x - 2 * y) + 0 * z)             # ky = -1 (index ny-1)
    f = o.horfft_fwd(s).view(np.complex128)
    assert abs(f[0, o.ny - 1, 1] - (-0.5j)) < 1e-15


def test_horfft_roundtrip():
    o = small()
    rng = np.random.default_rng(1)
    a = np.zeros((o.nzp, o.ny, o.nxpp))
    a[..., : o.nx] = rng.standard_normal((o.nzp, o.ny, o.nx))
    f = o.horfft_fwd(a)
    o.norm(f, f.copy(), f.copy(), f.copy())                             # ky-Nyquist removed
    g = o.horfft_fwd(o.horfft_inv(f))
    assert np.max(np.abs(g - f)) < 1e-15


def test_dfdx_dfdy_analytic():
    o = small()
    a = phys(o, lambda x, y, z: np.sin(3 * x) * np.cos(4 * y) + 0 * z)
    dx = o.horfft_inv(o.dfdx(o.horfft_fwd(a)))[..., : o.nx]
    dy = o.horfft_inv(o.dfdy(o.horfft_fwd(a)))[..., : o.nx]
    ex = phys(o, lambda x, y, z: 3 * np.cos(3 * x) * np.cos(4 * y) + 0 * z)[..., : o.nx]
    ey = phys(o, lambda x, y, z: -4 * np.sin(3 * x) * np.sin(4 * y) + 0 * z)[..., : o.nx]
    assert np.max(np.abs(dx - ex)) < 1e-13
    assert np.max(np.abs(dy - ey)) < 1e-13


def test_norm_hermitian():
    o = small()
    rng = np.random.default_rng(2)
    f = rng.standard_normal((o.nzp, o.ny, o.nx))
    o.norm(f, f.copy(), f.copy(), f.copy())
    c = f.view(np.complex128)
    for j in range(1, o.ny // 2):
        assert np.array_equal(c[:, o.ny - j, 0], np.conj(c[:, j, 0]))
    assert np.all(c[:, 0, 0].imag == 0)
    assert np.all(c[:, o.ny // 2, :] == 0)
    g = f.copy()
    o.norm(g, g.copy(), g.copy(), g.copy())
    assert np.array_equal(f, g)                                        # idempotent


# ------------------------------------------------------------ Chebyshev
@pytest.mark.parametrize("mapping,zlen", [(3, 2.0), (0, 1.0), (3, 5.0)])
def test_cgl_matrix_exact_on_polynomials(mapping, zlen):
    """build_cheb_zder_matrix's d differentiates every polynomial of degree <= nz exactly."""
    o = small(mapping=mapping, zlen=zlen, nz=24)
    xl = o.xl
    xi = 2 * (o.zpts - (o.zpts[-1] if mapping == 3 else -xl / 2)) / xl - 1
    for n in (1, 5, 17, 24):
        c = np.zeros(n + 1)
        c[n] = 1
        p = npcheb.chebval(xi, c)
        dp = npcheb.chebval(xi, npcheb.chebder(c)) * (2 / xl)
        assert np.max(np.abs(o.d @ p - dp)) < 1e-9 * max(1, np.max(np.abs(dp)))


def test_dct_diff_exact_below_top_mode_and_q2():
    """dct_diff equals d on degree < N; on T_N it is exact only when xl = 1 (quirk Q2)."""
    for mapping, zlen, exact_top in ((3, 2.0, False), (0, 1.0, True)):
        o = small(mapping=mapping, zlen=zlen, nz=40)
        xl = o.xl
        xi = (2 * (o.zpts - (o.zpts[-1] if mapping == 3 else -xl / 2)) / xl - 1)
        for n in (0, 3, 20, 39):
            c = np.zeros(n + 1)
            c[n] = 1
            p = npcheb.chebval(xi, c)
            assert np.max(np.abs(o.dct_diff(p[:, None], 1)[:, 0] - o.d @ p)) < 1e-10 * max(1.0, n * n)
        c = np.zeros(41)
        c[40] = 1
        p = npcheb.chebval(xi, c)
        err = np.max(np.abs(o.dct_diff(p[:, None], 1)[:, 0] - o.d @ p)) / np.max(np.abs(o.d @ p))
        assert (err < 1e-10) == exact_top
    # p = 2 is the operator applied twice (chebyshev.f90:238-259)
    o = small(nz=40)
    y = np.random.default_rng(3).standard_normal((o.nzp, 4))
    assert np.allclose(o.dct_diff(y, 2), o.dct_diff(o.dct_diff(y, 1), 1), rtol=0, atol=1e-9 * np.max(np.abs(o.dct_diff(y, 2))))


def test_eigendecomposition_reconstructs():
    o = small(nz=40)
    D = o.d2[1:o.nz, 1:o.nz]
    assert np.max(np.abs(o.smat @ np.diag(o.diag) @ o.simat - D)) < 1e-12 * np.max(np.abs(D))
    assert np.all(o.diag < 0)


# -------------------------------------------------------------- helmholtz
@pytest.mark.parametrize("sigma", [0.0, 400.0])
def test_helmholtz_residual(sigma):
    o = small(nz=40)
    rng = np.random.default_rng(4)
    r = rng.standard_normal((o.nzp, o.ny, o.nx))
    u = o.solve_helmholtz(r, sigma)
    k2 = o.wavxx[None, :] + o.wavyy[:, None]
    res = np.einsum("ij,jkl->ikl", o.d2, u) - (k2 + sigma) * u - r
    assert np.max(np.abs(res[1:-1])) < 1e-8 * np.max(np.abs(r)) * 1e3
    assert np.all(u[0] == 0) and np.all(u[-1] == 0)


def test_nhom_helmholtz_residual_and_bc():
    o = small(nz=40)
    rng = np.random.default_rng(5)
    r = rng.standard_normal((o.nzp, o.ny, o.nx))
    bc = rng.standard_normal((o.nzp, o.ny, o.nx))
    s = 123.0
    u = o.solve_nhom_helmholtz(r, bc, s)
    assert np.allclose(u[0], bc[0], atol=1e-12) and np.allclose(u[-1], bc[-1], atol=1e-12)
    k2 = o.wavxx[None, :] + o.wavyy[:, None]
    # the lifting functions are exact solutions of their own operator only up to d2h vs D2 h
    hb, ht, d2hb, d2ht = o.lifting_functions()
    res = (np.einsum("ij,jkl->ikl", o.d2, u) - (k2 + s) * u - r)
    lift_err = (np.einsum("ij,j->i", o.d2, hb) - d2hb)[:, None, None] * bc[-1] + \
               (np.einsum("ij,j->i", o.d2, ht) - d2ht)[:, None, None] * bc[0]
    assert np.max(np.abs((res - lift_err)[1:-1])) < 1e-6 * np.max(np.abs(r)) * 1e3


# ---------------------------------------------------------------- filters
def brute_filter(o, q, a, b):
    """filters.f90:383-453 written as the reference's loops (one field, no pads)."""
    nx, ny, nzp = o.nx, o.ny, o.nzp
    s = 2.0 * a + b
    qh = np.zeros((nzp, ny, nx))
    for k in range(nzp):
        for j in range(ny):
            for i in range(nx):
                qh[k, j, i] = (a * q[k, j, (i - 1) % nx] + b * q[k, j, i] + a * q[k, j, (i + 1) % nx]) / s
    qf = np.zeros_like(qh)
    for k in range(nzp):
        for j in range(ny):
            qf[k, j] = (a * qh[k, (j - 1) % ny] + b * qh[k, j] + a * qh[k, (j + 1) % ny]) / s
    az, bz, cz = o.filter_z_weights(a, b, 1)
    out = qf.copy()
    for k in range(1, nzp - 1):
        out[k] = (az[k] * qf[k - 1] + bz[k] * qf[k] + cz[k] * qf[k + 1]) / s
    return out


def test_filter_q_3pt_brute_force():
    o = small(nx=6, ny=4, nz=8)
    rng = np.random.default_rng(6)
    q = np.zeros((o.nzp, o.ny, o.nxpp))
    q[..., : o.nx] = rng.standard_normal((o.nzp, o.ny, o.nx))
    for a, b in ((1.0, 2.0), (1.0, 1.0), (1.0, 6.0)):
        got = o.filter_q_3pt(q, a, b, 1)[..., : o.nx]
        assert np.max(np.abs(got - brute_filter(o, q, a, b))) < 1e-15
    # the interpolating z weights sum to 2a+b (filter preserves constants)
    az, bz, cz = o.filter_z_weights(1.0, 2.0, 1)
    assert np.allclose(az + bz + cz, 4.0, rtol=0, atol=1e-13)


def test_spectral_filter_closed_form():
    o = small(nx=32, ny=16)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((o.nzp, o.ny, o.nx))
    f = o.specfilter_exp_xy(u, 0.6, 1.0, 20.0, 1).view(np.complex128)
    c = u.view(np.complex128)
    akmax = math.hypot(o.nxh * o.alpha, o.nyh * o.beta)
    ak = np.sqrt(o.wavxx[0::2][None, :] + o.wavyy[:, None])
    th = (ak - 0.6 * akmax) / (akmax - 0.6 * akmax)
    fac = np.where(ak >= 0.6 * akmax, np.exp(-20.0 * th ** 2), 1.0)
    # the reference filters through magnitude and phase (acos/cos/sin), whose round trip is
    # ill-conditioned near phase 0: allow that rounding, nothing more
    assert np.max(np.abs(f - fac[None] * c)) < 1e-12 * np.max(np.abs(c))
    assert np.max(np.abs(f - fac[None] * c)) > 0.0


# ------------------------------------------------------------------- SGS
def test_sigma_model_singular_values():
    """D_sigma from the invariant formula equals s3(s1-s2)(s2-s3)/s1^2 from numpy's SVD."""
    o = small(nx=8, ny=4, nz=6)
    rng = np.random.default_rng(8)
    g = np.zeros((3, 3, o.nzp, o.ny, o.nxpp))
    g[..., : o.nx] = rng.standard_normal((3, 3, o.nzp, o.ny, o.nx))
    tau = o.sigma_model(g)
    z = o.zpts
    for k in range(o.nzp):
        dz = z[0] - z[1] if k == 0 else (z[o.nz - 1] - z[-1] if k == o.nzp - 1 else 0.5 * (z[k - 1] - z[k + 1]))
        delta = (o.dx * o.dy * dz) ** (1 / 3)
        for j in range(o.ny):
            for i in range(o.nx):
                gg = g[:, :, k, j, i]
                s = np.linalg.svd(gg, compute_uv=False)
                ds = max(s[2] * (s[0] - s[1]) * (s[1] - s[2]) / s[0] ** 2, 0.0)
                nu = -(1.35 * delta) ** 2 * ds
                ref = 2 * nu * 0.5 * (gg + gg.T)
                assert np.allclose(tau[:, :, k, j, i], ref, rtol=1e-7, atol=1e-12 * np.max(np.abs(ref)) + 1e-300)


def test_static_smagorinsky_brute_force():
    o = small(nx=8, ny=4, nz=6)
    rng = np.random.default_rng(9)
    g = np.zeros((3, 3, o.nzp, o.ny, o.nxpp))
    g[..., : o.nx] = rng.standard_normal((3, 3, o.nzp, o.ny, o.nx))
    tau = o.static_smagorinsky(0.1, g)
    utau = math.sqrt(abs(np.mean(o.cfg.xnu * g[0, 2, -1, :, : o.nx])))
    delta = o.sgs_delta_static()
    for k in range(o.nzp):
        zp = o.zpts[k] * utau / o.cfg.xnu
        for j in range(o.ny):
            for i in range(o.nx):
                S = 0.5 * (g[:, :, k, j, i] + g[:, :, k, j, i].T)
                nu = -(0.1 * delta[k] * (1 - math.exp(-zp / 25))) ** 2 * math.sqrt(2 * np.sum(S * S))
                assert np.allclose(tau[:, :, k, j, i], nu * S, rtol=1e-12, atol=1e-300)


def test_ceiling_row_without_eddy_viscosity():
    """nut_ceil0 = 1: the static and sigma stresses vanish on the ceiling row (row 0) and
    are unchanged on every other row."""
    for model in ("static", "sigma"):
        out = {}
        for flag in (0, 1):
            o = small(nx=8, ny=4, nz=6, nut_ceil0=flag)
            g = np.zeros((3, 3, o.nzp, o.ny, o.nxpp))
            g[..., : o.nx] = np.random.default_rng(13).standard_normal((3, 3, o.nzp, o.ny, o.nx))
            out[flag] = o.static_smagorinsky(0.1, g) if model == "static" else o.sigma_model(g)
        assert np.max(np.abs(out[0][:, :, 0])) > 0.0, model
        assert np.all(out[1][:, :, 0] == 0.0), model
        assert np.array_equal(out[1][:, :, 1:], out[0][:, :, 1:]), model


def test_dynamic_smagorinsky_pointwise_algebra():
    """Restate sgsterm.f90:171-267 with explicit per-point loops on top of the filter."""
    o = small(nx=6, ny=4, nz=6, a=1.0, b=2.0)
    rng = np.random.default_rng(10)
    sh = (o.nzp, o.ny, o.nxpp)
    ui = np.zeros((3,) + sh)
    ui[..., : o.nx] = rng.standard_normal((3, o.nzp, o.ny, o.nx))
    g = np.zeros((3, 3) + sh)
    g[..., : o.nx] = rng.standard_normal((3, 3, o.nzp, o.ny, o.nx))
    uu = np.einsum("l...,m...->lm...", ui, ui)
    tau = o.dynamic_smagorinsky(ui, uu, g)
    z, nx = o.zpts, o.nx
    dl = np.array([(o.dx * o.dy * (z[0] - z[1])) ** (1 / 3)] +
                  [(o.dx * o.dy * 0.5 * (z[k - 1] - z[k + 1])) ** (1 / 3) for k in range(1, o.nzp - 1)] +
                  [(o.dx * o.dy * (z[-2] - z[-1])) ** (1 / 3)])[:, None, None]
    S = 0.5 * (g + np.swapaxes(g, 0, 1))
    Sn = np.sqrt(2 * np.sum(S * S, axis=(0, 1)))
    F = lambda q: o.filter_q_3pt(q, 1.0, 2.0, 1)  # noqa: E731  (pinned above by brute force)
    Sh = F(S)
    Shn = np.sqrt(2 * np.sum(Sh * Sh, axis=(0, 1)))
    M = 2 * (F(dl ** 2 * Sn * S) - (2 * dl) ** 2 * Shn * Sh)
    uh = F(ui)
    Lij = F(uu) - np.einsum("l...,m...->lm...", uh, uh)
    F11 = lambda q: o.filter_q_3pt(q, 1.0, 1.0, 1)  # noqa: E731  (sgsterm.f90:243-244)
    num = F11(np.sum(M * Lij, axis=(0, 1)))
    den = F11(np.sum(M * M, axis=(0, 1)))
    with np.errstate(invalid="ignore", divide="ignore"):
        C = np.clip(np.nan_to_num(num / den, nan=0.0), 0.0, 0.04)
    ref = -2 * C * dl ** 2 * Sn * S
    assert np.allclose(tau[..., :nx], ref[..., :nx], rtol=1e-10, atol=1e-14)


def _zweights(z, a, b):
    """filters.f90:396-448, channel mapping, interpflag 1 (rows top first)."""
    n = len(z)
    az, bz, cz = np.full(n, a), np.full(n, b), np.full(n, a)
    mid = (n + 1) // 2
    for k in range(2, mid):
        m, p = z[k - 1] - z[k - 2], z[k] - z[k - 1]
        az[k - 1], bz[k - 1], cz[k - 1] = 2 * a * m / (p + m), 2 * a * (1 - m / p) + b, 2 * a * m * m / (p * (p + m))
    for k in range(mid + 1, n):
        m, p = z[k - 1] - z[k - 2], z[k] - z[k - 1]
        az[k - 1], bz[k - 1], cz[k - 1] = 2 * a * p * p / (m * (m + p)), 2 * a * (1 - p / m) + b, 2 * a * p / (m + p)
    az[mid - 1], bz[mid - 1], cz[mid - 1] = a, b, a
    az[[0, -1]], bz[[0, -1]], cz[[0, -1]] = 0.0, 2 * a + b, 0.0
    return az, bz, cz


def _zpass(v, w, s):
    az, bz, cz = w
    out = np.array(v, copy=True)
    out[1:-1] = (az[1:-1, None] * v[:-2] + bz[1:-1, None] * v[1:-1] + cz[1:-1, None] * v[2:]) / s
    return out


def _dynamic_x_only(o, u1, du1, dyn_avg, a=1.0, b=2.0):
    """sgsterm.f90:123-280 for u = u1(x), v = w = 0 (no y or z dependence), written
    separably: every x pass of the 3-point filter through its exact transfer function
    (b + 2a cos(2 pi m / nx)) / (2a + b) on the DFT, the z pass (filters.f90:396-448) acting
    only on the row widths Delta(z) (its weights sum to 2a + b, so the z-independent factors
    pass through).  Returns tau_11 (nzp, nx) and the coefficient c (nzp, nx)."""
    nx, nzp = o.nx, o.nzp
    m = np.fft.fftfreq(nx, 1.0 / nx)

    def fx(f, a_=a, b_=b):
        t = (b_ + 2 * a_ * np.cos(2 * math.pi * m / nx)) / (2 * a_ + b_)
        return np.real(np.fft.ifft(t * np.fft.fft(f, axis=-1), axis=-1))

    z = o.zpts
    dz = np.empty(nzp)
    dz[0], dz[-1], dz[1:-1] = z[0] - z[1], z[-2] - z[-1], 0.5 * (z[:-2] - z[2:])
    D2 = ((o.dx * o.dy * dz) ** (1 / 3)) ** 2
    S, Sh = du1, fx(du1)
    L = fx(u1 * u1) - fx(u1) ** 2                                  # Leonard term L_11
    P = fx(math.sqrt(2) * np.abs(S) * S)                           # filter of |S| S_11 (x part)
    Q = math.sqrt(2) * np.abs(Sh) * Sh                             # |S^| S^_11
    ZD2 = _zpass(D2[:, None], _zweights(z, a, b), 2 * a + b)[:, 0]
    M = 2 * (ZD2[:, None] * P[None, :] - 4 * D2[:, None] * Q[None, :])
    N, D = M * L[None, :], M * M
    if dyn_avg:
        cn, cd = N.mean(axis=1, keepdims=True), D.mean(axis=1, keepdims=True)
    else:
        w1 = _zweights(z, 1.0, 1.0)
        cn, cd = _zpass(fx(N, 1.0, 1.0), w1, 3.0), _zpass(fx(D, 1.0, 1.0), w1, 3.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = cn / cd
    c = np.minimum(np.where((c < 0) | np.isnan(c), 0.0, c), 0.2 ** 2) * np.ones((1, nx))
    return -2 * c * D2[:, None] * math.sqrt(2) * np.abs(S)[None, :] * S[None, :], c


def _xonly_fields(o, k1, a1, k2, a2):
    nx = o.nx
    x = np.arange(nx) * (o.cfg.xlen / nx)
    u1 = a1 * np.sin(k1 * x) + a2 * np.cos(k2 * x)
    du1 = a1 * k1 * np.cos(k1 * x) - a2 * k2 * np.sin(k2 * x)
    sh = (o.nzp, o.ny, o.nxpp)
    ui = np.zeros((3,) + sh)
    ui[0, ..., :nx] = u1
    g = np.zeros((3, 3) + sh)
    g[0, 0, ..., :nx] = du1
    return ui, g, u1, du1


@pytest.mark.parametrize("dyn_avg,modes", [(1, (2, 1.0, 1, 0.7)), (1, (6, 1.0, 3, 1.0)), (0, (1, 1.0, 2, 0.5))])
def test_dynamic_smagorinsky_x_only_field(dyn_avg, modes):
    """The oracle's dynamic model on u = a1 sin(k1 x) + a2 cos(k2 x) against the separable
    spectral restatement (the Leonard term of a single mode is also checked in closed form:
    L_11 = A^2/2 (1 - T(2k) cos 2kx) - T(k)^2 A^2 sin^2 kx).  The mode pairs give
    coefficients inside (0, 0.04) on a good share of the points, so the test is not
    settled by the clipping alone."""
    o = small(nx=16, ny=8, nz=12, a=1.0, b=2.0, dyn_avg=dyn_avg)
    ui, g, u1, du1 = _xonly_fields(o, *modes)
    uu = np.einsum("l...,m...->lm...", ui, ui)
    tau = o.dynamic_smagorinsky(ui, uu, g)
    want, c = _dynamic_x_only(o, u1, du1, dyn_avg)
    assert np.mean((c > 1e-6) & (c < 0.04)) > 0.02, c
    assert np.max(np.abs(tau[0, 0, :, :, : o.nx] - want[:, None, :])) < 1e-13 * np.max(np.abs(want))
    others = np.delete(tau.reshape(9, *tau.shape[2:]), 0, axis=0)
    assert np.max(np.abs(others)) == 0.0
    # single-mode Leonard term in closed form
    A, k = 0.8, 3
    nx = o.nx
    x = np.arange(nx) * (o.cfg.xlen / nx)
    T = lambda kk: (2.0 + 2 * 1.0 * math.cos(kk * o.cfg.xlen / nx)) / 4.0  # noqa: E731
    u = A * np.sin(k * x)
    uu1 = np.zeros((o.nzp, o.ny, o.nxpp))
    uu1[..., :nx] = u * u
    u1f = np.zeros((o.nzp, o.ny, o.nxpp))
    u1f[..., :nx] = u
    Lo = (o.filter_q_3pt(uu1, 1.0, 2.0, 1) - o.filter_q_3pt(u1f, 1.0, 2.0, 1) ** 2)[..., :nx]
    Lc = A * A / 2 * (1 - T(2 * k) * np.cos(2 * k * x)) - T(k) ** 2 * A * A * np.sin(k * x) ** 2
    assert np.max(np.abs(Lo - Lc[None, None, :])) < 1e-14


def test_dynamic_smagorinsky_invariances():
    """Galilean invariance (a uniform velocity added to u_i changes neither L_ij nor M_ij),
    degree-2 homogeneity (c is scale free, tau ~ |u|^2), and zero stress when the test
    filter is the identity (a = 0: L_ij = 0)."""
    o = small(nx=8, ny=4, nz=10, a=1.0, b=2.0)
    rng = np.random.default_rng(21)
    sh = (o.nzp, o.ny, o.nxpp)
    ui = np.zeros((3,) + sh)
    ui[..., : o.nx] = rng.standard_normal((3, o.nzp, o.ny, o.nx))
    g = np.zeros((3, 3) + sh)
    g[..., : o.nx] = rng.standard_normal((3, 3, o.nzp, o.ny, o.nx))
    uu = lambda q: np.einsum("l...,m...->lm...", q, q)  # noqa: E731
    tau = o.dynamic_smagorinsky(ui, uu(ui), g)
    assert np.max(np.abs(tau)) > 0.0
    shifted = ui.copy()
    shifted[:, ..., : o.nx] += np.array([0.7, -1.3, 0.4])[:, None, None, None]
    tau_g = o.dynamic_smagorinsky(shifted, uu(shifted), g)
    assert np.max(np.abs(tau_g - tau)) < 1e-12 * np.max(np.abs(tau))
    tau_2 = o.dynamic_smagorinsky(2.0 * ui, uu(2.0 * ui), 2.0 * g)
    assert np.max(np.abs(tau_2 - 4.0 * tau)) < 1e-12 * np.max(np.abs(tau))
    o0 = small(nx=8, ny=4, nz=10, a=0.0, b=2.0)
    assert np.max(np.abs(o0.dynamic_smagorinsky(ui, uu(ui), g))) == 0.0


# ---------------------------------------------------------------- dealias
def test_dealiasing_correct_mode_exact_product():
    o = small(nx=16, ny=8, dealiasflag=1)
    a = phys(o, lambda x, y, z: np.cos(2 * x) * np.cos(2 * y) + 0 * z)
    b = phys(o, lambda x, y, z: np.cos(x) * np.sin(4 * y) + 0 * z)
    got = o.dealiasing_2d_fftw(a, b, ref_compat=False)[..., : o.nx]
    assert np.max(np.abs(got - (a * b)[..., : o.nx])) < 1e-14


def brute_dealias(o, a, b, ref_compat):
    """dealiasing.f90:322-441 with explicit DFT sums (FFTW definitions; c2r reads only the
    real part of the kx = 0 input, the kx-Nyquist input is zero)."""
    nx, ny = o.nx, o.ny
    nxl, nyl = o.dealias_sizes()

    def dft_fwd(f, Nx, Ny, keep_x):            # (Ny, Nx) real -> (Ny, keep_x), scaled 1/(Nx Ny)
        ex = np.exp(-2j * np.pi * np.outer(np.arange(Nx), np.arange(keep_x)) / Nx)
        ey = np.exp(-2j * np.pi * np.outer(np.arange(Ny), np.arange(Ny)) / Ny)
        return ey @ (f @ ex) / (Nx * Ny)

    def dft_inv(F, Nx, Ny):                     # (Ny, m < Nx/2) complex -> (Ny, Nx) real
        ey = np.exp(2j * np.pi * np.outer(np.arange(Ny), np.arange(Ny)) / Ny)
        Y = ey @ F
        out = np.tile(Y[:, :1].real, (1, Nx))
        xs = np.arange(Nx)
        for m in range(1, F.shape[1]):
            out += 2 * np.real(Y[:, m:m + 1] * np.exp(2j * np.pi * m * xs / Nx)[None, :])
        return out

    def place(j):
        return j if (ref_compat or j < ny // 2) else j - ny + nyl

    res = np.zeros((o.nzp, ny, nx))
    for k in range(o.nzp):
        A = dft_fwd(a[k, :, :nx], nx, ny, nx // 2)
        B = dft_fwd(b[k, :, :nx], nx, ny, nx // 2)
        Al = np.zeros((nyl, nx // 2), complex)
        Bl = np.zeros((nyl, nx // 2), complex)
        for j in range(ny):
            Al[place(j)] = A[j]
            Bl[place(j)] = B[j]
        prod = dft_inv(Al, nxl, nyl) * dft_inv(Bl, nxl, nyl)
        P = dft_fwd(prod, nxl, nyl, nx // 2)
        out = np.array([P[place(j)] for j in range(ny)])
        res[k] = dft_inv(out, nx, ny)
    return res


@pytest.mark.parametrize("ref_compat", [True, False])
def test_dealiasing_index_flow_brute_force(ref_compat):
    o = small(nx=8, ny=4, nz=4, dealiasflag=1)
    rng = np.random.default_rng(11)
    a = np.zeros((o.nzp, o.ny, o.nxpp))
    b = np.zeros_like(a)
    a[..., : o.nx] = rng.standard_normal((o.nzp, o.ny, o.nx))
    b[..., : o.nx] = rng.standard_normal((o.nzp, o.ny, o.nx))
    got = o.dealiasing_2d_fftw(a, b, ref_compat)[..., : o.nx]
    assert np.max(np.abs(got - brute_dealias(o, a, b, ref_compat))) < 1e-13


def test_dealiasing_ref_compat_q1():
    """The reference's ky placement (dealiasing.f90:329-330, 435) is not the 3/2 rule (Q1):
    the product of cos(ky y) with itself comes back with an O(1) error."""
    o = small(nx=16, ny=8, dealiasflag=1)
    a = phys(o, lambda x, y, z: np.cos(2 * y) + 0 * x + 0 * z)
    exact = (a * a)[..., : o.nx]
    got = o.dealiasing_2d_fftw(a, a, ref_compat=True)[..., : o.nx]
    assert np.max(np.abs(got - exact)) > 0.1
    good = o.dealiasing_2d_fftw(a, a, ref_compat=False)[..., : o.nx]
    assert np.max(np.abs(good - exact)) < 1e-14


# -------------------------------------------------------------- the step
def test_step_mean_flow_euler_closed_form():
    """First Euler step of a pure base flow: the (0,0) mode solves
    (d2 - 1/(nu dt)) u = -ubarpp with u = 0 at both walls (navierstokes.f90:1282-1296, bctype 0)."""
    o = small(nx=8, ny=4, nz=40, bctype=0, xnu=1e-3, dt=1e-2)
    z = o.zpts
    o.ubar = np.tanh(z / 0.2)
    o.ubarp = (1 - o.ubar ** 2) / 0.2
    o.ubarpp = -2 * o.ubar * (1 - o.ubar ** 2) / 0.04
    zero = np.zeros((o.nzp, o.ny, o.nx))
    st = o.new_state(zero, zero, zero)
    o.timestep(st)
    s = 1.0 / (o.cfg.xnu * o.cfg.dt)
    A = o.d2[1:o.nz, 1:o.nz] - s * np.eye(o.nzm)
    u_in = np.linalg.solve(A, -o.ubarpp[1:o.nz])
    got = st["uf"][:, 0, 0]
    assert np.max(np.abs(got[1:o.nz] - u_in)) < 1e-12 * np.max(np.abs(u_in))
    assert got[0] == 0 and got[-1] == 0
    assert np.max(np.abs(st["vf"])) == 0 and np.max(np.abs(st["wf"])) < 1e-18


def test_step_projection_is_divergence_free():
    import lsbflow_amd as L
    case = L.cases.CASES["c1_dns"].with_(nx=16, ny=8, nz=40, pert_rms=0.05)
    keys = OracleConfig.__dataclass_fields__
    o = Oracle(OracleConfig(**{k: v for k, v in case.as_dict().items() if k in keys}))
    o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(case)
    u, v, w = L.cases.initial_condition(case)
    assert o.divcheck(u, v, w) < 1e-14                                  # the IC is divergence free
    st = o.new_state(u, v, w)
    for _ in range(3):
        o.timestep(st)
    assert o.divcheck(st["uf"], st["vf"], st["wf"]) < 1e-10


def _interior_solve(o, k2, sigma, rhs):
    """u with u = 0 at both walls and (d2 - k2 - sigma) u = rhs on the interior rows."""
    A = o.d2[1:o.nz, 1:o.nz] - (k2 + sigma) * np.eye(o.nzm)
    u = np.zeros(o.nzp, dtype=np.result_type(rhs, float))
    u[1:o.nz] = np.linalg.solve(A, rhs[1:o.nz])
    return u


def test_temperature_advection_closed_form():
    """First (Euler / backward-Euler) step of the temperature equation under a uniform
    stream U (ubar = U, no perturbation, tbar = 0): the skew-symmetric form
    -1/2 u.grad T - 1/2 div(u T) (navierstokes.f90:800-845) is -U dT/dx, and
    (T1 - T0)/dt = kappa (d2 - kx^2) T1 - i kx U T0 with kappa = nu/Pr and T1 = 0 at the
    walls (navierstokes.f90:1251-1275)."""
    o = small(nx=16, ny=4, nz=24, bctype=0, xnu=2e-2, dt=5e-3, dealiasflag=0, prtl=0.7)
    U = 1.3
    o.ubar = np.full(o.nzp, U)
    o.ubarp = np.zeros(o.nzp)
    o.ubarpp = np.zeros(o.nzp)
    z = o.zpts
    g = (1 - z ** 2) * (0.4 + z)
    T0 = o.horfft_fwd(phys(o, lambda x, y, zz: np.cos(2 * x) * ((1 - zz ** 2) * (0.4 + zz))))
    zero = np.zeros((o.nzp, o.ny, o.nx))
    st = o.new_state(zero, zero, zero, tempf=T0)
    o.timestep(st)
    kx = o.wavx[4]
    assert kx == 2.0 * 2 * math.pi / o.cfg.xlen
    c0 = T0[:, 0, 4] + 1j * T0[:, 0, 5]                 # the cos(2x) coefficient
    assert np.max(np.abs(c0 - 0.5 * g)) < 1e-14
    kap, dt = o.cfg.xnu / o.cfg.prtl, o.cfg.dt
    rhs = -(c0 / dt - 1j * kx * U * c0) / kap
    want = _interior_solve(o, kx * kx, 1.0 / (kap * dt), rhs)
    got = st["tempf"][:, 0, 4] + 1j * st["tempf"][:, 0, 5]
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    mask = np.ones(st["tempf"].shape, bool)
    mask[:, 0, 4:6] = False
    assert np.max(np.abs(st["tempf"][mask])) < 1e-15
    assert np.max(np.abs(st["uf"])) < 1e-15 and np.max(np.abs(st["wf"])) < 1e-15


def test_buoyancy_closed_form():
    """No flow, a (0, 0)-mode temperature T0(z) and gravity coex*g: after the first step
    T1 solves (T1 - T0)/dt = kappa d2 T1, and the streamwise velocity responds to the
    buoyancy force coex*gx*T1: u1/dt = nu d2 u1 + coex gx T1 (navierstokes.f90:1296-1303).
    The wall-normal force is balanced by the pressure on the (0, 0) mode (w stays 0)."""
    o = small(nx=8, ny=4, nz=30, bctype=0, xnu=1e-2, dt=1e-2, dealiasflag=0, prtl=0.9,
              coex=3e-3, gx=2.0, gz=-9.81)
    z = o.zpts
    prof = (1 - z ** 2) * np.exp(z)
    T0 = o.horfft_fwd(phys(o, lambda x, y, zz: (1 - zz ** 2) * np.exp(zz) + 0 * x + 0 * y))
    assert np.max(np.abs(T0[:, 0, 0] - prof)) < 1e-14
    zero = np.zeros((o.nzp, o.ny, o.nx))
    st = o.new_state(zero, zero, zero, tempf=T0)
    o.timestep(st)
    kap, nu, dt = o.cfg.xnu / o.cfg.prtl, o.cfg.xnu, o.cfg.dt
    T1 = _interior_solve(o, 0.0, 1.0 / (kap * dt), -prof / (kap * dt))
    assert np.max(np.abs(st["tempf"][:, 0, 0] - T1)) < 1e-12 * np.max(np.abs(T1))
    u1 = _interior_solve(o, 0.0, 1.0 / (nu * dt), -o.cfg.coex * o.cfg.gx * T1 / nu)
    assert np.max(np.abs(st["uf"][:, 0, 0] - u1)) < 1e-12 * np.max(np.abs(u1))
    assert np.max(np.abs(st["wf"])) < 1e-12 * np.max(np.abs(u1)) and np.max(np.abs(st["vf"])) == 0


def test_temperature_base_profile_source():
    """A base temperature profile with tbarpp != 0 and no flow: the first step adds the
    base-profile diffusion, (T1 - 0)/dt = kappa (d2 T1 + tbarpp) (navierstokes.f90:1266-1275,
    the analogue of ubarpp for the velocity)."""
    o = small(nx=8, ny=4, nz=30, bctype=0, xnu=1e-2, dt=1e-2, dealiasflag=0, prtl=0.8)
    z = o.zpts
    o.tbar = 1.0 + 0.5 * np.tanh(z / 0.3)
    o.tbarp = 0.5 * (1 - np.tanh(z / 0.3) ** 2) / 0.3
    o.tbarpp = -2 * np.tanh(z / 0.3) * 0.5 * (1 - np.tanh(z / 0.3) ** 2) / 0.09
    zero = np.zeros((o.nzp, o.ny, o.nx))
    st = o.new_state(zero, zero, zero)
    o.timestep(st)
    kap, dt = o.cfg.xnu / o.cfg.prtl, o.cfg.dt
    T1 = _interior_solve(o, 0.0, 1.0 / (kap * dt), -o.tbarpp)
    got = st["tempf"][:, 0, 0]
    assert np.max(np.abs(got[1:o.nz] - T1[1:o.nz])) < 1e-12 * np.max(np.abs(T1))
    assert got[0] == 0 and got[-1] == 0


# ------------------------------------------- boundary test / integration routines
def test_dealiasf_is_the_spectral_core():
    """dealiasing_2d_fftw (dealiasing.f90:262-447) is horfft(+1) of dealiasf_2d_fftw
    (dealiasing.f90:449-618) of the forward transforms; on two single modes (textbook
    placement) dealiasf gives the coefficients of their product,
    cos(2x)cos(2y) cos(x)sin(4y) (four sine modes of amplitude 1/8)."""
    o = small(nx=16, ny=8, dealiasflag=1)
    rng = np.random.default_rng(12)
    a = phys(o, lambda x, y, z: rng.standard_normal(x.shape[:2] + (x.shape[2],)) + 0 * y + 0 * z)
    b = phys(o, lambda x, y, z: np.cos(x + 3 * y) + 0 * z)
    for rc in (True, False):
        got = o.horfft_inv(o.dealiasf_2d_fftw(o.horfft_fwd(a), o.horfft_fwd(b), rc))
        want = o.dealiasing_2d_fftw(a, b, rc)
        assert np.max(np.abs(got - want)[..., : o.nx]) < 1e-14
    a1 = phys(o, lambda x, y, z: np.cos(2 * x) * np.cos(2 * y) + 0 * z)
    b1 = phys(o, lambda x, y, z: np.cos(x) * np.sin(4 * y) + 0 * z)
    got = o.dealiasf_2d_fftw(o.horfft_fwd(a1), o.horfft_fwd(b1), False)
    assert np.max(np.abs(got - o.horfft_fwd(phys(o, lambda x, y, z: (np.cos(2 * x) * np.cos(2 * y) * np.cos(x)
                                                                        * np.sin(4 * y)) + 0 * z)))) < 1e-15


def test_test_fft_nyquist_closed_form():
    """test_fft (horfft.f90:7-58): a kx-Nyquist field is dropped by the forward transform, so
    the round trip returns 0 and the mean error is the mean |u| = 1; a resolved field
    comes back to rounding."""
    o = small(nx=16, ny=8, nz=8)
    alt = phys(o, lambda x, y, z: np.cos(math.pi * np.round(x * o.nx / (2 * math.pi))) + 0 * y + 0 * z)
    assert abs(o.test_fft(alt) - 1.0) < 1e-15
    ok = phys(o, lambda x, y, z: np.sin(3 * x - 2 * y) * (1 + z) + 0.5)
    assert o.test_fft(ok) < 1e-15


@pytest.mark.parametrize("mapping,zlen", [(0, 1.0), (3, 2.0)])
def test_zinteg_closed_forms(mapping, zlen):
    """zinteg (derivatives.f90:9-70): the trapezoid rule integrates a linear profile exactly
    from the top (rev = 0) and from plane 1 (rev = 1), the two adding to the full integral;
    per = 1 is Fejer's second rule on the interior Chebyshev points t_j = cos(pi j/nz),
    exact for polynomials of degree <= nz - 2: with tmp = p(t) gf / (0.5 xl) the value at
    plane 2 is the integral of p over [-1, 1]."""
    o = small(nx=4, ny=2, nz=24, mapping=mapping, zlen=zlen)
    z = o.zpts
    f = (0.7 + 1.3 * z)[:, None, None] * np.ones((1, o.ny, o.nx))
    F = 0.7 * z + 0.65 * z ** 2
    top = o.zinteg(f, mapping, 0, 0)
    bot = o.zinteg(f, mapping, 0, 1)
    assert np.max(np.abs(top[:, 0, 0] - (F - F[-1]))) < 1e-13
    assert np.max(np.abs(bot[:, 0, 0] - (F[0] - F))) < 1e-13
    assert np.max(np.abs(top + bot - top[0])) < 1e-13
    t = np.cos(np.pi * np.arange(o.nzp) / o.nz)
    p = t ** 4 + 0.3 * t - 0.2 * t ** 7
    tmp = np.zeros((o.nzp, o.ny, o.nx))
    tmp[1:o.nz] = (p * o.gf / (0.5 * o.xl))[1:o.nz, None, None]
    q = o.zinteg(tmp, mapping, 1, 0)
    assert np.max(np.abs(q[1] - 2.0 / 5.0)) < 1e-13          # int t^4 = 2/5, odd parts vanish
    assert np.all(q[-1] == 0) and np.all(q[0] == 0)
