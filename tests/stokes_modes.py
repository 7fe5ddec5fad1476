"""Exact decaying Stokes modes of the channel z in [0, 2] (mapping 3), written straight into
the spectral layout (kx mode 1 in columns 2, 3 of ky row 0: real and imaginary parts with
the 1/2 of a real cosine/sine), without any of the solver's transforms.

* shear mode:  v = exp(-lam t) sin(kx x) sin(n pi z / 2), u = w = p = 0,
  lam = nu (kx^2 + (n pi / 2)^2); its nonlinear term vanishes exactly (every product holds a
  y-derivative of a ky = 0 field or a zero component) and it is divergence free, so the step
  advances only its viscous part: the amplitude follows the scalar BDF3 recurrence
  (11/6 + lam dt) a_{n+1} = 3 a_n - 3/2 a_{n-1} + 1/3 a_{n-2}.
* pressure mode (ky = 0, zeta = z - 1): w = W(zeta) cos(kx x), u = -W'(zeta)/kx sin(kx x),
  p = P(zeta) cos(kx x), W = a1 cosh(kx zeta) + a2 sinh(kx zeta) + a3 cos(mu zeta)
  + a4 sin(mu zeta), lam = nu (kx^2 + mu^2), P from the x-momentum balance
  -lam u = -dp/dx + nu (D^2 - kx^2) u.  Two boundary sets:
    "noslip":   W = W' = 0 at both walls (u = w = 0);
    "freeflow": W = W' = 0 at the wall (zeta = -1), W = W''' = 0 at the ceiling
                (w = 0, D^2 u = 0: the ceiling condition of navierstokes.f90:1335-1347).
  mu is the smallest positive root of the 4 x 4 boundary determinant.
"""

import math

import numpy as np
from scipy.optimize import brentq

KX = 1.0


def _basis(mu, zeta, der):
    k = KX
    ck, sk = np.cosh(k * zeta), np.sinh(k * zeta)
    cm, sm = np.cos(mu * zeta), np.sin(mu * zeta)
    return {0: (ck, sk, cm, sm),
            1: (k * sk, k * ck, -mu * sm, mu * cm),
            3: (k ** 3 * sk, k ** 3 * ck, mu ** 3 * sm, -mu ** 3 * cm)}[der]


def _conditions(mu, kind):
    rows = [_basis(mu, -1.0, 0), _basis(mu, -1.0, 1), _basis(mu, 1.0, 0),
            _basis(mu, 1.0, 1 if kind == "noslip" else 3)]
    return np.array([[float(v) for v in r] for r in rows])


def pressure_mode_mu(kind):
    f = lambda m: np.linalg.det(_conditions(m, kind))   # noqa: E731
    grid = np.linspace(0.3, 8.0, 4000)
    vals = [f(g) for g in grid]
    for i in range(len(grid) - 1):
        if vals[i] * vals[i + 1] < 0:
            return brentq(f, grid[i], grid[i + 1], xtol=1e-15)
    raise RuntimeError("no root")


def _spectral(prof_sin, prof_cos, nzp, ny, nx):
    a = np.zeros((nzp, ny, nx))
    if prof_cos is not None:
        a[:, 0, 2] = 0.5 * prof_cos
    if prof_sin is not None:
        a[:, 0, 3] = -0.5 * prof_sin
    return a


def shear_mode(z, ny, nx, nu, n=1):
    """(v0 spectral, lam) of the shear mode on the collocation points z."""
    m = n * math.pi / 2.0
    lam = nu * (KX * KX + m * m)
    return _spectral(np.sin(m * z), None, len(z), ny, nx), lam


def pressure_mode(z, ny, nx, nu, kind, amp=1e-10):
    """(u0, w0, p0 spectral, lam) of the pressure mode, max |W| = amp."""
    mu = pressure_mode_mu(kind)
    _, _, vt = np.linalg.svd(_conditions(mu, kind))
    a = vt[-1]
    zeta = z - 1.0
    W, DW, D3W = (sum(ai * b for ai, b in zip(a, _basis(mu, zeta, d))) for d in (0, 1, 3))
    s = amp / np.max(np.abs(W))
    W, DW, D3W = W * s, DW * s, D3W * s
    lam = nu * (KX * KX + mu * mu)
    P = (lam * DW / KX + nu * (D3W - KX * KX * DW) / KX) / KX
    nzp = len(z)
    return (_spectral(-DW / KX, None, nzp, ny, nx), _spectral(None, W, nzp, ny, nx),
            _spectral(None, P, nzp, ny, nx), lam)


def bdf3_amplitude(lam, dt, nsteps):
    """Scalar BDF3 recurrence from the exact levels a_0 = 1, a_-1 = e^{lam dt}, a_-2 = e^{2 lam dt}."""
    a2, a1, a0 = math.exp(2 * lam * dt), math.exp(lam * dt), 1.0
    for _ in range(nsteps):
        a2, a1, a0 = a1, a0, (3.0 * a0 - 1.5 * a1 + a2 / 3.0) / (11.0 / 6.0 + lam * dt)
    return a0


def exact_history(fields, lam, dt):
    """State arrays of the exact start (nstart = 1): the fields at t = -dt and -2 dt."""
    return [f * math.exp(lam * dt) for f in fields], [f * math.exp(2 * lam * dt) for f in fields]
