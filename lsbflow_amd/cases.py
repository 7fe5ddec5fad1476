"""Benchmark/parity configurations and their seeded initial conditions.

The reference driver has no tanh base flow, no seeded perturbation and no
steady ceiling suction/blowing profile (its ``ic=1`` hook at init.f90:443-445
is empty and ``bound`` only knows bctype 1/2, navierstokes.f90:241-329).  This
module is the driver-side patch that supplies them, identically for the HIP
library and for the CPU oracle:

* base flow  ``ubar = tanh(z/delta)`` with analytic ``ubarp``/``ubarpp``
  (delta = 0.2), wall at z = 0 (reference index k = nzp), ceiling at z = zlen;
* perturbation = curl of a random vector potential
  ``A_c(x,y,z) = env(z) * sum c * T_n(xi) * exp(i(kx x + ky y))`` with N(0,1)
  coefficients on |kx index| <= nx/8, |ky index| <= ny/8, Chebyshev degree
  n <= nz/4 and ``env = (z(zlen-z))^2/(zlen/2)^4`` (A and dA/dz vanish at both
  walls, so the velocity does too); divergence free to rounding because the
  curl is formed with commuting spectral operators; scaled to an rms of 1e-3
  over the physical grid;
* ceiling profile (bctype 3): ``w(x, zlen) = samp*(-G(0.4 Lx) + G(0.7 Lx))``,
  ``G(c) = exp(-((x - x0 - c)/(0.05 Lx))^2)`` on the reference's ``xpts``.

The LES workloads (suction amplitude 0.3 at Re 5e4-1e5) run with two switches
on, applied identically in the library and the oracle:

* ``dct_exact = 1``: the Chebyshev derivative scales the top DCT coefficient
  like the others (chebyshev.f90:242 leaves it unscaled, which is exact only for
  xl = 1; the channel has xl = 2).  With the reference operator the
  non-homogeneous step has a growing mode of factor 1.585 per step at these
  viscosities (single-pencil step map), whatever the ceiling condition;
* ``top_bc = 2``: the tangential ceiling values are ``(1 - m(x)) f + m(x) g``
  with ``m`` the profile's blowing Gaussian G(0.4 Lx), ``f`` the reference's
  free-flow value from D^2(u,v) -> 0 (navierstokes.f90:1335-1347) and ``g`` the
  wall's Zang-Hussaini value (:1329-1331, 1365-1367), blended in physical space.
  The free-flow value advances the ceiling with the interior's wall-normal
  derivative, an outflow treatment: where the profile blows fluid in it
  extrapolates against the incoming characteristic and the run diverges within
  ~60 steps.  A Dirichlet ceiling everywhere (``top_bc = 1``) instead grows a
  suction boundary layer of thickness nu/V ~ 1e-5, twenty times finer than the
  first Chebyshev spacing, which diverges after ~200 steps.  The blend
  prescribes the incoming fluid's tangential velocity (u' = 0, i.e. the free
  stream ubar(zlen)) where fluid enters and keeps the outflow condition
  elsewhere; it stays bounded (2000 steps at reduced horizontal size).

The uneven static-Smagorinsky ensemble adds a third switch:

* ``nut_ceil0 = 1``: the static model's filter width is zero on the ceiling row, as
  it is on the wall row (sgsterm.f90:95-103 never sets delta(nzp)).  Where the blend
  prescribes the tangential velocity, the impulsively started blowing leaves a
  Stokes layer thinner than the first Chebyshev spacing; its strain (~samp/dz) times
  the undamped (Cs delta)^2 of the coarse horizontal grids gives the ceiling row an
  explicit eddy viscosity whose diffusion grows without bound at t ~ 2e-3 whatever
  the time step (dt from 2.5e-4 down to 2e-6, same blow-up time).

The DNS workload keeps the reference's operator and ceiling condition.

The random numbers come from SplitMix64 + Box-Muller so that every consumer
reproduces the same bits from the same seed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, asdict, replace

import numpy as np
from numpy.polynomial import chebyshev as npcheb

PI = 4.0 * math.atan(1.0)


@dataclass
class Case:
    """Namelist values of one configuration (reference names; common.f90:20-107)."""

    name: str
    nx: int
    ny: int
    nz: int
    xlen: float = 2.0 * PI
    ylen: float = PI
    zlen: float = 2.0
    x0: float = 0.0
    u0: float = 1.0
    xnu: float = 1.0e-3
    mapping: int = 3
    bctype: int = 3
    samp: float = 0.0
    sgsflag: int = 0
    sgsmodel: int = 0
    a: float = 1.0
    b: float = 2.0
    cs: float = 0.1
    kc: float = 0.6
    filt_order: float = 1.0
    filt_amp: float = 20.0
    dealiasflag: int = 2
    dct_exact: int = 0
    top_bc: int = 0
    dyn_avg: int = 0
    nut_ceil0: int = 0
    dt: float = 1.0e-3
    nsteps: int = 10
    delta: float = 0.2
    pert_rms: float = 1.0e-3
    seed: int = 1234

    @property
    def nzp(self) -> int:
        return self.nz + 1

    @property
    def points(self) -> int:
        """Grid points per field (nx * nzp * ny)."""
        return self.nx * self.nzp * self.ny

    def with_(self, **kw) -> "Case":
        return replace(self, **kw)

    def as_dict(self) -> dict:
        return asdict(self)


# Workload names list streamwise x planes (nx), wall-normal points (nzp = nz + 1) and
# spanwise points (ny): "512x193x256" is nx=512, nz=192, ny=256.
CASES = {
    "c1_dns": Case("c1_dns", nx=64, ny=32, nz=64, xnu=1.0e-3, dt=1.0e-3, nsteps=10),
    "c2_dyn_les": Case("c2_dyn_les", nx=256, ny=128, nz=128, xnu=0.2 / 5.0e4, samp=0.3,
                       sgsflag=1, sgsmodel=1, a=1.0, b=2.0, dt=5.0e-4, nsteps=100, dct_exact=1, top_bc=2),
    "c4_sigma_les": Case("c4_sigma_les", nx=512, ny=256, nz=192, xnu=0.2 / 1.0e5, samp=0.3,
                         sgsflag=1, sgsmodel=2, dt=2.5e-4, nsteps=100, dct_exact=1, top_bc=2),
    "c4_dyn_les": Case("c4_dyn_les", nx=512, ny=256, nz=192, xnu=0.2 / 1.0e5, samp=0.3,
                       sgsflag=1, sgsmodel=1, a=1.0, b=2.0, dt=2.5e-4, nsteps=100, dct_exact=1, top_bc=2),
}


def ensemble_members(n: int = 16, seed: int = 5150, nsteps: int = 200):
    """The uneven ensemble: n independent static-Smagorinsky (Cs = 0.1) LES instances whose
    grids are drawn from nx in {64, 96, 128, 192, 256}, nzp in {65, 97, 129} and ny in
    {32, 64, 128}, Re log-uniform in [1e4, 1e5] on the delta = 0.2 tanh thickness, ceiling
    amplitude uniform in [0.1, 0.4], each with its own seed for the initial condition.  The
    draws come from SplitMix64 (uniforms in (0, 1]) so that every consumer gets the same
    members.  Each member's dt is the smaller of 2.5e-4 and 0.15 dz_top / samp, a
    wall-normal Courant number of 0.15 for the blowing speed over the first Chebyshev
    spacing (members at 0.27 and 0.33 diverge within 12 steps, 0.20 stays bounded);
    the ceiling row carries no eddy viscosity (``nut_ceil0``)."""
    r = splitmix64(seed, 6 * n)
    uni = ((r >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    nxs, nzps, nys = (64, 96, 128, 192, 256), (65, 97, 129), (32, 64, 128)
    out = []
    for i in range(n):
        u = uni[6 * i: 6 * i + 6]
        nx = nxs[min(int(u[0] * len(nxs)), len(nxs) - 1)]
        nzp = nzps[min(int(u[1] * len(nzps)), len(nzps) - 1)]
        ny = nys[min(int(u[2] * len(nys)), len(nys) - 1)]
        re = 10.0 ** (4.0 + u[3])
        samp = 0.1 + 0.3 * u[4]
        dz_top = 0.5 * Case.zlen * (1.0 - math.cos(PI / (nzp - 1)))
        dt = min(2.5e-4, 0.15 * dz_top / samp)
        out.append(Case(f"c5_m{i:02d}", nx=nx, ny=ny, nz=nzp - 1, xnu=0.2 / re, samp=samp, sgsflag=1, sgsmodel=0,
                        cs=0.1, dt=dt, nsteps=nsteps, dct_exact=1, top_bc=2, nut_ceil0=1,
                        seed=seed + 7919 * (i + 1)))
    return out


# ---------------------------------------------------------------- random numbers
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs (uint64) starting from ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normal_stream(seed: int, n: int) -> np.ndarray:
    """n N(0,1) doubles by Box-Muller on SplitMix64 uniforms in (0, 1]."""
    m = (n + 1) // 2
    r = splitmix64(seed, 2 * m)
    u = ((r >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2.0 * PI * u2)
    z[1::2] = rad * np.sin(2.0 * PI * u2)
    return z[:n]


# ------------------------------------------------------------------ grid helpers
def zpts_mapping3(case: Case) -> np.ndarray:
    """init.f90:231-237 (mapping 3): z = zlen/2 + zlen/2 cos(pi k/nz), k = 0..nz.

    This is also the distance from the bottom wall for mapping 0 (init.f90:188-196,
    z = zlen/2 cos(pi k/nz), the same channel shifted by zlen/2), so the base flow
    and the seeded perturbation are the same arrays for both channel mappings."""
    xl = case.zlen
    nz = case.nz
    return np.array([xl / 2.0 + xl / 2.0 * math.cos(PI * (k + 0.0) / float(nz)) for k in range(nz + 1)])


def xpts(case: Case) -> np.ndarray:
    """init.f90:153-156 (dx = xlen/(nx-1), reference quirk Q4)."""
    dx = case.xlen / (case.nx - 1.0)
    return np.array([case.x0 + i * dx for i in range(case.nx)])


def base_flow(case: Case):
    """ubar, ubarp, ubarpp on the Chebyshev points (analytic tanh profile)."""
    z = zpts_mapping3(case)
    d = case.delta
    t = np.tanh(z / d)
    s2 = 1.0 - t * t
    return t, s2 / d, -2.0 * t * s2 / (d * d)


def ceiling_profile(case: Case) -> np.ndarray:
    """Physical ceiling w(x) for bctype 3 (length nx)."""
    L = case.xlen
    xr = xpts(case) - case.x0
    g1 = np.exp(-(((xr - 0.4 * L) / (0.05 * L)) ** 2))
    g2 = np.exp(-(((xr - 0.7 * L) / (0.05 * L)) ** 2))
    return case.samp * (g2 - g1)


def _spectral_to_physical(fh: np.ndarray, nx: int) -> np.ndarray:
    """(nzp, ny, nx/2) complex coefficients -> (nzp, ny, nx) physical (horfft +1 convention)."""
    X = np.fft.ifft(fh, axis=-2, norm="forward")
    Xf = np.zeros(X.shape[:-1] + (nx // 2 + 1,), dtype=np.complex128)
    Xf[..., : nx // 2] = X
    return np.fft.irfft(Xf, n=nx, axis=-1, norm="forward")


def initial_condition(case: Case):
    """Seeded divergence-free perturbation -> (uf, vf, wf) spectral (nzp, ny, nx) float64."""
    if case.mapping not in (0, 3):
        raise ValueError("the seeded perturbation is defined for the channel mappings 0 and 3")
    nx, ny, nz = case.nx, case.ny, case.nz
    nzp = nz + 1
    z = zpts_mapping3(case)
    xl = case.zlen
    xi = 2.0 * z / xl - 1.0
    env = (z * (xl - z)) ** 2 / (xl / 2.0) ** 4
    denv = 2.0 * z * (xl - z) * (xl - 2.0 * z) / (xl / 2.0) ** 4
    mx, my, nc = max(1, nx // 8), max(1, ny // 8), max(1, nz // 4)
    alpha = 2.0 * PI / case.xlen
    beta = 2.0 * PI / case.ylen
    jys = list(range(-my, my + 1))
    ncoef = 3 * (mx + 1) * len(jys) * (nc + 1) * 2
    g = normal_stream(case.seed, ncoef).reshape(3, mx + 1, len(jys), nc + 1, 2)
    coef = g[..., 0] + 1j * g[..., 1]
    # Chebyshev Vandermonde matrices: T[k, n] = T_n(xi_k), dT[k, n] = T_n'(xi_k) d(xi)/dz
    T = npcheb.chebvander(xi, nc)
    dT = np.stack([npcheb.chebval(xi, npcheb.chebder(np.eye(nc + 1)[n])) for n in range(nc + 1)], axis=1)
    dT = dT * (2.0 / xl)
    P = np.einsum("kn,cmjn->ckjm", T, coef)          # (3, nzp, 2my+1, mx+1), rows jy = -my..my
    dP = np.einsum("kn,cmjn->ckjm", dT, coef)
    A = env[None, :, None, None] * P
    dA = denv[None, :, None, None] * P + env[None, :, None, None] * dP
    # Hermitian symmetry of the kx = 0 column (real field): A(0, -jy) = conj A(0, jy)
    for arr in (A, dA):
        for j in range(1, my + 1):
            arr[:, :, my - j, 0] = np.conj(arr[:, :, my + j, 0])
        arr[:, :, my, 0] = arr[:, :, my, 0].real
    kx = (np.arange(mx + 1) * alpha)[None, None, :]
    ky = (np.arange(-my, my + 1) * beta)[None, :, None]
    ax, ay, az = A
    dax, day, daz = dA
    band = (1j * ky * az - day, dax - 1j * kx * az, 1j * kx * ay - 1j * ky * ax)
    # rms over the physical grid by Parseval: mean(u^2) = sum_kx=0 |u|^2 + 2 sum_kx>0 |u|^2 per plane
    ms = 0.0
    for f in band:
        e = np.abs(f) ** 2
        ms += (float(np.sum(e[..., 0])) + 2.0 * float(np.sum(e[..., 1:]))) / nzp
    rms = math.sqrt(ms / 3.0)
    s = case.pert_rms / rms
    rows = np.array([jy % ny for jy in jys])
    out = []
    for f in band:
        full = np.zeros((nzp, ny, nx // 2), dtype=np.complex128)
        full[:, rows, : mx + 1] = f * s
        out.append(full.view(np.float64).reshape(nzp, ny, nx))
    return tuple(out)


def helmholtz_rhs(nzp: int, nx: int, ny: int, nfields: int, seed: int, zlen: float = 2.0):
    """C3 right-hand sides: Chebyshev coefficients N(0,1) exp(-0.1 n), to collocation.
    Returns (nfields, nzp, ny, nx) float64 in the spectral layout."""
    nz = nzp - 1
    xi = np.array([math.cos(PI * k / nz) for k in range(nzp)])
    ncol = nx * ny
    g = normal_stream(seed, nfields * nzp * ncol).reshape(nfields, nzp, ncol)
    g *= np.exp(-0.1 * np.arange(nzp))[None, :, None]
    T = npcheb.chebvander(xi, nz)                      # (nzp, nzp): T_n(xi_k)
    out = np.einsum("kn,fnc->fkc", T, g)
    return np.ascontiguousarray(out.reshape(nfields, nzp, ny, nx))
