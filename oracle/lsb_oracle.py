"""FP64 numpy restatement of lsbflow's per-time-step hot path (TEST ORACLE ONLY).

Layout convention (identical to the reference's Fortran memory layout):

* spectral field  Fortran ``qf(nxpl, ny, nzp)``  ->  numpy ``(nzp, ny, nx)``;
  real index ``2m`` / ``2m+1`` hold Re/Im of kx = ``wavx[2m]``.
* physical field  Fortran ``q(nxpp, nypl, nzp)``  ->  numpy ``(nzp, ny, nx+2)``.
* rank-1 vector   Fortran ``(..., 3)``  ->  numpy ``(3, nzp, ...)`` indexed ``[l]``.
* rank-2 tensor   Fortran ``(..., 3, 3)`` element ``(l, m)``  ->  numpy
  ``(3, 3, nzp, ...)`` indexed ``[l][m]`` (the C-ABI buffers of the product are
  Fortran-ordered, i.e. ``[m][l]`` in memory; tests convert).

One MPI rank (P = 1): nxpl = nx, nypl = ny, nzpl = nzp.

Pinning (tests/test_oracle_*.py):
  setup_grid           -- run/grid.dat, bit for bit
  build_cheb_zder_matrix + diagonalize_matrix -- run/output.txt:53-84 eigenvalues
  blasfl + dpdz/dct_diff -- run/paral.dat
  horfft/norm/dfdx/dfdy -- single-mode closed forms, Nyquist drop, round trip
  dct_diff             -- exactness on low Chebyshev polynomials, Q2 top mode
  solve_helmholtz/nhom -- residual of (D2-k2-s)u = rhs, E L E^-1 = D2
  sigma_model          -- singular values of g from numpy SVD
  static_smagorinsky / filter_q_3pt / specfilter_exp_xy / dealiasing -- brute
                          force and closed forms on tiny grids
  navierstokes_update  -- divergence-free projection, Stokes (0,0)-mode decay,
                          linear-operator invariants; parity unpinned against
                          reference-produced fields (the reference cannot run here)
  snapshot_record / read_snapshot -- run/q00019.dat round trip (read, transform,
                          write back to 2 float32 spacings), run/prev_q00001.dat
                          header and base profiles
  restart_record / read_restart -- element order of the writer's loop nest, reader
                          inverse, dimension check (no reference restart file ships)
  maxcalc              -- run/output.txt Max|u+ubar|, x-CFL, Max|rho| lines; closed
                          forms for the CFL planes, fringe range and the dt rule
  wall_shear / nancheck -- single-mode closed forms (cf and its zero crossings);
                          Inf vs NaN
  dealiasf_2d_fftw     -- the composition with dealiasing_2d_fftw; two single modes'
                          product coefficients
  test_fft             -- a kx-Nyquist field returns its mean |u| = 1
  zinteg               -- trapezoid exact on linear profiles, top + wall sums = total;
                          per = 1 is Fejer's second rule (exact on a degree-7 polynomial)
  time_average         -- window (its > istartavg) and scale (q/iomod) on fields that
                          grow linearly with the step
  sponge_lambda        -- the fringe's plateau, zero region and rise against the
                          published Herbst-Henningson form (ftype 1)
  modecut_fc / filterall3pt / tns_automatic_filtering -- one mode times a linear z
                          profile: 1 - (1 - T)^6; the TNS energy ratio of one mode
                          ((1 - T16)/(1 - T12))^12 and both decision branches
  temperature path     -- first step in closed form: advection by a uniform stream
                          (-U dT/dx, kappa = nu/Pr), buoyancy coex*g*T on the (0, 0)
                          mode, base-profile diffusion with tbarpp
  ener / lintest       -- one streamwise mode over ubar (cross term vanishes, mean
                          removed); a scaled state gives (s^2 - 1)/(2 s^2 dt)
  solve_orrsomm        -- run/output_lintest.txt omega (10 digits) and run/energy.dat
                          omega_i of the reference's sweep on its recovered grid;
                          the pencil residual and energy normalisation of the mode
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import scipy.fft as sfft

PI = 4.0 * math.atan(1.0)


@dataclass
class OracleConfig:
    """The namelist values the hot path reads (common.f90:20,65,78,92,107)."""

    nx: int = 64
    ny: int = 32
    nz: int = 64
    xlen: float = 2.0 * PI
    ylen: float = PI
    zlen: float = 2.0
    x0: float = 0.0
    u0: float = 1.0
    xnu: float = 1.0e-3
    prtl: float = 0.71
    coex: float = 0.0
    gx: float = 0.0
    gy: float = 0.0
    gz: float = 0.0
    mapping: int = 3
    blfac: float = 8.0
    bctype: int = 3
    samp: float = 0.0
    istartbound: float = 0.0
    forcing: int = 0
    sgsflag: int = 0
    sgsmodel: int = 0
    a: float = 1.0
    b: float = 2.0
    cs: float = 0.1
    kc: float = 0.6
    filt_order: float = 1.0
    filt_amp: float = 20.0
    dealiasflag: int = 2
    dealias_ref_compat: int = 1
    dct_exact: int = 0
    top_bc: int = 0
    dyn_avg: int = 0
    nut_ceil0: int = 0
    ifrz: int = 0
    nstart: int = 0
    dt: float = 1.0e-3
    fsize: float = 0.15
    ifilt: int = 0
    modfilt: int = 0
    lfsize: float = 0.25
    sfsize: float = 0.70

    def with_(self, **kw) -> "OracleConfig":
        return replace(self, **kw)


class Oracle:
    """Module state (dim, grid, mats, paral, time, core) + the hot-path routines."""

    def __init__(self, cfg: OracleConfig, build_mats: bool = True):
        self.cfg = cfg
        c = cfg
        # init_dim, common.f90:23-54 (P = 1)
        self.nx, self.ny, self.nz = c.nx, c.ny, c.nz
        self.nzp = c.nz + 1
        self.nzm = c.nz - 1
        self.nxh = c.nx // 2
        self.nyh = c.ny // 2
        self.nxpp = c.nx + 2
        self.setup_grid()
        if build_mats:
            self.setup_zder()
        self.ubar = np.zeros(self.nzp)
        self.ubarp = np.zeros(self.nzp)
        self.ubarpp = np.zeros(self.nzp)
        self.wbar = np.zeros(self.nzp)
        self.wbarp = np.zeros(self.nzp)
        self.wbarpp = np.zeros(self.nzp)
        self.tbar = np.zeros(self.nzp)
        self.tbarp = np.zeros(self.nzp)
        self.tbarpp = np.zeros(self.nzp)

    # ------------------------------------------------------------------ setup
    def setup_grid(self):
        """init.f90:130-273 (linstab = 0)."""
        c = self.cfg
        nx, ny, nz, nzp = self.nx, self.ny, self.nz, self.nzp
        twopi = 8.0 * math.atan(1.0)
        self.dx = c.xlen / (nx - 1.0)                       # Q4: xlen/(nx-1)
        self.xpts = np.array([c.x0 + i * self.dx for i in range(nx)])
        self.dy = c.ylen / (ny - 1.0)
        self.ypts = np.array([j * self.dy for j in range(ny)])
        self.alpha = twopi / c.xlen
        wavx = np.zeros(nx)
        for i in range(nx // 2):
            wavx[2 * i] = (i + 0.0) * self.alpha            # (myid*nxhpl + i - 1.)*alpha
            wavx[2 * i + 1] = wavx[2 * i]
        self.wavx = wavx
        self.wavxx = wavx * wavx
        self.beta = twopi / c.ylen
        wavy = np.zeros(ny)
        for j in range(1, ny // 2 + 1):
            wavy[ny - j] = -j * self.beta
            wavy[j - 1] = (j - 1.0) * self.beta
        self.wavy = wavy
        self.wavyy = wavy * wavy
        xi = np.zeros(nzp)
        zpts = np.zeros(nzp)
        gf = np.zeros(nzp)
        self.aa = 0.0
        self.bb = 0.0
        m = c.mapping
        if m == 0:
            self.xl = c.zlen
            for k in range(nzp):
                xi[k] = self.xl / 2.0 * math.cos(PI * (k + 0.0) / nz)
                zpts[k] = xi[k]
                gf[k] = 1.0
        elif m == 1:
            self.xl = 1.0
            self.aa = 1.0
            self.bb = 1.875 * math.sqrt(c.xnu * c.x0 / c.u0)
            for k in range(1, nzp):
                xi[k] = self.xl / 2.0 * (1.0 + math.cos(PI * (k + 0.0) / nz))
                zpts[k] = self.bb * xi[k] / (self.aa - xi[k])
                gf[k] = (self.aa - xi[k]) ** 2 / (self.bb * self.aa)
            xi[0] = 1.0
            zpts[0] = 1.0e20
            gf[0] = 0.0
        elif m == 2:
            self.xl = 1.0
            self.bb = 1.875 * math.sqrt(c.xnu * (c.x0 + c.blfac * c.xlen) / c.u0)
            self.aa = 1.0 + self.bb / c.zlen
            for k in range(nzp):
                xi[k] = self.xl / 2.0 * (1.0 + math.cos(PI * (k + 0.0) / float(nz)))
                zpts[k] = self.bb * xi[k] / (self.aa - xi[k])
                gf[k] = (self.aa - xi[k]) ** 2 / (self.aa * self.bb)
        elif m == 3:
            self.xl = c.zlen
            for k in range(nzp):
                xi[k] = self.xl / 2.0 * math.cos(PI * (k + 0.0) / float(nz))
                zpts[k] = self.xl / 2.0 + xi[k]
                gf[k] = 1.0
        else:
            raise ValueError("mapping must be 0..3")
        self.xi, self.zpts, self.gf = xi, zpts, gf

    # chebyshev.f90:51-159 -------------------------------------------------
    @staticmethod
    def _ff(i, nr):
        if i == 0:
            return 0.0
        v = nr * 0.5 / math.tan(PI * i / (2.0 * nr))
        if i % 2 == 0:
            v = -v
        return v

    @classmethod
    def _xnum(cls, kr, jr, nr):
        if kr == 0:
            if jr == 0:
                return (1.0 / 3.0) * nr * nr + 1.0 / 6.0
            if jr == nr:
                v = 0.5
                if nr % 2 == 1:
                    v = -v
                return v
            t = 1.0 / math.tan(PI * jr / (2.0 * nr))
            s = math.sin(PI * jr / (2.0 * nr))
            return (0.5 * (nr + 0.5) * ((nr + 0.5) + t ** 2) + 1.0 / 8.0
                    - 0.25 / (s ** 2.0) - 0.5 * nr * nr)
        return cls._ff(kr + jr, nr) + cls._ff(kr - jr, nr)

    @staticmethod
    def _den(kr, jr, nr):
        if kr == 0:
            d = 0.5 * nr
            if jr % 2 == 1:
                d = -d
            if jr == 0 or jr == nr:
                d = 1.0
            return d
        d = nr * math.sin(PI * kr / nr)
        if jr == 0 or jr == nr:
            d = d * 2.0
        return d

    def _del(self, k, j, n):
        """chebyshev.f90:51-77 (1-based k, j)."""
        kr, jr, nr = k - 1, j - 1, n - 1
        fac = 1.0
        if kr == nr:
            kr = 0
            jr = nr - jr
            fac = -1.0
        d = fac * self._xnum(kr, jr, nr) / self._den(kr, jr, nr)
        return d * (2.0 / self.xl)

    def build_cheb_zder_matrix(self):
        """chebyshev.f90:9-48: dor, d = diag(gf) dor, d2 = d d."""
        n = self.nzp
        dor = np.zeros((n, n))
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                dor[i - 1, j - 1] = self._del(i, j, n)
        d = dor * self.gf[:, None]
        d2 = d @ d
        return dor, d, d2

    @staticmethod
    def diagonalize_matrix(D2):
        """helmholtz.f90:13-126: D2 = E diag(lamda) E^-1 (dggev with B = I; dgetrf/dgetri)."""
        w, v = np.linalg.eig(D2)
        if np.max(np.abs(w.imag)) > 1e-8 * np.max(np.abs(w.real)):
            raise RuntimeError("complex eigenvalues in D2 interior block")
        lam = w.real.copy()
        e = v.real.copy()
        ei = np.linalg.inv(e)
        return lam, e, ei

    def setup_zder(self):
        """init.f90:276-293."""
        self.dor, self.d, self.d2 = self.build_cheb_zder_matrix()
        nz = self.nz
        self.diag, self.smat, self.simat = self.diagonalize_matrix(self.d2[1:nz, 1:nz].copy())

    # ---------------------------------------------------------- blasfl/setup_ic
    def blasfl(self, eps, iters, alph, zpts, dor, d, d2):
        """init.f90:497-586 (Newton iteration with dgesv)."""
        nzp = zpts.shape[0]
        nzpp = nzp + 1
        f = (-1.0 + np.exp(-zpts * alph)) / alph
        d3 = d2 @ d
        rat = 1.0 / (1.0 + zpts)
        it_done = iters
        for ii in range(1, iters + 1):
            rhs = d2 @ f
            A = np.zeros((nzpp, nzpp))
            A[:nzp, :nzp] = (0.5 * rat * f)[:, None] * d2 + rat[:, None] * d3 + (0.5 * (zpts * rat))[:, None] * d2
            A[np.arange(nzp), np.arange(nzp)] += 0.5 * rat * rhs
            A[:nzp, nzpp - 1] = d[:nzp, nzp - 1]
            A[0, :nzp] = dor[0, :nzp]
            A[nzp - 1, :nzp] = 0.0
            A[nzpp - 1, :nzp] = d[nzp - 1, :nzp]
            A[0, nzpp - 1] = 0.0
            A[nzp - 1, nzpp - 1] = 0.0
            A[nzpp - 1, nzpp - 1] = 0.0
            A[nzp - 1, nzp - 1] = 1.0
            t1 = d3 @ f
            t2 = d2 @ f
            df = np.zeros(nzpp)
            df[:nzp] = -0.5 * rat * f * t2 - 0.5 * zpts * rat * t2 - rat * t1
            df[0] = 0.0
            df[nzp - 1] = 0.0
            df[nzpp - 1] = 0.0
            df = np.linalg.solve(A, df)
            f = f + df[:nzp]
            err = np.sum(np.abs(df[:nzp]))
            if err <= eps:
                it_done = ii
                break
        ubar = 1.0 + d @ f
        wbar = 0.5 * (zpts * ubar - (f + zpts))
        return ubar, wbar, it_done

    def setup_blasius(self):
        """The ic=0 base flow of setup_ic, init.f90:353-424."""
        c = self.cfg
        okr = math.sqrt(c.x0 * c.xnu / c.u0)
        bz = self.zpts / okr
        bd = self.d * okr
        bd2 = bd @ bd
        ubar, wbar, it = self.blasfl(1.0e-11, 100, 0.33, bz, self.dor, bd, bd2)
        ubar = c.u0 * ubar
        wbar = wbar * math.sqrt(c.xnu * c.u0 / c.x0)
        wbar[0] = wbar[1]
        wbar[-1] = 0.0
        self.ubar = ubar
        self.wbar = wbar
        self.ubarp = self.dpdz(ubar[:, None], 1)[:, 0]
        self.ubarpp = self.dpdz(ubar[:, None], 2)[:, 0]
        self.wbarp = self.dpdz(wbar[:, None], 1)[:, 0]
        self.wbarpp = self.dpdz(wbar[:, None], 2)[:, 0]
        self.blasfl_iters = it

    # ------------------------------------------------------------------ horfft
    def horfft_fwd(self, a):
        """horfft(ain, ainf, -1), horfft.f90:94-119: r2c x (Nyquist dropped :303-315),
        c2c y forward, scale 1/(nx*ny).  ``a`` (..., nzp, ny, nxpp) -> (..., nzp, ny, nx)."""
        nx, ny = self.nx, self.ny
        X = np.fft.rfft(a[..., :nx], axis=-1)[..., : nx // 2]
        X = np.fft.fft(X, axis=-2)
        X = X * (1.0 / float(nx * ny))
        X = np.ascontiguousarray(X)
        return X.view(np.float64).reshape(a.shape[:-1] + (nx,))

    def horfft_inv(self, af):
        """horfft(ain, ainf, +1), horfft.f90:121-144: c2c y backward, c2r x (unscaled).
        The kx-Nyquist slot is zero (scratch zeroed at :90); pads are returned as 0."""
        nx, ny = self.nx, self.ny
        af = np.ascontiguousarray(af)
        X = af.view(np.complex128)
        X = np.fft.ifft(X, axis=-2, norm="forward")
        Xf = np.zeros(X.shape[:-1] + (nx // 2 + 1,), dtype=np.complex128)
        Xf[..., : nx // 2] = X
        x = np.fft.irfft(Xf, n=nx, axis=-1, norm="forward")
        out = np.zeros(af.shape[:-1] + (nx + 2,))
        out[..., :nx] = x
        return out

    def norm(self, *fields):
        """horfft.f90:153-224 (in place on spectral fields)."""
        ny, nyh = self.ny, self.nyh
        for f in fields:
            if self.wavx[0] == 0:
                f[:, 0, 1] = 0.0
        if ny > 2:
            for f in fields:
                if self.wavx[0] == 0:
                    for j in range(2, nyh + 1):
                        jp = ny + 2 - j
                        f[:, j - 1, 0] = 0.5 * (f[:, j - 1, 0] + f[:, jp - 1, 0])
                        f[:, j - 1, 1] = 0.5 * (f[:, j - 1, 1] - f[:, jp - 1, 1])
                    for j in range(2, nyh + 1):
                        jp = ny + 2 - j
                        f[:, jp - 1, 0] = f[:, j - 1, 0]
                        f[:, jp - 1, 1] = -f[:, j - 1, 1]
                f[:, nyh, :] = 0.0

    # ------------------------------------------------------------- derivatives
    def dfdx(self, uf):
        """derivatives.f90:261-280."""
        out = np.empty_like(uf)
        w = self.wavx[0::2]
        out[..., 0::2] = -w * uf[..., 1::2]
        out[..., 1::2] = w * uf[..., 0::2]
        return out

    def dfdy(self, uf):
        """derivatives.f90:283-303."""
        out = np.empty_like(uf)
        w = self.wavy[:, None]
        out[..., 0::2] = -w * uf[..., 1::2]
        out[..., 1::2] = w * uf[..., 0::2]
        return out

    def divuv(self, uf, vf):
        """derivatives.f90:394-409."""
        out = np.empty_like(uf)
        wx = self.wavx[0::2]
        wy = self.wavy[:, None]
        out[..., 0::2] = -wx * uf[..., 1::2] - wy * vf[..., 1::2]
        out[..., 1::2] = wx * uf[..., 0::2] + wy * vf[..., 0::2]
        return out

    def dct_diff(self, y, p):
        """chebyshev.f90:204-270 along axis 0 of y (n, ...): DCT-I, recurrence, DCT-I, *gf/2(n-1).

        The reference divides only alpha(2:n-1) by the half interval (chebyshev.f90:242), so
        the top coefficient alpha(n) misses the factor 1/(c_N * half) = 1/xl it needs; that
        is exact for xl = 1 and doubles the derivative of T_N for the channel's xl = 2.
        ``dct_exact = 1`` divides alpha(n) by xl, which makes the operator the exact
        Gauss-Lobatto derivative on every mapping (and changes nothing when xl = 1)."""
        n = y.shape[0]
        half = 0.5 * self.xl
        nlog = 2 * (n - 1)
        beta = np.array(y, dtype=np.float64, copy=True)
        gf = self.gf.reshape((n,) + (1,) * (y.ndim - 1))
        for _ in range(p):
            alpha = sfft.dct(beta, type=1, axis=0)
            alpha[1 : n - 1] = alpha[1 : n - 1] / half
            if self.cfg.dct_exact:
                alpha[n - 1] = alpha[n - 1] / self.xl
            beta = np.zeros_like(alpha)
            beta[n - 2] = 2.0 * (n - 1.0) * alpha[n - 1]
            for k in range(n - 1, 1, -1):               # 1-based k = n-1 .. 2
                beta[k - 2] = beta[k] + 2.0 * (k - 1.0) * alpha[k - 1]
            alpha = sfft.dct(beta, type=1, axis=0)
            beta = alpha * gf / nlog
        return beta

    def _zapply(self, mat, u):
        n = u.shape[0]
        return (mat @ u.reshape(n, -1)).reshape(u.shape)

    def dfdz(self, uf):
        """derivatives.f90:112-132 (dct_diff if nz > 31 else matmul(d, .))."""
        if self.nz > 31:
            return self.dct_diff(uf, 1)
        return self._zapply(self.d, uf)

    def d2fdz2(self, uf):
        """derivatives.f90:161-188."""
        if self.nz > 31:
            return self.dct_diff(uf, 2)
        return self._zapply(self.d2, uf)

    def dpdz(self, u, p):
        """derivatives.f90:135-158 on arrays (n, ...); n <= 31 applies d once for any p."""
        if u.shape[0] > 31:
            return self.dct_diff(u, p)
        return self._zapply(self.d, u)

    # ----------------------------------------------------------------- helmholtz
    def solve_helmholtz(self, rhs, dtnui, diag=None, smat=None, simat=None):
        """helmholtz.f90:130-208: interior rows 2..nz, E^-1, divide, E, zero end rows."""
        diag = self.diag if diag is None else diag
        smat = self.smat if smat is None else smat
        simat = self.simat if simat is None else simat
        nz = self.nz
        n = diag.shape[0]
        out = np.array(rhs, copy=True)
        q = out[1 : n + 1].reshape(n, -1)
        t = simat @ q
        dfac = (-self.wavxx[None, :] - self.wavyy[:, None] - dtnui).reshape(-1)
        t = t / (diag[:, None] + dfac[None, :])
        q = smat @ t
        out[1 : n + 1] = q.reshape((n,) + out.shape[1:])
        out[0] = 0.0
        out[self.nzp - 1] = 0.0
        del nz
        return out

    def lifting_functions(self):
        """helmholtz.f90:240-307: hbot, htop, d2hbot, d2htop per mapping."""
        z = self.zpts
        nzp = self.nzp
        m = self.cfg.mapping
        if m == 0:
            width = z[0] - z[nzp - 1]
            htop = (z - z[nzp - 1]) / width
            hbot = -(z - z[0]) / width
            d2htop = np.zeros(nzp)
            d2hbot = np.zeros(nzp)
        elif m == 1:
            bbot = 50.0
            htop = np.zeros(nzp)
            d2htop = np.zeros(nzp)
            hbot = np.exp(-bbot * z)
            hbot[0] = 0.0
            d2hbot = bbot ** 2.0 * hbot
        elif m == 2:
            h = z[0] - z[nzp - 1]
            bbot = 5.0
            hbot = (np.exp(-bbot * z) - np.exp(bbot * (z - 2.0 * h)))
            hbot = hbot / (1.0 - math.exp(-2.0 * bbot * h))
            d2hbot = bbot ** 2 * hbot
            htop = self.wbar / self.wbar[0]
            htop[nzp - 1] = 0.0
            d2htop = self.wbarpp / self.wbar[0]
        else:
            btop = 50.0
            bbot = 50.0
            h = z[0] - z[nzp - 1]
            hbot = (np.exp(-bbot * z) - np.exp(bbot * (z - 2.0 * h)))
            hbot = hbot / (1.0 - math.exp(-2.0 * bbot * h))
            d2hbot = bbot ** 2.0 * hbot
            htop = np.exp(btop * (z - h)) * (1.0 - np.exp(-2.0 * btop * z))
            htop = htop / (1.0 - math.exp(-2.0 * btop * h))
            d2htop = btop ** 2.0 * htop
        return hbot, htop, d2hbot, d2htop

    def solve_nhom_helmholtz(self, rhs, bc, dtnui):
        """helmholtz.f90:211-358 (Green's-function lifting, only bc planes 1 and nzp used)."""
        hbot, htop, d2hbot, d2htop = self.lifting_functions()
        nzp = self.nzp
        dfac = -self.wavxx[None, :] - self.wavyy[:, None] - dtnui
        r = np.array(rhs, copy=True)
        bcb = bc[nzp - 1]
        bct = bc[0]
        for k in range(nzp):
            abot = d2hbot[k] + dfac * hbot[k]
            r[k] = r[k] - abot * bcb
            atop = d2htop[k] + dfac * htop[k]
            r[k] = r[k] - atop * bct
        r = self.solve_helmholtz(r, dtnui)
        for k in range(nzp):
            r[k] = r[k] + hbot[k] * bcb + htop[k] * bct
        return r

    # ------------------------------------------------------------------ filters
    def specfilter_fourier2d(self, ak, akmax, kc, p, amp):
        """filters.f90:338-360 (vectorised over ak)."""
        akc = akmax * kc
        theta = (np.abs(ak) - akc) / (akmax - akc)
        return np.where(np.abs(ak) >= akc, np.exp(-amp * theta ** (2 * p)), 1.0)

    def specfilter_exp_xy(self, uf, kc, p, amp, filtery=1):
        """filters.f90:281-335: magnitude/phase exponential filter."""
        wavxmax = self.nxh * self.alpha
        wavymax = self.nyh * self.beta
        akmax = wavxmax
        if filtery == 1:
            akmax = math.sqrt(wavxmax ** 2 + wavymax ** 2)
        fr = uf[..., 0::2]
        fi = uf[..., 1::2]
        fmag = np.sqrt(fr * fr + fi * fi)
        nzero = fmag != 0.0
        with np.errstate(invalid="ignore", divide="ignore"):
            cosc = np.where(nzero, fr / np.where(nzero, fmag, 1.0), 1.0)
            sinc = np.where(nzero, fi / np.where(nzero, fmag, 1.0), 0.0)
        ac = np.arccos(cosc)
        phase = np.where(sinc >= 0.0, ac, -ac)
        if filtery == 1:
            ak = np.sqrt(self.wavxx[0::2][None, :] + self.wavyy[:, None])
        else:
            ak = np.broadcast_to(self.wavx[0::2][None, :], (self.ny, self.nxh))
        filt = self.specfilter_fourier2d(ak, akmax, kc, p, amp)
        fmagf = fmag * filt
        out = np.empty_like(uf)
        out[..., 0::2] = np.where(nzero, fmagf * np.cos(phase), 0.0)
        out[..., 1::2] = np.where(nzero, fmagf * np.sin(phase), 0.0)
        return out

    def filterqspec(self, qf):
        """filters.f90:164-192 (qf: (..., nzp, ny, nx))."""
        c = self.cfg
        return self.specfilter_exp_xy(qf, c.kc, c.filt_order, c.filt_amp, 1)

    def filter_z_weights(self, a, b, interpflag=1):
        """filters.f90:396-448: az, bz, cz for the z pass."""
        z = self.zpts
        nzp = self.nzp
        az = np.full(nzp, a)
        bz = np.full(nzp, b)
        cz = np.full(nzp, a)
        if interpflag == 1:
            m = self.cfg.mapping
            if m in (1, 2):
                for k in range(2, nzp):                     # 1-based k = 2..nzp-1
                    dltm = z[k - 1] - z[k - 2]
                    dltp = z[k] - z[k - 1]
                    az[k - 1] = a * 2.0 * dltm / (dltp + dltm)
                    bz[k - 1] = a * 2.0 * (1.0 - dltm / dltp) + b
                    cz[k - 1] = a * 2.0 * dltm * dltm / (dltp * (dltp + dltm))
                if m == 1:
                    az[1] = 0.0
                    bz[1] = 2.0 * a + b
                    cz[1] = 0.0
            else:
                for k in range(2, (nzp + 1) // 2):         # 2 .. (nzp+1)/2-1
                    dltm = z[k - 1] - z[k - 2]
                    dltp = z[k] - z[k - 1]
                    az[k - 1] = a * 2.0 * dltm / (dltp + dltm)
                    bz[k - 1] = a * 2.0 * (1.0 - dltm / dltp) + b
                    cz[k - 1] = a * 2.0 * dltm * dltm / (dltp * (dltp + dltm))
                for k in range((nzp + 1) // 2 + 1, nzp):   # (nzp+1)/2+1 .. nzp-1
                    dltm = z[k - 1] - z[k - 2]
                    dltp = z[k] - z[k - 1]
                    az[k - 1] = a * 2.0 * dltp * dltp / (dltm * (dltm + dltp))
                    bz[k - 1] = a * 2.0 * (1.0 - dltp / dltm) + b
                    cz[k - 1] = a * 2.0 * dltp / (dltm + dltp)
                mid = (nzp + 1) // 2 - 1
                az[mid] = a
                bz[mid] = b
                cz[mid] = a
            az[0] = 0.0
            bz[0] = 2.0 * a + b
            cz[0] = 0.0
            az[nzp - 1] = 0.0
            bz[nzp - 1] = 2.0 * a + b
            cz[nzp - 1] = 0.0
        return az, bz, cz

    def filter_q_3pt(self, q, a, b, interpflag=1):
        """filters.f90:363-458 + filtery_q_3pt :461-613 at P = 1.
        q: (..., nzp, ny, nxpp) physical; returns the filtered field, pads = 0.
        x: periodic (a,b,a)/(2a+b); y: periodic; z: interpolating weights on
        interior rows only (k = 1 and nzp keep the xy-filtered value, Q7)."""
        nx = self.nx
        s = 2.0 * a + b
        qx = q[..., :nx]
        qh = (a * np.roll(qx, 1, axis=-1) + b * qx + a * np.roll(qx, -1, axis=-1)) / s
        qy = (a * np.roll(qh, 1, axis=-2) + b * qh + a * np.roll(qh, -1, axis=-2)) / s
        az, bz, cz = self.filter_z_weights(a, b, interpflag)
        out = np.array(qy, copy=True)
        nzp = self.nzp
        sh = (nzp - 2,) + (1, 1)
        out[..., 1 : nzp - 1, :, :] = (az[1 : nzp - 1].reshape(sh) * qy[..., 0 : nzp - 2, :, :]
                                       + bz[1 : nzp - 1].reshape(sh) * qy[..., 1 : nzp - 1, :, :]
                                       + cz[1 : nzp - 1].reshape(sh) * qy[..., 2:nzp, :, :]) / s
        res = np.zeros(q.shape)
        res[..., :nx] = out
        return res

    # ---------------------------------------------------------------------- SGS
    def sgs_delta_static(self):
        """sgsterm.f90:95-103 (Q8: delta(nzp) is never set; the Van Driest factor is 0 there).

        ``nut_ceil0 = 1`` also zeroes the ceiling row's width (no eddy viscosity there)."""
        z = self.zpts
        nzp, nz = self.nzp, self.nz
        dx, dy = self.dx, self.dy
        delta = np.zeros(nzp)
        dz = abs(z[0] - z[1])
        delta[0] = (dx * dy * dz) ** (1.0 / 3.0)
        for k in range(2, nz + 1):
            dz = (z[k - 2] - z[k]) / 2.0
            delta[k - 1] = (dx * dy * dz) ** (1.0 / 3.0)
        dz = abs(z[nzp - 1] - z[nzp - 2])
        delta[0] = (dx * dy * dz) ** (1.0 / 3.0)
        delta[nzp - 1] = 0.0
        if self.cfg.nut_ceil0:
            delta[0] = 0.0
        return delta

    def static_smagorinsky(self, cs, duidxj):
        """sgsterm.f90:39-120.  duidxj[l][m] physical (nzp, ny, nxpp)."""
        c = self.cfg
        z = self.zpts
        nx, ny = self.nx, self.ny
        utau = np.zeros(duidxj.shape[2:][1:])
        wallcount = 0
        for k in range(self.nzp):
            if z[k] == 0.0 or abs(z[k]) == 0.5:
                utau = utau + c.xnu * duidxj[0, 2, k]
                wallcount += 1
        utau = utau / wallcount
        utau_avg = np.sum(utau[:, :nx]) / (nx * ny)
        utau_avg = math.sqrt(abs(utau_avg))
        zplus = z * utau_avg / c.xnu
        snorm = np.zeros(duidxj.shape[2:])
        for j in range(3):
            for i in range(3):
                snorm = snorm + (0.5 * (duidxj[i, j] + duidxj[j, i])) ** 2
        snorm = np.sqrt(2.0 * snorm)
        delta = self.sgs_delta_static()
        fac = -((cs * delta * (1.0 - np.exp(-zplus / 25.0))) ** 2)
        nu = fac[:, None, None] * snorm
        tau = np.empty_like(duidxj)
        for j in range(3):
            for i in range(3):
                tau[i, j] = nu * 0.5 * (duidxj[i, j] + duidxj[j, i])
        return tau

    def dynamic_smagorinsky(self, ui, uiuj, duidxj):
        """sgsterm.f90:123-280 (local 3-pt average of MijLij / MijMij, Q9).

        ``dyn_avg = 1``: <.> is the mean over each x-y plane (the nx physical points of
        every row) instead of the local filter of sgsterm.f90:243-244, so the coefficient
        is a function of z only (the plane-averaged dynamic procedure)."""
        c = self.cfg
        a, b = c.a, c.b
        z = self.zpts
        nzp, nx = self.nzp, self.nx
        dx, dy = self.dx, self.dy
        dz1 = np.zeros(nzp)
        dz1[0] = (dx * dy * (z[0] - z[1])) ** (1.0 / 3.0)
        for k in range(2, nzp):
            dzk = 0.5 * (z[k - 2] - z[k])
            dz1[k - 1] = (dx * dy * dzk) ** (1.0 / 3.0)
        dz1[nzp - 1] = (dx * dy * (z[nzp - 2] - z[nzp - 1])) ** (1.0 / 3.0)
        delta = dz1[:, None, None]
        cd = 2.0 * delta
        sij = np.empty_like(duidxj)
        snorm = np.zeros(duidxj.shape[2:])
        for m in range(3):
            for l in range(3):
                sij[l, m] = 0.5 * (duidxj[l, m] + duidxj[m, l])
                snorm = snorm + 2.0 * sij[l, m] * sij[l, m]
        snorm = np.sqrt(snorm)
        sijhat = self.filter_q_3pt(sij, a, b, 1)
        snormhat = np.zeros(duidxj.shape[2:])
        for m in range(3):
            for l in range(3):
                snormhat = snormhat + 2.0 * sijhat[l, m] * sijhat[l, m]
        snormhat = np.sqrt(snormhat)
        lij = np.empty_like(duidxj)
        for m in range(3):
            for l in range(3):
                lij[l, m] = delta * delta * snorm * sij[l, m]
                sijhat[l, m] = cd * cd * snormhat * sijhat[l, m]
        mij = self.filter_q_3pt(lij, a, b, 1)
        mij = 2.0 * (mij - sijhat)
        uihat = self.filter_q_3pt(ui, a, b, 1)
        lij = self.filter_q_3pt(uiuj, a, b, 1)
        for m in range(3):
            for l in range(3):
                lij[l, m] = lij[l, m] - uihat[l] * uihat[m]
        numc = np.zeros(duidxj.shape[2:])
        denomc = np.zeros(duidxj.shape[2:])
        for m in range(3):
            for l in range(3):
                numc = numc + mij[l, m] * lij[l, m]
                denomc = denomc + mij[l, m] * mij[l, m]
        if c.dyn_avg == 1:
            ctot = np.zeros_like(numc)
            cdd = np.zeros_like(denomc)
            ctot[..., :nx] = numc[..., :nx].mean(axis=(1, 2))[:, None, None]
            cdd[..., :nx] = denomc[..., :nx].mean(axis=(1, 2))[:, None, None]
        else:
            ctot = self.filter_q_3pt(numc, 1.0, 1.0, 1)
            cdd = self.filter_q_3pt(denomc, 1.0, 1.0, 1)
        with np.errstate(invalid="ignore", divide="ignore"):
            cc = ctot[..., :nx] / cdd[..., :nx]
        cc = np.where((cc < 0.0) | np.isnan(cc), 0.0, cc)
        cc = np.where(cc > 0.2 ** 2, 0.2 ** 2, cc)
        ctot[..., :nx] = cc
        nu = ctot * delta * delta * snorm
        tau = np.empty_like(duidxj)
        for m in range(3):
            for l in range(3):
                tau[l, m] = -2.0 * nu * sij[l, m]
        return tau

    def sigma_model(self, duidxj):
        """sgsterm.f90:283-423 (per-point invariants of G = g^T g, acos angles)."""
        nx, nzp, nz = self.nx, self.nzp, self.nz
        z = self.zpts
        g = duidxj[..., :nx]                                # g[l][m] = du_l/dx_m
        G = [[g[0, a] * g[0, b] + g[1, a] * g[1, b] + g[2, a] * g[2, b] for b in range(3)] for a in range(3)]
        G2 = [[G[a][0] * G[0][b] + G[a][1] * G[1][b] + G[a][2] * G[2][b] for b in range(3)] for a in range(3)]
        i1 = G[0][0] + G[1][1] + G[2][2]
        i2 = 0.5 * (i1 ** 2 - (G2[0][0] + G2[1][1] + G2[2][2]))
        i3 = (G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1])
              - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0])
              + G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]))
        a1 = i1 ** 2 / 9.0 - i2 / 3.0
        a2 = i1 ** 3 / 27.0 - i1 * i2 / 6.0 + 0.5 * i3
        a1 = np.where(a1 < 0.0, np.abs(a1), a1)
        with np.errstate(invalid="ignore", divide="ignore"):
            pre = a2 / np.sqrt(a1) ** 3.0
            a3 = np.where((pre >= 1.0) | np.isnan(pre), 0.0,
                          np.where(pre <= -1.0, PI / 3.0, np.arccos(np.clip(pre, -1.0, 1.0)) / 3.0))
        sa = np.sqrt(a1)
        s1 = np.sqrt(np.abs(i1 / 3.0 + 2.0 * sa * np.cos(a3)))
        s2 = np.sqrt(np.abs(i1 / 3.0 - 2.0 * sa * np.cos(PI / 3.0 + a3)))
        s3 = np.sqrt(np.abs(i1 / 3.0 - 2.0 * sa * np.cos(PI / 3.0 - a3)))
        ds = s3 * (s1 - s2) * (s2 - s3)
        with np.errstate(invalid="ignore", divide="ignore"):
            ds = np.where(s1 != 0.0, ds / np.where(s1 != 0.0, s1, 1.0) ** 2, ds)
        ds = np.where((ds < 0.0) | np.isnan(ds), 0.0, ds)
        dz = np.zeros(nzp)
        dz[0] = z[0] - z[1]
        dz[nzp - 1] = z[nz - 1] - z[nzp - 1]
        for k in range(1, nzp - 1):
            dz[k] = 0.5 * (z[k - 1] - z[k + 1])
        delta = (self.dx * self.dy * dz) ** (1.0 / 3.0)
        if self.cfg.nut_ceil0:
            delta[0] = 0.0
        nu = np.zeros(duidxj.shape[2:])
        nu[..., :nx] = -((1.35 * delta) ** 2)[:, None, None] * ds
        tau = np.empty_like(duidxj)
        for m in range(3):
            for l in range(3):
                tau[l, m] = 2.0 * nu * 0.5 * (duidxj[l, m] + duidxj[m, l])
        return tau

    def interscale_model(self, ui):
        """sgsterm.f90:426-467."""
        c = self.cfg
        ub = self.filter_q_3pt(ui, c.a, c.b, 1)
        ub2 = self.filter_q_3pt(ub, c.a, c.b, 1)
        uibuj = np.empty((3, 3) + ui.shape[1:])
        for m in range(3):
            for l in range(3):
                uibuj[l, m] = ub[l] * ui[m]
        ubb = self.filter_q_3pt(uibuj, c.a, c.b, 1)
        tau = np.empty_like(uibuj)
        for m in range(3):
            for l in range(3):
                tau[l, m] = ubb[l, m] - ub2[l] * ub2[m] - ub[l] * (ui[m] - ub[m])
        return tau

    def tau_sgs(self, ui, uiuj, duidxj):
        """sgsterm.f90:10-37."""
        c = self.cfg
        if c.sgsmodel == 0:
            return self.static_smagorinsky(c.cs, duidxj)
        if c.sgsmodel == 1:
            return self.dynamic_smagorinsky(ui, uiuj, duidxj)
        if c.sgsmodel == 2:
            return self.sigma_model(duidxj)
        if c.sgsmodel == 3:
            return self.interscale_model(ui)
        if c.sgsmodel == 4:   # TNS: no case; tauij stays 0 (navierstokes.f90:696)
            return np.zeros_like(uiuj)
        raise ValueError("sgsmodel")

    # --------------------------------------------------------------- dealiasing
    def dealias_sizes(self):
        """dealiasing.f90:133-154 at P = 1."""
        nx, ny = self.nx, self.ny
        if nx % 2 == 0 and nx % 4 == 0:
            nxl = 3 * nx // 2
        elif nx % 3 == 0:
            nxl = 5 * nx // 3
        elif nx % 5 == 0:
            nxl = 8 * nx // 5
        else:
            nxl = 2 * nx
        if ny % 2 == 0:
            nyl = 3 * ny // 2
        elif ny % 3 == 0:
            nyl = 5 * ny // 3
        elif ny % 5 == 0:
            nyl = 8 * ny // 5
        else:
            nyl = 2 * ny
        return nxl, nyl

    def dealiasing_2d_fftw(self, a, b, ref_compat=True):
        """dealiasing.f90:262-447.  ref_compat reproduces the reference's ky placement
        (ky slots copied to the first ny of ny_lg, Q1); otherwise negative ky go to the
        top of the padded spectrum (textbook 3/2 rule)."""
        abf = self.dealiasf_2d_fftw(self.horfft_fwd(a), self.horfft_fwd(b), ref_compat)
        return self.horfft_inv(abf)

    def dealiasf_2d_fftw(self, af, bf, ref_compat=True):
        """dealiasing.f90:449-618: the padded product of two spectral fields (nzp, ny, nx),
        spectral out (the middle of dealiasing_2d_fftw)."""
        nx, ny = self.nx, self.ny
        nxl, nyl = self.dealias_sizes()
        af = np.ascontiguousarray(af, dtype=np.float64).view(np.complex128)
        bf = np.ascontiguousarray(bf, dtype=np.float64).view(np.complex128)

        def pad(xf):
            big = np.zeros(xf.shape[:-2] + (nyl, nxl // 2 + 1), dtype=np.complex128)
            if ref_compat:
                big[..., :ny, : nx // 2] = xf
            else:
                h = ny // 2
                big[..., :h, : nx // 2] = xf[..., :h, :]
                big[..., nyl - (ny - h):, : nx // 2] = xf[..., h:, :]
            big = np.fft.ifft(big, axis=-2, norm="forward")
            return np.fft.irfft(big, n=nxl, axis=-1, norm="forward")

        prod = pad(af) * pad(bf)
        P = np.fft.rfft(prod, axis=-1)
        P = np.fft.fft(P, axis=-2)
        P = P * (1.0 / (float(nxl) * float(nyl)))
        out = np.zeros(P.shape[:-2] + (ny, nx // 2), dtype=np.complex128)
        if ref_compat:
            out[...] = P[..., :ny, : nx // 2]
        else:
            h = ny // 2
            out[..., :h, :] = P[..., :h, : nx // 2]
            out[..., h:, :] = P[..., nyl - (ny - h):, : nx // 2]
        return np.ascontiguousarray(out).view(np.float64)

    def test_fft(self, u):
        """horfft.f90:7-58: mean |u - q| over the nx physical points, q the inverse of the
        forward transform of u."""
        q = self.horfft_inv(self.horfft_fwd(u))
        n = self.nx * self.ny * self.nzp
        return float(np.sum(np.abs(u[..., : self.nx] - q[..., : self.nx]) / n))

    def zinteg(self, tmp, mapping, per, rev):
        """derivatives.f90:9-70 on tmp (nzp, n2, n1): per = 1 the interior quadrature
        weights wts(j) = 2/nz sin(t_j) sum_m sin(m t_j)(1 - cos(pi m))/m, t_j = pi j/nz,
        with the change of variable 0.5 xl/gf(k), summed from plane nz down; per = 0 the
        trapezoid rule on zpts from the top (rev = 0) or from plane 1 (rev = 1)."""
        nz, nzm, nzp = self.nz, self.nzm, self.nzp
        z = self.zpts
        tint = np.zeros(tmp.shape)
        if per == 1:
            pi = 4.0 * math.atan(1.0)
            wts = np.zeros(nzm + 1)
            for j in range(1, nzm + 1):
                tj = pi * float(j) / float(nz)
                acc = 0.0
                for m in range(1, nzm + 1):
                    acc = acc + math.sin(float(m) * tj) * (1.0 - math.cos(pi * float(m))) / float(m)
                wts[j] = 2.00 / float(nz) * math.sin(tj) * acc
            integ = np.zeros(tmp.shape[1:])
            for k in range(nz, 1, -1):               # 1-based k = nz..2
                integ = integ + wts[k - 1] * 0.5 * self.xl / self.gf[k - 1] * tmp[k - 1]
                tint[k - 1] = integ
            if mapping == 1:
                tint[0] = tint[1]
        elif rev == 0:
            integ = np.zeros(tmp.shape[1:])
            for k in range(nzp, 1, -1):              # 1-based k = nzp..2
                integ = integ + 0.5 * (tmp[k - 2] + tmp[k - 1]) * (z[k - 2] - z[k - 1])
                tint[k - 2] = integ
            if mapping == 1:
                tint[0] = tint[1]
        else:
            integ = np.zeros(tmp.shape[1:])
            for k in range(2, nzp + 1):
                integ = integ + 0.5 * (tmp[k - 2] + tmp[k - 1]) * (z[k - 2] - z[k - 1])
                tint[k - 1] = integ
        return tint

    # ---------------------------------------------------------------- boundary
    def bound(self, t):
        """navierstokes.f90:214-355 -> spectral bc field (nzp, ny, nx).
        bctype 1: Spalart ceiling suction (+ fringe blowing rescaled for mass balance);
        bctype 3: steady ceiling profile V(x) = samp*(-G(0.4 Lx) + G(0.7 Lx)) with
        Gaussian width 0.05 Lx, no-slip wall (the workloads' driver-side patch)."""
        c = self.cfg
        nx, nzp = self.nx, self.nzp
        bc = np.zeros((nzp, self.ny, self.nxpp))
        x = self.xpts
        if c.bctype == 1:
            pi = PI
            h = self.zpts[0]
            xloc = 1.0e5 * c.xnu / c.u0
            if x[nx // 2 - 1] < xloc:
                xloc = c.x0 + 3.0 / 7.0 * c.xlen
            xwidth = 0.24 * h
            amp = c.samp * c.u0 / (0.24 * math.sqrt(pi))
            bamp = c.samp / 2.25
            ref_time = c.xlen / c.u0
            if t - c.istartbound <= ref_time:
                s = math.sin(0.5 * pi * (t - c.istartbound) / ref_time) ** 2.0
                amp, bamp = amp * s, bamp * s
            bwidth = xwidth * 2.25
            x1 = x[nx - 1] - 2.0 * bwidth
            eta = (x - xloc) / xwidth
            eta0 = (x - x1) / bwidth
            bc[0, :, :nx] = (amp * np.exp(-eta ** 2.0))[None, :]
            bc[nzp - 1, :, :nx] = (bamp * np.exp(-eta0 ** 4.0))[None, :]
        elif c.bctype == 3:
            L = c.xlen
            xr = x - c.x0
            g1 = np.exp(-(((xr - 0.4 * L) / (0.05 * L)) ** 2))
            g2 = np.exp(-(((xr - 0.7 * L) / (0.05 * L)) ** 2))
            bc[0, :, :nx] = (c.samp * (g2 - g1))[None, :]
        bcf = self.horfft_fwd(bc)
        if c.bctype == 1:
            ampf = bcf[0, 0, 0]
            bampf = bcf[nzp - 1, 0, 0]
            bfac = ampf / bampf if bampf != 0 else 0.0
            bcf[nzp - 1] = bfac * bcf[nzp - 1]
        return bcf

    def ceiling_weight(self):
        """Inflow weight m(x) of the bctype-3 ceiling: the blowing Gaussian G(0.4 Lx)."""
        c = self.cfg
        L = c.xlen
        xr = self.xpts - c.x0
        return np.exp(-(((xr - 0.4 * L) / (0.05 * L)) ** 2))

    def ceiling_blend(self, free, dirichlet):
        """Spectral ceiling plane (1 - m(x)) free + m(x) dirichlet, the product in physical space."""
        m = np.zeros(self.nxpp)
        m[: self.nx] = self.ceiling_weight()
        a = self.horfft_inv(free[None])
        b = self.horfft_inv(dirichlet[None])
        return self.horfft_fwd((1.0 - m) * a + m * b)[0]

    # ------------------------------------------------------------ time coeffs
    @staticmethod
    def abcoef(nstart, istep):
        """navierstokes.f90:1500-1542 (EX3 3,-3,1 with Euler/SS2 start, Q13)."""
        if nstart == 0:
            if istep == 1:
                return np.array([1.0, 0.0, 0.0])
            if istep == 2:
                return np.array([2.0, -1.0, 0.0])
        return np.array([3.0, -3.0, 1.0])

    @staticmethod
    def bdfcoef(nstart, istep, dt1, dt2, dt3):
        """navierstokes.f90:1545-1590 + bdfvar* :1623-1662."""
        def var():
            dt12 = dt1 + dt2
            dt23 = dt2 + dt3
            dt123 = dt1 + dt2 + dt3
            v = 1.0 + dt1 / dt12 + dt1 / dt123
            b1 = -(-1.0 - dt1 / dt2 - dt1 / dt123 - dt1 * dt1 / (dt123 * dt2) - dt1 * dt1 * dt12 / (dt123 * dt23 * dt2))
            b2 = -(dt1 * dt1 / (dt12 * dt2) + dt1 * dt1 / (dt2 * dt123) + dt1 * dt1 * dt12 / (dt2 * dt23 * dt123)
                   + dt1 * dt1 * dt12 / (dt3 * dt23 * dt123))
            b3 = -(-dt1 * dt1 * dt12 / (dt3 * dt23 * dt123))
            return np.array([b1, b2, b3]), 1.0 / v
        if nstart == 0:
            if istep == 1:
                return np.array([1.0, 0.0, 0.0]), 1.0
            if istep == 2:
                return np.array([2.0, -0.5, 0.0]), 2.0 / 3.0
            if dt1 != dt2 or dt2 != dt3:
                return var()
            return np.array([3.0, -1.5, 1.0 / 3.0]), 6.0 / 11.0
        if dt1 != dt2 or dt2 != dt3:
            return var()
        return np.array([3.0, -1.5, 1.0 / 3.0]), 6.0 / 11.0

    # ------------------------------------------------------------------- state
    def new_state(self, uf, vf, wf, tempf=None):
        """Core arrays after setup_ic (init.f90:473-486): histories and pressure zero."""
        sh = uf.shape
        z = lambda: np.zeros(sh)  # noqa: E731
        st = dict(uf=np.array(uf, copy=True), vf=np.array(vf, copy=True), wf=np.array(wf, copy=True),
                  tempf=z() if tempf is None else np.array(tempf, copy=True),
                  pf=z(), pnf=z(), tnf=z(), tnmf=z(), nltnf=z(), nltnmf=z(),
                  unif=np.zeros((3,) + sh), unmif=np.zeros((3,) + sh),
                  nlnif=np.zeros((3,) + sh), nlnmif=np.zeros((3,) + sh),
                  istep=0, t=0.0, dt=self.cfg.dt, dt2=self.cfg.dt, dt3=self.cfg.dt,
                  abfac=np.zeros(3), bdf=np.zeros(3), bdfv=0.0)
        return st

    # ----------------------------------------------------------- nonlinear term
    def nlterm_skewsym(self, uif, tempf, sgsflag):
        """navierstokes.f90:658-855.  uif (3, nzp, ny, nx) spectral -> (nltermif, nltempf)."""
        c = self.cfg
        nzp = self.nzp
        duidxjf = np.empty((3, 3) + uif.shape[1:])
        for l in range(3):
            duidxjf[l, 0] = self.dfdx(uif[l])
            duidxjf[l, 1] = self.dfdy(uif[l])
            duidxjf[l, 2] = self.dfdz(uif[l])
        duidxj = self.horfft_inv(duidxjf)
        ui = self.horfft_inv(uif)
        ubk = self.ubar.reshape(nzp, 1, 1)
        ui[0] = ui[0] + ubk
        duidxj[0, 2] = duidxj[0, 2] + self.ubarp.reshape(nzp, 1, 1)
        nlt = np.zeros(ui.shape)
        uiuj = np.empty((3, 3) + ui.shape[1:])
        if c.dealiasflag == 1:
            rc = bool(c.dealias_ref_compat)
            for l in range(3):
                for m in range(3):
                    nlt[l] = nlt[l] - 0.5 * self.dealiasing_2d_fftw(ui[m], duidxj[l, m], rc)
                    uiuj[m, l] = self.dealiasing_2d_fftw(ui[m], ui[l], rc)
        else:
            for l in range(3):
                for m in range(3):
                    nlt[l] = nlt[l] - 0.5 * ui[m] * duidxj[l, m]
                    uiuj[m, l] = ui[m] * ui[l]
        if c.forcing == 1:  # navierstokes.f90:750-756: nltermi += F_i
            lam = np.zeros(self.nxpp)          # pads carry no forcing
            lam[: self.nx] = self.sponge_lambda()
            lam = lam[None, None, :]
            nlt[0] = nlt[0] + lam * (ubk - ui[0])
            nlt[1] = nlt[1] + lam * (0.0 - ui[1])
            nlt[2] = nlt[2] + lam * (0.0 - ui[2])
        tau = np.zeros_like(uiuj)
        if sgsflag == 1:
            tau = self.tau_sgs(ui, uiuj, duidxj)
        h = 0.5 * uiuj + tau
        nltermif = self.horfft_fwd(nlt)
        hijf = self.horfft_fwd(h)
        for l in range(3):
            nltermif[l] = nltermif[l] - self.dfdx(hijf[l, 0])
            nltermif[l] = nltermif[l] - self.dfdy(hijf[l, 1])
            nltermif[l] = nltermif[l] - self.dfdz(hijf[l, 2])
        # temperature (navierstokes.f90:800-845); exact zero when T == 0 and tbar == 0
        if np.any(np.abs(tempf) > 0.0) or np.any(self.tbar != 0.0):
            dtdxif = np.stack([self.dfdx(tempf), self.dfdy(tempf), self.dfdz(tempf)])
            if np.any(np.abs(tempf) > 0.0):
                dtdxi = self.horfft_inv(dtdxif)
                temp = self.horfft_inv(tempf)
            else:
                dtdxi = np.zeros(ui.shape)
                temp = np.zeros(ui.shape[1:])
            temp = temp + self.tbar.reshape(nzp, 1, 1)
            dtdxi[2] = dtdxi[2] + np.sum(self.tbarp)
            nltemp = np.zeros(temp.shape)
            uitemp = np.empty(ui.shape)
            for l in range(3):
                nltemp = nltemp - 0.5 * ui[l] * dtdxi[l]
                uitemp[l] = ui[l] * temp
            nltempf = self.horfft_fwd(nltemp)
            hf = self.horfft_fwd(uitemp)
            nltempf = nltempf - 0.5 * self.dfdx(hf[0])
            nltempf = nltempf - 0.5 * self.dfdy(hf[1])
            nltempf = nltempf - 0.5 * self.dfdz(hf[2])
        else:
            nltempf = np.zeros(tempf.shape)
        return nltermif, nltempf

    # -------------------------------------------------------- step (a20 rows)
    def navierstokes_update(self, st):
        """navierstokes.f90:1157-1496 (bctype > 0 path uses solve_nhom_helmholtz)."""
        c = self.cfg
        nzp = self.nzp
        uf, vf, wf = st["uf"], st["vf"], st["wf"]
        istep = st["istep"]
        dt, dt2, dt3 = st["dt"], st["dt2"], st["dt3"]
        uif = np.stack([uf, vf, wf])
        tempsf = np.array(st["tempf"], copy=True)
        uif_nl = uif
        if c.dealiasflag >= 2:
            uif_nl = self.filterqspec(uif)
            if np.any(np.abs(tempsf) > 1.0e-10):
                tempsf = self.filterqspec(tempsf)
        nltermif, nltempf = self.nlterm_skewsym(uif_nl, tempsf, c.sgsflag)
        if c.dealiasflag != 2:
            uif = uif_nl
        tmp = nltermif.copy()
        tempsf = nltempf.copy()
        if istep <= 3 or dt != dt2 or dt2 != dt3:
            st["abfac"] = self.abcoef(c.nstart, istep)
        ab = st["abfac"]
        nltermif = ab[0] * nltermif + ab[1] * st["nlnif"] + ab[2] * st["nlnmif"]
        nltempf = ab[0] * nltempf + ab[1] * st["nltnf"] + ab[2] * st["nltnmf"]
        if istep <= 3 or dt != dt2 or dt2 != dt3:
            st["bdf"], st["bdfv"] = self.bdfcoef(c.nstart, istep, dt, dt2, dt3)
        bdf, bdfv = st["bdf"], st["bdfv"]
        # temperature (navierstokes.f90:1251-1275)
        xkap = c.xnu / c.prtl
        dtkapi = 1.0 / (xkap * dt)
        rhs_t = -dtkapi * (dt * nltempf + bdf[0] * st["tempf"] + bdf[1] * st["tnf"] + bdf[2] * st["tnmf"])
        if np.sum(np.abs(self.tbarpp)) > 0.0:
            du0 = np.zeros((nzp, self.ny, self.nxpp))
            du0[:, :, : self.nx] = self.tbarpp.reshape(nzp, 1, 1)
            rhs_t = rhs_t - self.horfft_fwd(du0)
        rhs_t = self.solve_helmholtz(rhs_t, dtkapi / bdfv)
        st["nltnmf"] = st["nltnf"]
        st["tnmf"] = st["tnf"]
        st["nltnf"] = tempsf
        st["tnf"] = st["tempf"]
        st["tempf"] = rhs_t
        tempf = rhs_t
        # velocity rhs (navierstokes.f90:1282-1303)
        dtnui = 1.0 / (c.xnu * dt)
        rhsf = -dtnui * (dt * nltermif + bdf[0] * uif + bdf[1] * st["unif"] + bdf[2] * st["unmif"])
        if np.sum(np.abs(self.ubarpp)) > 0.0:
            du0 = np.zeros((nzp, self.ny, self.nxpp))
            du0[:, :, : self.nx] = self.ubarpp.reshape(nzp, 1, 1)
            rhsf[0] = rhsf[0] - self.horfft_fwd(du0)
        if np.any(np.abs(tempf) > 1.0e-10):
            rhsf[0] = rhsf[0] - c.coex * c.gx * tempf / c.xnu
            rhsf[1] = rhsf[1] - c.coex * c.gy * tempf / c.xnu
            rhsf[2] = rhsf[2] - c.coex * c.gz * tempf / c.xnu
        st["nlnmif"] = st["nlnif"]
        st["unmif"] = st["unif"]
        st["nlnif"] = tmp
        st["unif"] = uif.copy()
        uif = uif.copy()
        # BCs (navierstokes.f90:1321-1373)
        if istep == 1:
            pfac = (0.0, 0.0)
        elif istep == 2:
            pfac = (1.0, 0.0)
        else:
            pfac = (2.0, -1.0)
        psf = pfac[0] * st["pf"] + pfac[1] * st["pnf"]
        st["pnf"] = st["pf"]
        dpf = self.dfdx(psf)
        uif[0, nzp - 1] = dt * bdfv * dpf[nzp - 1]
        wav2 = self.wavxx[None, :] + self.wavyy[:, None]
        if c.top_bc == 1:
            # Dirichlet ceiling: the wall's Zang-Hussaini value mirrored at the top
            uif[0, 0] = dt * bdfv * dpf[0]
        else:
            uif[0, 0] = -rhsf[0, 0] / (wav2 + dtnui / bdfv)
            uif[1, 0] = -rhsf[1, 0] / (wav2 + dtnui / bdfv)
        if c.top_bc == 2:
            dpy = self.dfdy(psf)
            uif[0, 0] = self.ceiling_blend(uif[0, 0], dt * bdfv * dpf[0])
            uif[1, 0] = self.ceiling_blend(uif[1, 0], dt * bdfv * dpy[0])
        dwdzf = self.d2fdz2(uif[2])
        nhom = (c.bctype > 0 and st["t"] >= c.istartbound) or c.forcing == 1
        if c.bctype > 0 and st["t"] >= c.istartbound:
            uif[2] = self.bound(st["t"])
            uif[2, 0] = uif[2, 0] + c.xnu * dt * bdfv * dwdzf[0]
        else:
            uif[2, 0] = c.xnu * dt * bdfv * dwdzf[0]
            uif[2, nzp - 1] = c.xnu * dt * bdfv * dwdzf[nzp - 1]
        dpf = self.dfdy(psf)
        uif[1, nzp - 1] = dt * bdfv * dpf[nzp - 1]
        if c.top_bc == 1:
            uif[1, 0] = dt * bdfv * dpf[0]
        for i in range(3):
            if nhom:
                rhsf[i] = self.solve_nhom_helmholtz(rhsf[i], uif[i], dtnui / bdfv)
            else:
                rhsf[i] = self.solve_helmholtz(rhsf[i], dtnui / bdfv)
        uf = rhsf[0].copy()
        vf = rhsf[1].copy()
        wf = rhsf[2].copy()
        # pressure step (navierstokes.f90:1411-1488)
        dpf = self.dfdz(uf)
        dwdzf = self.dfdz(vf)
        divdzf = self.divuv(dpf, dwdzf)
        divf = self.divuv(uf, vf)
        rhs3 = -divdzf - wav2[None] * wf
        if c.bctype > 0 and st["t"] >= c.istartbound:
            wf = self.bound(st["t"])
        else:
            wf[0] = 0.0
            wf[nzp - 1] = 0.0
        wf = self.solve_nhom_helmholtz(rhs3, wf, 0.0)
        dwdzf = self.dfdz(wf)
        with np.errstate(invalid="ignore", divide="ignore"):
            pf = np.where(wav2[None] == 0.0, 0.0, -(dwdzf + divf) / (wav2[None] * dt * bdfv))
        dpf = self.dfdx(pf)
        uf = uf - dt * bdfv * dpf
        dpf = self.dfdy(pf)
        vf = vf - dt * bdfv * dpf
        st["uf"], st["vf"], st["wf"], st["pf"] = uf, vf, wf, pf

    def timestep(self, st):
        """navierstokes.f90:10-90 with dtadapt = 0 (the per-step part; the filtering of
        ifilt >= 1 every modfilt steps before the update)."""
        c = self.cfg
        st["istep"] += 1
        if c.ifilt >= 1 and st["istep"] % c.modfilt == 0:   # navierstokes.f90:31-39
            if c.sgsflag == 1 and c.sgsmodel == 4:
                r = self.tns_automatic_filtering(st["uf"], st["vf"], st["wf"], st["tempf"], st["pf"])
            else:
                r = self.filterall3pt(st["uf"], st["vf"], st["wf"], st["tempf"], st["pf"])
            st["uf"], st["vf"], st["wf"], st["tempf"], st["pf"] = r[:5]
        self.navierstokes_update(st)
        self.norm(st["uf"], st["vf"], st["wf"], st["tempf"])
        if c.ifrz == 1 and self.wavx[0] == 0:
            for k in ("uf", "vf", "wf"):
                st[k][:, 0, 0:2] = 0.0
        st["t"] = st["t"] + st["dt"]
        st["dt3"] = st["dt2"]
        st["dt2"] = st["dt"]
        return st

    AVG_FIELDS = ("uf", "vf", "wf", "tempf", "pf")

    def time_average(self, st, avg, istartavg, iomod, its=None):
        """main.f90:69-76, after timestep(): q_tavg = q_tavg + q/iomod for uf, vf, wf, tempf,
        pf once its (the total step count, navierstokes.f90:63; the run's istep unless a
        restart set it) exceeds istartavg.  avg: dict of the five averages, updated."""
        its = st["istep"] if its is None else its
        if its > istartavg:
            for k in self.AVG_FIELDS:
                avg[k] = avg[k] + st[k] / iomod
        return avg

    def divcheck(self, uf, vf, wf):
        """derivatives.f90:348-392: mean |div u| in physical space (pads are 0 here)."""
        divf = self.dfdx(uf) + self.dfdy(vf) + self.dfdz(wf)
        divp = self.horfft_inv(divf)
        return float(np.sum(np.abs(divp)) / (self.nx * self.ny * self.nzp))

    # ----------------------------------------------------- TNS / mode-cut filters
    MODECUT_COEF = (6.0, -15.0, 20.0, -15.0, 6.0, -1.0)

    def modecut_fc(self, q, a, b):
        """filters.f90:617-654 modecutFC: qb = sum_l coef_l F^l q with F the 3-pt filter
        (filter_q_3pt with interpflag 1) applied l times, coef = (6, -15, 20, -15, 6, -1)."""
        qb = np.zeros(q.shape)
        f = q
        for cl in self.MODECUT_COEF:
            f = self.filter_q_3pt(f, a, b, 1)
            qb = qb + cl * f
        return qb

    def filterall3pt(self, uf, vf, wf, tempf, pf):
        """filters.f90:9-49: physical u, v, w, p (temp when any |T| > 1e-10), modecutFC with the
        SGS test-filter pair (a, b), back to spectral space.  Returns (uf, vf, wf, tempf, pf)."""
        c = self.cfg
        hot = tempf is not None and bool(np.any(np.abs(tempf) > 1.0e-10))
        fields = [uf, vf, wf, pf] + ([tempf] if hot else [])
        q = np.stack([self.horfft_inv(f) for f in fields])
        qb = self.modecut_fc(q, c.a, c.b)
        out = [self.horfft_fwd(qb[l]) for l in range(len(fields))]
        return out[0], out[1], out[2], (out[4] if hot else tempf), out[3]

    def tns_automatic_filtering(self, uf, vf, wf, tempf, pf):
        """filters.f90:52-161: modecutFC of physical u, v, w, p with (1, 6) and with (1, 2);
        per plane the kinetic energy removed by each, summed over x and y over nx*ny (the
        nx physical points of each row; the reference's sum also covers the two pad reals,
        quirk Q10); trapezoidal z integrals I(h) (the (1, 6) filter) and I(delta) (the
        (1, 2) filter); when I(h)/I(delta) > 0.007 the fields take the (1, 6) result.
        Returns (uf, vf, wf, tempf, pf, Ih, Idelta, applied)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        q = np.stack([self.horfft_inv(f) for f in (uf, vf, wf, pf)])
        q2kc = self.modecut_fc(q, 1.0, 6.0)
        qf = self.modecut_fc(q, 1.0, 2.0)
        e_delta = np.zeros(self.nzp)
        e_h = np.zeros(self.nzp)
        for k in range(self.nzp):
            for l in range(3):
                e_delta[k] += np.sum((q[l, k, :, :nx] - qf[l, k, :, :nx]) ** 2) / (ny * nx)
                e_h[k] += np.sum((q[l, k, :, :nx] - q2kc[l, k, :, :nx]) ** 2) / (ny * nx)
        Ih = 0.0
        Idelta = 0.0
        for k in range(nz - 1, -1, -1):                 # Fortran k = nz .. 1
            dz = self.zpts[k] - self.zpts[k + 1]
            Ih = Ih + 0.5 * (e_h[k] + e_h[k + 1]) * dz
            Idelta = Idelta + 0.5 * (e_delta[k] + e_delta[k + 1]) * dz
        applied = Ih / Idelta > 0.007
        if applied:
            uf, vf, wf, pf = (self.horfft_fwd(q2kc[l]) for l in range(4))
        return uf, vf, wf, tempf, pf, Ih, Idelta, applied

    # ------------------------------------------------- Orr-Sommerfeld mode
    def orrsomm_matrices(self, alphax, betax):
        """orrsomm.f90:159-213 (iglb <= 0): the n = nzp + 2 complex pair (A, B) of the
        Orr-Sommerfeld problem on the Chebyshev grid, with the membrane-bottom-wall rows
        (c01..c04 = 1e-40) and the boundary rows 1, 2, 3, n-1, n."""
        nzp = self.nzp
        n = nzp + 2
        c01 = c02 = c03 = c04 = 1.0e-40
        rey = 1.0 / self.cfg.xnu
        wav2 = alphax * alphax + betax * betax
        d, d2 = self.d, self.d2
        d4 = d2 @ d2
        I = np.eye(nzp)
        a = np.zeros((n, n), dtype=np.complex128)
        b = np.zeros((n, n), dtype=np.complex128)
        ub = self.ubar[:, None]
        a[:nzp, :nzp] = (d4 - 2.0 * wav2 * d2 + wav2 * wav2 * I
                         - 1j * alphax * rey * (ub * (d2 - wav2 * I) - self.ubarpp[:, None] * I))
        b[:nzp, :nzp] = -1j * rey * (d2 - wav2 * I)
        row = np.zeros(nzp, dtype=np.complex128)
        for j in range(nzp):
            acc = 0.0 + 0.0j
            for k in range(nzp):
                acc = acc - 1j * c01 / (alphax * rey) * d2[nzp - 1, k] * d[k, j]
            row[j] = acc
        a[n - 1, :] = 0.0
        a[n - 1, :nzp] = row
        for r in (0, 1, 2, n - 2, n - 1):
            b[r, :] = 0.0
        for r in (0, 2, n - 2):
            a[r, :] = 0.0
        a[1, :nzp] = d[nzp - 1, :]
        a[0, 0] = 1.0
        a[2, nzp - 1] = 1.0
        b[2, n - 2] = 1.0 / alphax
        a[1, n - 2] = self.ubarp[nzp - 1]
        a[1, n - 1] = 0.0
        a[n - 2, n - 1] = 1.0
        b[n - 2, n - 2] = -1j
        a[n - 1, n - 2] = c03 + c04 * wav2 - 1j * wav2 * (c02 / rey) * (self.ubarp[nzp - 1] / alphax)
        a[n - 1, n - 1] = c01
        b[n - 1, n - 1] = 1j
        return a, b

    def solve_orrsomm(self, eps, alphax, betax, omega, iglb=-1):
        """orrsomm.f90:110-562: (uvec, vvec, wvec, omega).  iglb < 0 first takes the most
        unstable eigenvalue with 0 < c_r < 0.42 of the whole generalized spectrum (zggevx there,
        scipy's QZ here; eigenvalues phi/|gam|, :238-244, 255); then the two-sided inverse
        iteration with Rayleigh-quotient updates (:328-470; outer and inner loops of 10);
        streamfunction -> (u, v, w) = (D phi, 0, -i alpha phi), scaled to energy eps.  The
        reference's random nudges of a stalled iteration (:383-396) are left out, and where it
        restarts from a random omega because a converged Rayleigh quotient still leaves a
        residual above 1.5e-6 (:447-453) the restatement keeps the mode."""
        import scipy.linalg as sla
        a, b = self.orrsomm_matrices(alphax, betax)
        n = a.shape[0]
        if iglb < 0:
            phi, gam = sla.eig(a, b, homogeneous_eigvals=True)[0]
            mag = np.abs(gam)
            reig = np.where(mag > 0, phi.real / np.where(mag > 0, mag, 1.0), phi.real)
            aeig = np.where(mag > 0, phi.imag / np.where(mag > 0, mag, 1.0), phi.imag)
            ok = (reig / alphax > 0) & (reig / alphax < 0.42)
            j = int(np.argmax(np.where(ok, aeig / alphax, -np.inf)))
            omega = complex(reig[j], aeig[j])
        omk = complex(omega)
        ph = np.ones(n, dtype=np.complex128)
        ga = np.ones(n, dtype=np.complex128)
        bh = b.conj().T
        l1 = lambda x: float(np.sum(np.abs(x)))  # noqa: E731

        phk1, gak1 = ph, ga
        for _ in range(10):
            ck = a - omk * b
            rk = l1(ck @ ph)
            lu = sla.lu_factor(ck)
            for ii in range(1, 11):
                inner_rk = l1(ck @ ph)
                rphi = sla.lu_solve(lu, b @ ph, trans=0)
                rgam = sla.lu_solve(lu, bh @ ga, trans=2)
                phk1 = rphi / l1(rphi)
                gak1 = rgam / l1(rgam)
                inner_rk1 = l1(ck @ phk1)
                if l1(phk1 - ph) < 5.0e-5 or inner_rk1 < 5.0e-6:
                    if inner_rk1 > inner_rk:
                        phk1, gak1 = ph, ga
                    ph, ga = phk1, gak1
                    break
                # (a residual that grew here makes the reference add 1e-3 random noise to the
                # iterates, :383-396; the restatement keeps iterating without it)
                ph, ga = phk1, gak1
            omk1 = np.vdot(gak1, a @ phk1) / np.vdot(gak1, b @ phk1)
            rk1 = l1((a - omk1 * b) @ phk1)
            if abs(omk1 - omk) < 1.0e-14 and rk1 < 1.5e-6:
                if rk1 >= rk:
                    omk1, phk1, gak1 = omk, ph, ga
                omk = omk1
                break
            if abs(omk1 - omk) < 1.0e-14:
                # stationary Rayleigh quotient with a residual above 1.5e-6 (a test near the
                # rounding floor of (A - omega B) phi): the reference restarts from a random
                # omega (:447-453); the restatement keeps the mode
                omk = omk1
                break
            if rk1 >= rk:
                omk1, phk1, gak1 = omk, ph, ga
            omk = omk1
        nzp = self.nzp
        w = phk1[:nzp].copy()
        u = self.d @ w
        w = -1j * alphax * w
        v = np.zeros(nzp, dtype=np.complex128)

        def energy(q):
            e = 0.0
            for k in range(1, nzp - 1):          # Fortran k = 2 .. nzp-1
                e = e - 0.5 * (abs(q[k + 1]) * abs(q[k + 1]) + abs(q[k]) * abs(q[k])) * (self.zpts[k + 1] - self.zpts[k])
            return e * 2.0 / 7.0
        etot = energy(u) + energy(v) + energy(w)
        xn = math.sqrt(eps / etot)
        return u * xn, v * xn, w * xn, omk

    # ------------------------------------------------- linear-stability energy
    def ener(self, uf, vf, wf, itot):
        """orrsomm.f90:625-666: ee(k) = sum over the nx*ny physical points of
        0.5 (u^2 + v^2 + w^2) (u + ubar when itot = 1); en = sum_{k=nz..k0} 0.5 (ee(k+1) +
        ee(k)) (zpts(k) - zpts(k+1)) with k0 = 2, i.e. the top interval is left out."""
        nx = self.nx
        u = self.horfft_inv(uf)[:, :, :nx]
        v = self.horfft_inv(vf)[:, :, :nx]
        w = self.horfft_inv(wf)[:, :, :nx]
        if itot == 1:
            u = u + self.ubar[:, None, None]
        ee = np.sum(0.5 * (u * u + v * v + w * w), axis=(1, 2))
        en = 0.0
        for k in range(self.nz - 1, 0, -1):              # Fortran k = nz .. 2
            en = en + 0.5 * (ee[k + 1] + ee[k]) * (self.zpts[k] - self.zpts[k + 1])
        return float(en)

    def lintest(self, uf, vf, wf, istep, dt, enr1, enr2, grow1, grow2):
        """orrsomm.f90:566-622 for step istep (1-based): enr1 = ener(total), enr2 = ener with
        the (kx, ky) = (0, 0) mode removed (no ubar); grow = (E(istep) - E(istep-1)) /
        (2 E(istep) dt), 0 at istep 1.  Fills the arrays in place."""
        i = istep - 1
        enr1[i] = self.ener(uf, vf, wf, 1)
        up, vp, wp = (np.array(f, copy=True) for f in (uf, vf, wf))
        for f in (up, vp, wp):
            f[:, 0, 0:2] = 0.0
        enr2[i] = self.ener(up, vp, wp, 0)
        if istep > 1:
            grow1[i] = 1.0 / (2.0 * enr1[i]) * (enr1[i] - enr1[i - 1]) / dt
            grow2[i] = 1.0 / (2.0 * enr2[i]) * (enr2[i] - enr2[i - 1]) / dt
        else:
            grow1[i] = 0.0
            grow2[i] = 0.0

    # ------------------------------------------------------------ sponge layer
    def sponge_lambda(self):
        """navierstokes.f90:119-187 force(): the Rayleigh-damping coefficient lambda(x) of the
        fringe (ftype = 1, Herbst & Henningson), nx values.  kappa = 2 u0/dxf (-ln 1e-5);
        r(i) = kappa (s0 + s) with s0 = exp(-eta0^2) rising over dxs (1 on the plateau,
        0 once eta >= 0) and s = exp(-eta^2) decaying over the left part dxl; then shifted
        right by xshift = the last i <= nx/4 with xpts(i) < xpts(1) + dxl/1.5."""
        c = self.cfg
        nx, xp = self.nx, self.xpts
        epsx = 1.0e-5
        sfs = c.sfsize * c.fsize
        lfs = c.lfsize * c.fsize
        rfsize = c.fsize - sfs - lfs
        dxs = sfs * c.xlen
        dxl = lfs * c.xlen
        dxr = rfsize * c.xlen
        dxf = dxs + dxl + dxr
        kappa = 2.0 * c.u0 / dxf * (-math.log(epsx))
        r = np.zeros(nx)
        for i in range(nx):
            eta0 = (xp[i] - (xp[nx - 1] - dxf + dxs)) / (dxs / 2.0)
            eta = (xp[i] - (xp[nx - 1] - dxl)) / (dxl / 2.5)
            s0 = math.exp(-(eta0 * eta0))    # Fortran x**2 with a constant exponent is x*x
            s = math.exp(-(eta * eta))
            if eta0 > 0.0 and eta < 0.0:
                s0 = 1.0
            if eta < 0.0:
                s = 0.0
            if eta >= 0.0:
                s0 = 0.0
            r[i] = kappa * (s0 + s)
        xshift = 0
        for i in range(1, nx // 4 + 1):          # Fortran i = 1..nx/4
            if xp[i - 1] < xp[0] + dxl / 1.5:
                xshift = i
        lam = np.zeros(nx)
        if xshift > 1:
            lam[xshift:] = r[: nx - xshift]
            lam[:xshift] = r[nx - xshift:]
        else:
            lam[:] = r
        return lam

    # ------------------------------------------------- per-step diagnostics
    def nancheck(self, *fields):
        """navierstokes.f90:77-81 / main.f90:79: any(isnan(.)) over spectral fields (Inf passes)."""
        return any(bool(np.isnan(f).any()) for f in fields if f is not None)

    def maxcalc(self, uf, vf, wf, tempf, dtmax, dt, dt2, dt3, t, fsize=0.0):
        """navierstokes.f90:358-511 maxcalc (one rank; the allreduces are identities).

        Returns (stats, dt, dt2, dt3, change_dt); stats = [umax, umin, vmax, vmin, wmax, wmin,
        tmax, tmin, cflx, cfly, cflz] with u = u + ubar(k), temp = temp + tbar(k) (:416-419).
        The extrema run over the nx physical points of each x-row; the reference's maxval also
        reads the two pad reals of each row, which hold FFT junk there (quirk Q10).
        z-CFL (:428-437): (w dt)/(zpts(k) - zpts(k+1)) on planes 1..nz, then plane 1 a second
        time with the same factor, plane nzp with (zpts(nz) - zpts(nzp)); over x < findex =
        int((1 - 0.8 fsize) nx) once t > istartbound (:442-450)."""
        c = self.cfg
        nx, nz, nzp = self.nx, self.nz, self.nzp
        zp = self.zpts
        u = self.horfft_inv(uf)[:, :, :nx] + self.ubar[:, None, None]
        v = self.horfft_inv(vf)[:, :, :nx]
        w = self.horfft_inv(wf)[:, :, :nx]
        tp = np.zeros_like(u) if tempf is None else self.horfft_inv(tempf)[:, :, :nx]
        tp = tp + self.tbar[:, None, None]
        vmax, vmin, wmax, wmin = v.max(), v.min(), w.max(), w.min()
        umax, umin, tmax, tmin = u.max(), u.min(), tp.max(), tp.min()
        ws = np.empty_like(w)
        for k in range(nz):
            ws[k] = (w[k] * dt) / (zp[k] - zp[k + 1])
        ws[0] = (ws[0] * dt) / (zp[0] - zp[1])
        ws[nz] = (w[nz] * dt) / (zp[nz - 1] - zp[nz])
        cflx = (np.abs(u).max() * dt) / self.dx
        cfly = (np.abs(v).max() * dt) / self.dy
        findex = int((1.0 - 0.8 * fsize) * nx) if t > c.istartbound else nx
        # maxval of an empty section is -huge (findex = 0)
        cflz = np.abs(ws[:, :, :findex]).max() if findex > 0 else -np.finfo(np.float64).max
        exceed = cflz > 0.45 or cflx > 0.25 or cfly > 0.25
        below = cflz < 0.25 and cflx < 0.10 and cfly < 0.12 and dt < dtmax
        change = exceed or below
        if change:
            cdt = 1.25 if exceed else 1.0 / 1.25
            dt2 = dt
            dt = dt / cdt
        stats = np.array([umax, umin, vmax, vmin, wmax, wmin, tmax, tmin, cflx, cfly, cflz])
        return stats, dt, dt2, dt3, change

    def wall_shear(self, uf):
        """Skin friction along x as run/compare_results.m:63,76-79 computes it from a snapshot:
        u = spanwise mean of u + ubar (spanwiseavg.m:3), dz = zpts(nz) - zpts(nzp),
        utau = xnu (u(:, nz) - u(:, nzp))/dz, cf = utau/(0.5 u0^2).

        x_sep: first index i with cf(i-1) >= 0 > cf(i), x_reatt: the first later i with
        cf(i-1) < 0 <= cf(i); both placed at the linear zero of cf between xpts(i-1) and
        xpts(i), NaN when there is none.  Returns (cf, x_sep, x_reatt)."""
        c = self.cfg
        nx, ny, nz = self.nx, self.ny, self.nz
        ph = self.horfft_inv(uf[nz - 1 : nz + 1])[:, :, :nx]
        um = ph.sum(axis=1) / float(ny) + self.ubar[nz - 1 : nz + 1, None]
        dz = self.zpts[nz - 1] - self.zpts[nz]
        cf = ((c.xnu * (um[0] - um[1])) / dz) / (0.5 * c.u0**2)
        return (cf,) + self._cf_crossings(cf)

    def _cf_crossings(self, cf):
        xs = self.xpts
        sep = reatt = float("nan")
        i0 = None
        for i in range(1, len(cf)):
            if cf[i - 1] >= 0.0 > cf[i]:
                sep = xs[i - 1] + (xs[i] - xs[i - 1]) * cf[i - 1] / (cf[i - 1] - cf[i])
                i0 = i
                break
        if i0 is not None:
            for i in range(i0 + 1, len(cf)):
                if cf[i - 1] < 0.0 <= cf[i]:
                    reatt = xs[i - 1] + (xs[i] - xs[i - 1]) * cf[i - 1] / (cf[i - 1] - cf[i])
                    break
        return float(sep), float(reatt)

    # --------------------------------------------- snapshot / restart records
    SNAP_HEADER = 7
    RESTART_HEADER = 9
    RESTART_FIELDS = 20

    def snapshot_record(self, uf, vf, wf, tempf, pf, name, isum, t, dt):
        """mpi_io.f90:339-497 mpi_snapshot as the float32 values of the file it writes.

        Header (:408) [isum, t, dt, nx, ny, nzp, nypl]; then the physical u, v, w, temp, p
        (:380-384), each in x, z, y order with x fastest (:411-424); u += ubar(k) and
        temp += tbar(k) unless name == 2 (:387-393).  With P ranks each rank writes its y-slab
        at offset (ivar-1) res + myid bufsize (:449), so the file is the same for any P except
        the nypl entry of the header (ny here, one rank)."""
        nx, ny, nzp = self.nx, self.ny, self.nzp
        head = np.array([isum, t, dt, nx, ny, nzp, ny], dtype=np.float64).astype(np.float32)
        phys = [None if f is None else self.horfft_inv(f) for f in (uf, vf, wf, tempf, pf)]
        return np.concatenate([head, self.snapshot_pack(phys, name)])

    def snapshot_pack(self, phys, name):
        """Body of the snapshot record from the five physical fields ((nzp, ny, nxpp) each,
        None = zero): mpi_io.f90:387-393 base profiles, :411-424 (x, z, y) order, real(4)."""
        nx, ny, nzp = self.nx, self.ny, self.nzp
        body = np.empty((5, ny, nzp, nx), dtype=np.float32)
        for l, ph in enumerate(phys):
            p = np.zeros((nzp, ny, nx)) if ph is None else ph[:, :, :nx]
            if name != 2 and l == 0:
                p = p + self.ubar[:, None, None]
            if name != 2 and l == 3:
                p = p + self.tbar[:, None, None]
            body[l] = p.transpose(1, 0, 2)
        return body.ravel()

    @staticmethod
    def read_snapshot(rec, nx, ny, nzp):
        """mpi_io.f90:500-650 mpi_read_snapshot (and run/readsnapshot.m): header values and
        q(l, k, j, i) = buf((j-1) nzp nx + (k-1) nx + i, l), i.e. numpy (5, nzp, ny, nx)."""
        head = np.asarray(rec[:7], dtype=np.float64)
        if (int(round(head[3])), int(round(head[4])), int(round(head[5]))) != (nx, ny, nzp):
            raise ValueError("snapshot: dimension mismatch")
        body = np.asarray(rec[7 : 7 + 5 * nx * ny * nzp]).reshape(5, ny, nzp, nx)
        return head, body.transpose(0, 2, 1, 3).astype(np.float64)

    def restart_record(self, st, q12_compat=False):
        """mpi_io.f90:9-160 mpi_write_restart_file as the float32 values of the file.

        Header (:44) [isum, t, dt, dt2, dt3, nx, ny, nzp, nxpl]; 20 spectral fields (:53-70)
        uf, vf, wf, tempf, unif(1:3), unmif(1:3), nlnif(1:3), slots 14-16, tnf, tnmf, nltnf,
        nltnmf, each in z, y, x order with z fastest (:49-51).  The reference fills slots
        14-16 with nlnif again (:61-62, quirk Q12) while mpi_read_restart_file reads them
        into nlnmif (:264-267); this writer stores nlnmif there (q12_compat keeps the
        duplicate)."""
        nx, ny, nzp = self.nx, self.ny, self.nzp
        head = np.array([st.get("isum", st["istep"]), st["t"], st["dt"], st["dt2"], st["dt3"], nx, ny, nzp, nx],
                        dtype=np.float64).astype(np.float32)
        slot14 = st["nlnif"] if q12_compat else st["nlnmif"]
        fields = [st["uf"], st["vf"], st["wf"], st["tempf"]]
        fields += [st["unif"][l] for l in range(3)] + [st["unmif"][l] for l in range(3)]
        fields += [st["nlnif"][l] for l in range(3)] + [slot14[l] for l in range(3)]
        fields += [st.get(k, np.zeros((nzp, ny, nx))) for k in ("tnf", "tnmf", "nltnf", "nltnmf")]
        body = np.stack([np.asarray(f).transpose(2, 1, 0) for f in fields]).astype(np.float32)
        return np.concatenate([head, body.ravel()])

    def read_restart(self, rec):
        """mpi_io.f90:163-336 mpi_read_restart_file: header (:229-240, isum = nint), the
        dimension check (:243-252) and the 20 fields back into the core arrays (:255-285).
        The restarted run's own step counter istep starts again at 0 (main.f90:61 loops
        istep = 1..nsteps), so its first steps follow the nstart start-up rule."""
        nx, ny, nzp = self.nx, self.ny, self.nzp
        h = np.asarray(rec[:9], dtype=np.float64)
        if (int(round(h[5])), int(round(h[6])), int(round(h[7]))) != (nx, ny, nzp):
            raise ValueError("restart: dimension mismatch")
        body = np.asarray(rec[9 : 9 + 20 * nx * ny * nzp]).reshape(20, nx, ny, nzp)
        f = [b.transpose(2, 1, 0).astype(np.float64) for b in body]
        st = dict(isum=int(round(h[0])), istep=0, t=float(h[1]), dt=float(h[2]), dt2=float(h[3]), dt3=float(h[4]),
                  uf=f[0], vf=f[1], wf=f[2], tempf=f[3], unif=np.stack(f[4:7]), unmif=np.stack(f[7:10]),
                  nlnif=np.stack(f[10:13]), nlnmif=np.stack(f[13:16]), tnf=f[16], tnmf=f[17],
                  nltnf=f[18], nltnmf=f[19])
        return st


__all__ = ["Oracle", "OracleConfig", "PI", "field"]
