"""Host-side mirror of the reference's hot-path modules over the C-ABI.

Method names and argument meanings follow the Fortran routines they replace
(horfft.f90, derivatives.f90, chebyshev.f90, helmholtz.f90, sgsterm.f90,
filters.f90, dealiasing.f90, navierstokes.f90).  Arrays are numpy float64 in
the reference's memory layout:

* spectral  ``(nzp, ny, nx)``      (Fortran ``qf(nxpl, ny, nzp)``)
* physical  ``(nzp, ny, nx + 2)``  (Fortran ``q(nxpp, nypl, nzp)``)
* vectors   ``(3, nzp, ...)``
* tensors   ``(3, 3, nzp, ...)`` indexed ``[l][m]``; converted to the
  reference's ``(..., 3, 3)`` order (element (l,m) = field l + 3m) at the call.

Every call runs the HIP kernels; there is no CPU path.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LsbConfig, check

PI = 3.141592653589793


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _to_fortran_tensor(t):
    """[l][m] numpy tensor -> contiguous (m, l, ...) buffer (element (l,m) at field l + 3m)."""
    return np.ascontiguousarray(np.swapaxes(np.asarray(t, dtype=np.float64), 0, 1))


def _from_fortran_tensor(buf):
    return np.ascontiguousarray(np.swapaxes(buf, 0, 1))


@dataclass
class Config:
    """Namelist snapshot passed to ``lsb_create`` (reference names)."""

    nx: int
    ny: int
    nz: int
    mapping: int = 3
    bctype: int = 3
    forcing: int = 0
    sgsflag: int = 0
    sgsmodel: int = 0
    dealiasflag: int = 2
    dealias_ref_compat: int = 1
    dct_exact: int = 0
    top_bc: int = 0
    dyn_avg: int = 0
    nut_ceil0: int = 0
    ifrz: int = 0
    nstart: int = 0
    precision: int = 0
    device: int = 0
    xlen: float = 2.0 * PI
    ylen: float = PI
    zlen: float = 2.0
    x0: float = 0.0
    u0: float = 1.0
    xnu: float = 1.0e-3
    blfac: float = 8.0
    prtl: float = 0.71
    a: float = 1.0
    b: float = 2.0
    cs: float = 0.1
    kc: float = 0.6
    filt_order: float = 1.0
    filt_amp: float = 20.0
    samp: float = 0.0
    istartbound: float = 0.0
    dt: float = 1.0e-3

    @classmethod
    def from_case(cls, case, **kw) -> "Config":
        keys = {f for f in cls.__dataclass_fields__}
        d = {k: v for k, v in case.as_dict().items() if k in keys}
        d.update(kw)
        return cls(**d)

    def to_c(self) -> LsbConfig:
        c = LsbConfig()
        for name, _ in LsbConfig._fields_:
            setattr(c, name, getattr(self, name))
        return c


# ----------------------------------------------------------- host-only setup
def build_cheb_zder_matrix(nzp: int, xl: float, gf):
    """chebyshev.f90:9-48 -> (dor, d, d2) as (nzp, nzp) arrays [row, col]."""
    gf = _f64(gf)
    dor = np.zeros((nzp, nzp), order="F")
    d = np.zeros((nzp, nzp), order="F")
    d2 = np.zeros((nzp, nzp), order="F")
    check(_lib.lib().lsb_build_cheb_zder_matrix(nzp, xl, _p(gf), _pf(dor), _pf(d), _pf(d2)), "build_cheb_zder_matrix")
    return np.ascontiguousarray(dor), np.ascontiguousarray(d), np.ascontiguousarray(d2)


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def sponge_lambda(xpts, xlen: float, u0: float, fsize: float, lfsize: float, sfsize: float):
    """navierstokes.f90:119-187: the fringe coefficient lambda(x) on xpts (host)."""
    xp = _f64(xpts)
    lam = np.zeros(xp.size)
    check(_lib.lib().lsb_sponge_lambda(xp.size, _p(xp), xlen, u0, fsize, lfsize, sfsize, _p(lam)), "sponge_lambda")
    return lam


def solve_orrsomm(d, d2, ubar, ubarp, ubarpp, zpts, xnu: float, eps: float, alphax: float, betax: float,
                  omega: complex, iglb: int = 0):
    """orrsomm.f90:110-562 on the host: (uvec, vvec, wvec, omega); d, d2 as (nzp, nzp) [row, col]."""
    nzp = len(zpts)
    a = [np.asfortranarray(d, dtype=np.float64), np.asfortranarray(d2, dtype=np.float64)]
    v = [_f64(x) for x in (ubar, ubarp, ubarpp, zpts)]
    om = np.array([complex(omega).real, complex(omega).imag])
    out = [np.zeros(2 * nzp) for _ in range(3)]
    check(_lib.lib().lsb_solve_orrsomm(nzp, _pf(a[0]), _pf(a[1]), *[_p(x) for x in v], xnu, eps, alphax, betax,
                                       _p(om), iglb, *[_p(x) for x in out]), "solve_orrsomm")
    u, vv, w = (x[0::2] + 1j * x[1::2] for x in out)
    return u, vv, w, complex(om[0], om[1])


def dct_diff(y, xl: float, gf, p: int):
    """chebyshev.f90:204-270 along axis 0 of y (n, ...)."""
    y = _f64(y)
    n = y.shape[0]
    P = int(np.prod(y.shape[1:])) if y.ndim > 1 else 1
    out = np.empty_like(y)
    check(_lib.lib().lsb_dct_diff(P, 1, n, xl, _p(_f64(gf)), _p(y), _p(out), p), "dct_diff")
    return out


def diagonalize_matrix(D2):
    """helmholtz.f90:13-126: returns (lamda, E, E^-1) with E[:, i] the eigenvectors."""
    D2 = np.asfortranarray(D2, dtype=np.float64)
    n = D2.shape[0]
    lam = np.zeros(n)
    e = np.zeros((n, n), order="F")
    ei = np.zeros((n, n), order="F")
    check(_lib.lib().lsb_diagonalize_matrix(_pf(D2), n, _p(lam), _pf(e), _pf(ei)), "diagonalize_matrix")
    return lam, np.ascontiguousarray(e), np.ascontiguousarray(ei)


# ---------------------------------------------------------------- context
PROF_CLASSES = 6  # LSB_PROF_CLASSES
class Context:
    """One HIP context: module globals snapshot + device operators + state."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._L = _lib.lib()
        h = C.c_void_p()
        cc = cfg.to_c()
        check(self._L.lsb_create(C.byref(cc), C.byref(h)), "lsb_create")
        self.h = h
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        self.nzp = cfg.nz + 1
        self.nzm = cfg.nz - 1
        self.nxpp = cfg.nx + 2
        self.sshape = (self.nzp, self.ny, self.nx)
        self.pshape = (self.nzp, self.ny, self.nxpp)

    def close(self):
        if self.h:
            self._L.lsb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------- snapshot
    def set_paral(self, ubar, ubarp, ubarpp, wbar=None, wbarpp=None):
        a = [None if x is None else _f64(x) for x in (ubar, ubarp, ubarpp, wbar, wbarpp)]
        check(self._L.lsb_set_paral(self.h, *[_p(x) for x in a]), "set_paral")

    def set_grid(self, wavx, wavy, zpts, gf, xpts, dx, dy, xl):
        a = [_f64(x) for x in (wavx, wavy, zpts, gf, xpts)]
        check(self._L.lsb_set_grid(self.h, *[_p(x) for x in a], dx, dy, xl), "set_grid")

    def set_mats(self, d=None, d2=None, diag=None, smat=None, simat=None):
        def fo(x):
            return None if x is None else np.asfortranarray(x, dtype=np.float64)
        a = [fo(d), fo(d2), None if diag is None else _f64(diag), fo(smat), fo(simat)]
        check(self._L.lsb_set_mats(self.h, *[None if x is None else _pf(x) for x in a]), "set_mats")

    def grid(self):
        wavx, wavy = np.zeros(self.nx), np.zeros(self.ny)
        zpts, gf, xpts = np.zeros(self.nzp), np.zeros(self.nzp), np.zeros(self.nx)
        dd = np.zeros(3)
        check(self._L.lsb_get_grid(self.h, _p(wavx), _p(wavy), _p(zpts), _p(gf), _p(xpts), _p(dd)), "get_grid")
        return dict(wavx=wavx, wavy=wavy, zpts=zpts, gf=gf, xpts=xpts, dx=dd[0], dy=dd[1], xl=dd[2])

    def mats(self):
        n, m = self.nzp, self.nzm
        arrs = {k: np.zeros((n, n), order="F") for k in ("dor", "d", "d2", "mdz", "md2")}
        diag = np.zeros(m)
        smat, simat = np.zeros((m, m), order="F"), np.zeros((m, m), order="F")
        check(self._L.lsb_get_mats(self.h, *[_pf(arrs[k]) for k in ("dor", "d", "d2", "mdz", "md2")],
                                   _p(diag), _pf(smat), _pf(simat)), "get_mats")
        out = {k: np.ascontiguousarray(v) for k, v in arrs.items()}
        out.update(diag=diag, smat=np.ascontiguousarray(smat), simat=np.ascontiguousarray(simat))
        return out

    # ------------------------------------------------------------ modhorfft
    def horfft(self, ain, ainf, is_):
        """horfft.f90:61-150 in place: is_=-1 fills ainf from ain, is_=+1 fills ain from ainf.
        Leading axes beyond (nzp, ny, nx[pp]) are independent fields."""
        assert ain.dtype == np.float64 and ainf.dtype == np.float64
        assert ain.flags.c_contiguous and ainf.flags.c_contiguous
        nf = int(np.prod(ain.shape[:-3])) if ain.ndim > 3 else 1
        check(self._L.lsb_horfft(self.h, _p(ain), _p(ainf), is_, nf), "horfft")

    def fwd(self, phys):
        phys = _f64(phys)
        out = np.zeros(phys.shape[:-1] + (self.nx,))
        self.horfft(phys, out, -1)
        return out

    def inv(self, spec):
        spec = _f64(spec)
        out = np.zeros(spec.shape[:-1] + (self.nxpp,))
        self.horfft(out, spec, 1)
        return out

    def norm(self, uf, vf, wf, tempf=None):
        """horfft.f90:153-224 in place."""
        check(self._L.lsb_norm(self.h, _p(uf), _p(vf), _p(wf), _p(tempf)), "norm")

    # ---------------------------------------------------------- derivatives
    def _unary(self, fn, uf):
        uf = _f64(uf)
        out = np.empty_like(uf)
        check(fn(self.h, _p(uf), _p(out)), fn.__name__)
        return out

    def dfdx(self, uf):
        return self._unary(self._L.lsb_dfdx, uf)

    def dfdy(self, uf):
        return self._unary(self._L.lsb_dfdy, uf)

    def dfdz(self, uf):
        return self._unary(self._L.lsb_dfdz, uf)

    def d2fdz2(self, uf):
        return self._unary(self._L.lsb_d2fdz2, uf)

    def lapl(self, uf):
        return self._unary(self._L.lsb_lapl, uf)

    def dpdz(self, u, p):
        """derivatives.f90:135-158 on (nzp, n2, n1) arrays."""
        u = _f64(u)
        out = np.empty_like(u)
        n1 = u.shape[-1]
        n2 = int(np.prod(u.shape[1:-1])) if u.ndim > 2 else 1
        check(self._L.lsb_dpdz(self.h, n1, n2, _p(u), _p(out), p), "dpdz")
        return out

    def divuv(self, uf, vf):
        uf, vf = _f64(uf), _f64(vf)
        out = np.empty_like(uf)
        check(self._L.lsb_divuv(self.h, _p(uf), _p(vf), _p(out)), "divuv")
        return out

    def div(self, uf, vf, wf):
        uf, vf, wf = _f64(uf), _f64(vf), _f64(wf)
        out = np.empty_like(uf)
        check(self._L.lsb_div(self.h, _p(uf), _p(vf), _p(wf), _p(out)), "div")
        return out

    def divcheck(self, uf, vf, wf):
        uf, vf, wf = _f64(uf), _f64(vf), _f64(wf)
        r = C.c_double()
        check(self._L.lsb_divcheck(self.h, _p(uf), _p(vf), _p(wf), C.byref(r)), "divcheck")
        return r.value

    # ------------------------------------------------------------ helmholtz
    def solve_helmholtz(self, rhs, dtnui, diag=None, smat=None, simat=None):
        out = np.array(rhs, dtype=np.float64, copy=True, order="C")
        m = self.mats() if diag is None else None
        diag = m["diag"] if diag is None else _f64(diag)
        smat = np.asfortranarray(m["smat"] if smat is None else smat)
        simat = np.asfortranarray(m["simat"] if simat is None else simat)
        check(self._L.lsb_solve_helmholtz(self.h, _p(out), dtnui, self.nzm, _p(diag), _pf(smat), _pf(simat)),
              "solve_helmholtz")
        return out

    def solve_nhom_helmholtz(self, rhs, bc, dtnui):
        out = np.array(rhs, dtype=np.float64, copy=True, order="C")
        bc = _f64(bc)
        check(self._L.lsb_solve_nhom_helmholtz(self.h, _p(out), _p(bc), dtnui, None, None, None),
              "solve_nhom_helmholtz")
        return out

    # ----------------------------------------------------------- sgs_models
    def _tau(self, fn, *args):
        out = np.zeros((3, 3) + self.pshape)
        check(fn(self.h, *args, _p(out)), fn.__name__)
        return _from_fortran_tensor(out)

    def sigma_model(self, duidxj):
        g = _to_fortran_tensor(duidxj)
        return self._tau(self._L.lsb_sigma_model, _p(g))

    def static_smagorinsky(self, cs, duidxj):
        g = _to_fortran_tensor(duidxj)
        out = np.zeros((3, 3) + self.pshape)
        check(self._L.lsb_static_smagorinsky(self.h, cs, _p(g), _p(out)), "static_smagorinsky")
        return _from_fortran_tensor(out)

    def dynamic_smagorinsky(self, ui, uiuj, duidxj):
        u = _f64(ui)
        uu = _to_fortran_tensor(uiuj)
        g = _to_fortran_tensor(duidxj)
        return self._tau(self._L.lsb_dynamic_smagorinsky, _p(u), _p(uu), _p(g))

    def interscale_model(self, ui):
        return self._tau(self._L.lsb_interscale_model, _p(_f64(ui)))

    def tau_sgs(self, ui, uiuj, duidxj):
        u = _f64(ui)
        uu = _to_fortran_tensor(uiuj)
        g = _to_fortran_tensor(duidxj)
        return self._tau(self._L.lsb_tau_sgs, _p(u), _p(uu), _p(g))

    # -------------------------------------------------------------- filters
    def filter_q_3pt(self, q, a, b, interpflag=1):
        """filters.f90:363-458; q (..., nzp, ny, nxpp) with the leading axes as (n5, n4)."""
        q = _f64(q)
        nf = int(np.prod(q.shape[:-3])) if q.ndim > 3 else 1
        out = np.zeros_like(q)
        check(self._L.lsb_filter_q_3pt(self.h, _p(q), _p(out), nf, 1, a, b, interpflag), "filter_q_3pt")
        return out

    def specfilter_exp_xy(self, uf, kc, p, amp, filtery=1):
        uf = _f64(uf)
        out = np.empty_like(uf)
        check(self._L.lsb_specfilter_exp_xy(self.h, _p(uf), _p(out), kc, p, amp, filtery), "specfilter_exp_xy")
        return out

    def filterqspec(self, qf):
        out = np.array(qf, dtype=np.float64, copy=True, order="C")
        nq = int(np.prod(out.shape[:-3])) if out.ndim > 3 else 1
        check(self._L.lsb_filterqspec(self.h, _p(out), nq), "filterqspec")
        return out

    # --------------------------------------------------------------- dealias
    def setup_dealiasing(self):
        a, b = C.c_int(), C.c_int()
        check(self._L.lsb_setup_dealiasing(self.h, C.byref(a), C.byref(b)), "setup_dealiasing")
        return a.value, b.value

    def dealiasing_2d_fftw(self, a, b):
        a, b = _f64(a), _f64(b)
        out = np.zeros_like(a)
        check(self._L.lsb_dealiasing_2d_fftw(self.h, _p(a), _p(b), _p(out)), "dealiasing_2d_fftw")
        return out

    # ------------------------------------------------------------------ step
    def nlterm_skewsym(self, uif):
        uif = _f64(uif)
        out = np.zeros_like(uif)
        check(self._L.lsb_nlterm_skewsym(self.h, _p(uif), _p(out)), "nlterm_skewsym")
        return out

    def state_set(self, uf, vf, wf, pf=None, pnf=None, unif=None, unmif=None, nlnif=None, nlnmif=None, istep=0,
                  t=0.0):
        arrs = [None if x is None else _f64(x) for x in (uf, vf, wf, pf, pnf, unif, unmif, nlnif, nlnmif)]
        check(self._L.lsb_state_set(self.h, *[_p(x) for x in arrs], istep, t), "state_set")

    def state_get(self):
        s = self.sshape
        d = dict(uf=np.zeros(s), vf=np.zeros(s), wf=np.zeros(s), pf=np.zeros(s), pnf=np.zeros(s),
                 unif=np.zeros((3,) + s), unmif=np.zeros((3,) + s), nlnif=np.zeros((3,) + s),
                 nlnmif=np.zeros((3,) + s))
        ist = C.c_int()
        t = C.c_double()
        keys = ("uf", "vf", "wf", "pf", "pnf", "unif", "unmif", "nlnif", "nlnmif")
        check(self._L.lsb_state_get(self.h, *[_p(d[k]) for k in keys], C.byref(ist), C.byref(t)), "state_get")
        d["istep"] = ist.value
        d["t"] = t.value
        return d

    def timestep(self, nsteps=1):
        check(self._L.lsb_timestep(self.h, nsteps), "timestep")

    def navierstokes_update(self):
        check(self._L.lsb_navierstokes_update(self.h), "navierstokes_update")

    def device_ptrs(self):
        ps = [C.c_void_p() for _ in range(4)]
        check(self._L.lsb_state_device_ptrs(self.h, *[C.byref(p) for p in ps]), "state_device_ptrs")
        return [p.value for p in ps]

    def stream(self):
        return self._L.lsb_stream(self.h)

    def profile_begin(self):
        check(self._L.lsb_profile_begin(self.h), "profile_begin")

    def profile_begin_classes(self, classes):
        """Profile only the launch groups of the named classes (see profile_end)."""
        keys = ("fft", "zcontract", "helmholtz", "pointwise", "rhs", "nonlinear")
        mask = 0
        for k in classes:
            mask |= 1 << keys.index(k)
        check(self._L.lsb_profile_begin_classes(self.h, mask), "profile_begin_classes")

    def profile_end(self):
        """{class: dict(ms=total device ms, n=launch groups, flops=..., bytes=...)}."""
        ms = np.zeros(PROF_CLASSES)
        n = np.zeros(PROF_CLASSES, dtype=np.int32)
        fl = np.zeros(PROF_CLASSES)
        by = np.zeros(PROF_CLASSES)
        check(self._L.lsb_profile_end(self.h, _p(ms), n.ctypes.data_as(C.POINTER(C.c_int)), _p(fl), _p(by)),
              "profile_end")
        keys = ("fft", "zcontract", "helmholtz", "pointwise", "rhs", "nonlinear")
        return {k: dict(ms=float(ms[i]), n=int(n[i]), flops=float(fl[i]), bytes=float(by[i])) for i, k in enumerate(keys)}

    def synchronize(self):
        check(self._L.lsb_synchronize(self.h), "synchronize")

    # the filter families hold a launch variant, the y-pass families the kx columns per
    # tile (-1: the built-in default)
    TUNE_FAMILIES = ("filter3_f64", "filter3_f32", "yfwd_f64", "yfwd_f32", "yinv3_f64", "yinv3_f32", "yrhs_f32")

    def autotune(self):
        """Time the launch variants of the tunable kernels on this grid; returns the table."""
        check(self._L.lsb_autotune(self.h), "autotune")
        return self.tune_table()

    def tune_table(self):
        n = len(self.TUNE_FAMILIES)
        v = (C.c_int * n)()
        ms = (C.c_float * n)()
        check(self._L.lsb_autotune_get(self.h, v, ms, n), "autotune_get")
        return {f: {"variant": v[i], "ms": round(float(ms[i]), 4)} for i, f in enumerate(self.TUNE_FAMILIES)}

    def set_tune_table(self, table):
        """Set the given families (the others keep their current choice)."""
        cur = self.tune_table()
        n = len(self.TUNE_FAMILIES)
        v = (C.c_int * n)(*[int(table.get(f, cur[f])["variant"]) for f in self.TUNE_FAMILIES])
        check(self._L.lsb_autotune_set(self.h, v, n), "autotune_set")

    # ------------------------------------------ snapshots, restart, diagnostics
    def _fp(self, a):
        return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None

    def snapshot_record(self, uf, vf, wf, tempf, pf, name=0, isum=0, t=0.0, dt=0.0, tbar=None):
        """mpi_snapshot record (7 + 5*nx*ny*nzp float32) of host spectral fields."""
        n = self.nx * self.ny * self.nzp
        rec = np.zeros(7 + 5 * n, dtype=np.float32)
        a = [None if f is None else _f64(f) for f in (uf, vf, wf, tempf, pf, tbar)]
        check(self._L.lsb_snapshot_record(self.h, *[_p(x) for x in a], name, isum, t, dt, self._fp(rec)),
              "snapshot_record")
        return rec

    def state_snapshot(self, name=0, path=None, record=True):
        n = self.nx * self.ny * self.nzp
        rec = np.zeros(7 + 5 * n, dtype=np.float32) if record else None
        check(self._L.lsb_state_snapshot(self.h, name, None if path is None else os.fsencode(path), self._fp(rec)),
              "state_snapshot")
        return rec

    def dealiasf_2d_fftw(self, af, bf):
        """dealiasing.f90:449-618: padded product of two spectral fields, spectral out."""
        out = np.zeros(self.sshape)
        check(self._L.lsb_dealiasf_2d_fftw(self.h, _p(_f64(af)), _p(_f64(bf)), _p(out)), "dealiasf_2d_fftw")
        return out

    def test_fft(self, u):
        """horfft.f90:7-58: mean |u - inverse(forward(u))| over the physical points."""
        e = C.c_double()
        check(self._L.lsb_test_fft(self.h, _p(_f64(u)), C.byref(e)), "test_fft")
        return e.value

    def zinteg(self, tmp, mapping=None, per=0, rev=0):
        """derivatives.f90:9-70 on tmp (nzp, n2, n1) (numpy order of Fortran (n1, n2, nzp))."""
        t = _f64(tmp)
        out = np.zeros(t.shape)
        mp = self.cfg.mapping if mapping is None else mapping
        check(self._L.lsb_zinteg(self.h, _p(t), _p(out), t.shape[2], t.shape[1], mp, per, rev), "zinteg")
        return out

    def set_time_average(self, otfavg=1, istartavg=0, iomod=1):
        """main.f90:69-76: on-the-fly averages of uf, vf, wf, tempf, pf over the steps of
        timestep() once the total step count exceeds istartavg (q/iomod per step)."""
        check(self._L.lsb_set_time_average(self.h, otfavg, istartavg, iomod), "set_time_average")

    def state_snapshot_average(self, path=None, record=True):
        """main.f90:98-106: the qavg snapshot record of the averages, which then restart at zero."""
        n = self.nx * self.ny * self.nzp
        rec = np.zeros(7 + 5 * n, dtype=np.float32) if record else None
        check(self._L.lsb_state_snapshot_average(self.h, None if path is None else os.fsencode(path),
                                                 self._fp(rec)), "state_snapshot_average")
        return rec

    def state_get_time_average(self):
        """(uf, vf, wf, tempf, pf) averages, spectral."""
        out = [np.zeros(self.sshape) for _ in range(5)]
        check(self._L.lsb_state_get_time_average(self.h, *[_p(x) for x in out]), "state_get_time_average")
        return tuple(out)

    def read_snapshot(self, path):
        """(q (5, nzp, ny, nx), isum, t, dt) of a snapshot file."""
        q = np.zeros((5,) + self.sshape)
        isum, t, dt = C.c_int(), C.c_double(), C.c_double()
        check(self._L.lsb_read_snapshot(self.h, os.fsencode(path), _p(q), C.byref(isum), C.byref(t), C.byref(dt)),
              "read_snapshot")
        return q, isum.value, t.value, dt.value

    def write_restart(self, path=None, record=True):
        n = self.nx * self.ny * self.nzp
        rec = np.zeros(9 + 20 * n, dtype=np.float32) if record else None
        check(self._L.lsb_write_restart(self.h, None if path is None else os.fsencode(path), self._fp(rec)),
              "write_restart")
        return rec

    def read_restart(self, path=None, record=None):
        rec = None if record is None else np.ascontiguousarray(record, dtype=np.float32)
        check(self._L.lsb_read_restart(self.h, None if path is None else os.fsencode(path), self._fp(rec)),
              "read_restart")

    def maxcalc(self, uf, vf, wf, tempf=None, dtmax=1.0, dt=None, dt2=None, dt3=None, t=0.0, fsize=0.0, tbar=None):
        """(stats[11], dt, dt2, dt3, change_dt) as navierstokes.f90 maxcalc."""
        a = [None if f is None else _f64(f) for f in (uf, vf, wf, tempf, tbar)]
        d1, d2, d3 = (C.c_double(self.cfg.dt if x is None else x) for x in (dt, dt2, dt3))
        stats = np.zeros(11)
        ch = C.c_int()
        check(self._L.lsb_maxcalc(self.h, *[_p(x) for x in a], dtmax, t, fsize, C.byref(d1), C.byref(d2),
                                  C.byref(d3), _p(stats), C.byref(ch)), "maxcalc")
        return stats, d1.value, d2.value, d3.value, bool(ch.value)

    def state_maxcalc(self, dtmax=1.0, fsize=0.0):
        stats = np.zeros(11)
        ch = C.c_int()
        check(self._L.lsb_state_maxcalc(self.h, dtmax, fsize, _p(stats), C.byref(ch)), "state_maxcalc")
        return stats, bool(ch.value)

    def nancheck(self, uf, vf=None, wf=None, tempf=None):
        a = [None if f is None else _f64(f) for f in (uf, vf, wf, tempf)]
        r = C.c_int()
        check(self._L.lsb_nancheck(self.h, *[_p(x) for x in a], C.byref(r)), "nancheck")
        return bool(r.value)

    def state_nancheck(self):
        r = C.c_int()
        check(self._L.lsb_state_nancheck(self.h, C.byref(r)), "state_nancheck")
        return bool(r.value)

    def wall_shear(self, uf=None):
        """(cf (nx,), x_sep, x_reatt) of host spectral uf, or of the resident state."""
        cf = np.zeros(self.nx)
        xs, xr = C.c_double(), C.c_double()
        if uf is None:
            check(self._L.lsb_state_wall_shear(self.h, _p(cf), C.byref(xs), C.byref(xr)), "state_wall_shear")
        else:
            u = _f64(uf)
            check(self._L.lsb_wall_shear(self.h, _p(u), _p(cf), C.byref(xs), C.byref(xr)), "wall_shear")
        return cf, xs.value, xr.value

    def set_adaptive_dt(self, dtadapt=1, iomod=5, fsize=0.0):
        check(self._L.lsb_set_adaptive_dt(self.h, dtadapt, iomod, fsize), "set_adaptive_dt")

    def set_sponge(self, fsize=0.15, lfsize=0.25, sfsize=0.70):
        check(self._L.lsb_set_sponge(self.h, fsize, lfsize, sfsize), "set_sponge")

    def filterall3pt(self, uf, vf, wf, tempf, pf):
        """filters.f90:9-49 on copies; returns (uf, vf, wf, tempf, pf)."""
        a = [None if f is None else np.array(f, dtype=np.float64, copy=True, order="C")
             for f in (uf, vf, wf, tempf, pf)]
        check(self._L.lsb_filterall3pt(self.h, *[_p(x) for x in a]), "filterall3pt")
        return tuple(a)

    def tns_automatic_filtering(self, uf, vf, wf, tempf, pf):
        """filters.f90:52-161 on copies; returns (uf, vf, wf, tempf, pf, Ih, Idelta, applied)."""
        a = [None if f is None else np.array(f, dtype=np.float64, copy=True, order="C")
             for f in (uf, vf, wf, tempf, pf)]
        ih, idl, ap = C.c_double(), C.c_double(), C.c_int()
        check(self._L.lsb_tns_automatic_filtering(self.h, *[_p(x) for x in a], C.byref(ih), C.byref(idl),
                                                  C.byref(ap)), "tns_automatic_filtering")
        return tuple(a) + (ih.value, idl.value, bool(ap.value))

    def state_filterall3pt(self):
        check(self._L.lsb_state_filterall3pt(self.h), "state_filterall3pt")

    def state_tns_filter(self):
        ih, idl, ap = C.c_double(), C.c_double(), C.c_int()
        check(self._L.lsb_state_tns_filter(self.h, C.byref(ih), C.byref(idl), C.byref(ap)), "state_tns_filter")
        return ih.value, idl.value, bool(ap.value)

    def ener(self, uf, vf, wf, itot=1):
        a = [_f64(f) for f in (uf, vf, wf)]
        en = C.c_double()
        check(self._L.lsb_ener(self.h, *[_p(x) for x in a], itot, C.byref(en)), "ener")
        return en.value

    def lintest(self, uf, vf, wf, istep, dt, enr1, enr2, grow1, grow2):
        """orrsomm.f90 lintest: fills entry istep-1 of the four float64 arrays in place."""
        a = [_f64(f) for f in (uf, vf, wf)]
        for arr in (enr1, enr2, grow1, grow2):
            assert arr.dtype == np.float64 and arr.flags.c_contiguous and arr.size >= istep
        check(self._L.lsb_lintest(self.h, *[_p(x) for x in a], istep, dt, _p(enr1), _p(enr2), _p(grow1),
                                  _p(grow2)), "lintest")

    def state_lintest(self, enr1, enr2, grow1, grow2):
        for arr in (enr1, enr2, grow1, grow2):
            assert arr.dtype == np.float64 and arr.flags.c_contiguous
        check(self._L.lsb_state_lintest(self.h, _p(enr1), _p(enr2), _p(grow1), _p(grow2)), "state_lintest")

    def set_temperature(self, tbar=None, tbarp=None, tbarpp=None, coex=0.0, gx=0.0, gy=0.0, gz=0.0):
        a = [None if x is None else _f64(x) for x in (tbar, tbarp, tbarpp)]
        check(self._L.lsb_set_temperature(self.h, *[_p(x) for x in a], coex, gx, gy, gz), "set_temperature")

    def state_set_temperature(self, tempf=None, tnf=None, tnmf=None, nltnf=None, nltnmf=None):
        a = [None if x is None else _f64(x) for x in (tempf, tnf, tnmf, nltnf, nltnmf)]
        check(self._L.lsb_state_set_temperature(self.h, *[_p(x) for x in a]), "state_set_temperature")

    def state_get_temperature(self):
        keys = ("tempf", "tnf", "tnmf", "nltnf", "nltnmf")
        d = {k: np.zeros(self.sshape) for k in keys}
        check(self._L.lsb_state_get_temperature(self.h, *[_p(d[k]) for k in keys]), "state_get_temperature")
        return d

    def set_filtering(self, ifilt=1, modfilt=1):
        check(self._L.lsb_set_filtering(self.h, ifilt, modfilt), "set_filtering")


class Ensemble:
    """Contexts stepped together from captured HIP graphs (lsb_ensemble_*): each member must
    have its state set and have run its three start-up steps."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        self.h = C.c_void_p()
        self._L = _lib.lib()
        check(self._L.lsb_ensemble_create(arr, len(self.members), C.byref(self.h)), "ensemble_create")

    def timestep(self, nsteps=1):
        check(self._L.lsb_ensemble_timestep(self.h, nsteps), "ensemble_timestep")

    def synchronize(self):
        check(self._L.lsb_ensemble_synchronize(self.h), "ensemble_synchronize")

    def info(self):
        a, b, p = C.c_int(), C.c_int(), C.c_int()
        check(self._L.lsb_ensemble_info(self.h, C.byref(a), C.byref(b), C.byref(p)), "ensemble_info")
        return {"members": a.value, "graphs": b.value, "phase": p.value}

    def close(self):
        if self.h:
            check(self._L.lsb_ensemble_destroy(self.h), "ensemble_destroy")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def setup_case(case, precision: int = 0, device: int = 0, **kw) -> Context:
    """Context for a configuration of lsbflow_amd.cases with its base flow installed."""
    from . import cases
    cfg = Config.from_case(case, precision=precision, device=device, **kw)
    ctx = Context(cfg)
    ub, ubp, ubpp = cases.base_flow(case)
    ctx.set_paral(ub, ubp, ubpp)
    return ctx


__all__ = ["Config", "Context", "Ensemble", "build_cheb_zder_matrix", "dct_diff", "diagonalize_matrix", "setup_case", "field"]
