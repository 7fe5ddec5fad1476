"""The batched Helmholtz solve alone at its configured size (helmholtz.f90:130-208).

All (kx, ky) pencils of a 512 x 193 x 256 spectral grid (nx x nzp x ny; 256 x 256
pencils, the kx-Nyquist column dropped as the reference does) x 4 right-hand-side fields
in one call, with seeded N(0, 1) Chebyshev coefficients decaying as exp(-0.1 n) in every
real column, for sigma = c / (nu dt) with c in {1, 1.5} (nu = 2e-6, dt = 2.5e-4) and
sigma = 0 (the pressure Poisson problem).  Checks:
  * FP64: against the eigen-solution E diag(1/(lambda - k^2 - sigma)) E^-1 q of numpy's
    own eigen-decomposition of d2(2:nz, 2:nz) on 64 pencils; the residual of
    (D2 - k^2 - sigma) u = rhs on the interior rows of every pencil;
  * FP32 projections, and FP32 with one FP64 refinement: their error against the FP64
    solution, stated, the refinement required to gain at least four digits;
  * the mixed form's velocity-solve form (FP64 leading term -rhs/(k^2 + sigma), FP32
    correction; LSB_HELM_SPLIT32) against the FP64 solution at sigma = c/(nu dt), where
    sigma is over 16 x max|lambda|; at sigma = 0 it is the FP64 solve itself.
"""

import ctypes as C

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

import lsbflow_amd as L
from lsbflow_amd import _lib
from lsbflow_amd.core import setup_case

pytestmark = pytest.mark.gpu

NU, DT = 2.0e-6, 2.5e-4
H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, src, shape):
        a = np.empty(shape)
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, D2H) == 0
        return a


def c3_rhs(nzp, ncols, nf, seed=20261203):
    """(nf, nzp, ncols) collocation values of sum_n a_n T_n(xi), a_n ~ N(0,1) exp(-0.1 n)."""
    rng = np.random.default_rng(seed)
    n = np.arange(nzp)
    xi = np.cos(np.pi * n / (nzp - 1))              # Gauss-Lobatto points, top first
    T = npcheb.chebvander(xi, nzp - 1)               # T[k, n] = T_n(xi_k)
    out = np.empty((nf, nzp, ncols))
    for f in range(nf):
        a = rng.standard_normal((nzp, ncols)) * np.exp(-0.1 * n)[:, None]
        out[f] = T @ a
    return out


@pytest.fixture(scope="module")
def c3():
    case = L.cases.CASES["c4_sigma_les"]
    ctx = setup_case(case)
    nzp, ny, nx = ctx.nzp, ctx.ny, ctx.nx
    rhs = c3_rhs(nzp, ny * nx, 4).reshape(4, nzp, ny, nx)
    g = ctx.grid()
    m = ctx.mats()
    yield ctx, rhs, g, m
    ctx.close()


def solve_dev(ctx, hip, rhs, sigma, mode):
    nb = rhs.nbytes
    d = hip.alloc(nb)
    try:
        hip.put(d, rhs)
        rc = _lib.lib().lsb_solve_helmholtz_mixed_dev(ctx.h, d, sigma, rhs.shape[0], mode, None)
        assert rc == 0, _lib.lib().lsb_last_error()
        assert _lib.lib().lsb_synchronize(ctx.h) == 0
        return hip.get(d, rhs.shape)
    finally:
        hip.lib.hipFree(d)


@pytest.mark.parametrize("c", [1.0, 1.5, 0.0])
def test_c3_batched_helmholtz(c3, c):
    ctx, rhs, g, m = c3
    hip = Hip()
    sigma = c / (NU * DT)
    nzp, ny, nx = ctx.nzp, ctx.ny, ctx.nx
    nz = nzp - 1
    u64 = solve_dev(ctx, hip, rhs, sigma, 0)
    u32 = solve_dev(ctx, hip, rhs, sigma, 1)
    ur = solve_dev(ctx, hip, rhs, sigma, 2)
    us = solve_dev(ctx, hip, rhs, sigma, 3)
    k2 = (g["wavx"][None, :] ** 2 + g["wavy"][:, None] ** 2)          # (ny, nx) per real column
    # FP64 against numpy's eigen-decomposition on 64 pencils
    D2 = m["d2"]
    lam, E = np.linalg.eig(D2[1:nz, 1:nz])
    lam, E = lam.real, E.real
    Ei = np.linalg.inv(E)
    rng = np.random.default_rng(5)
    cols = rng.choice(ny * nx, 64, replace=False)
    jj, ii = cols // nx, cols % nx
    worst = 0.0
    for f in range(4):
        q = rhs[f, 1:nz, jj, ii].T                                      # (nzm, 64)
        ref = E @ ((Ei @ q) / (lam[:, None] - k2[jj, ii][None, :] - sigma))
        got = u64[f, 1:nz, jj, ii].T
        worst = max(worst, np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # residual on every pencil, interior rows, relative to |D2| |u| + |rhs|
    res_w = 0.0
    nD2 = np.abs(D2[1:nz]).sum(axis=1).max()
    for f in range(4):
        u = u64[f].reshape(nzp, -1)
        r = D2[1:nz] @ u - (k2.reshape(-1)[None, :] + sigma) * u[1:nz] - rhs[f].reshape(nzp, -1)[1:nz]
        scale = nD2 * np.abs(u).max() + (k2.max() + sigma) * np.abs(u).max() + np.abs(rhs[f]).max()
        res_w = max(res_w, np.abs(r).max() / scale)
    e32 = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    eref = np.linalg.norm(ur - u64) / np.linalg.norm(u64)
    espl = np.linalg.norm(us - u64) / np.linalg.norm(u64)
    print(f"c3 sigma={sigma:.3g}: FP64 vs numpy eigen-solution {worst:.2e}, FP64 residual {res_w:.2e}, "
          f"FP32 {e32:.2e}, FP32 + 1 FP64 refinement {eref:.2e}, FP64 leading + FP32 correction {espl:.2e} "
          f"(rel L2 vs FP64)", flush=True)
    if sigma == 0.0:
        assert np.array_equal(us, u64)          # below the 16 x max|lambda| margin: the FP64 solve
    else:
        assert espl < 1e-8 and espl < 1e-2 * e32
        assert np.all(us[:, 0] == 0.0) and np.all(us[:, nzp - 1] == 0.0)
    assert worst < 1e-11
    assert res_w < 1e-13
    assert np.all(u64[:, 0] == 0.0) and np.all(u64[:, nzp - 1] == 0.0)
    assert e32 < 1e-4
    assert eref < 1e-4 * e32 and eref < 1e-9
