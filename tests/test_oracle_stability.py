"""Stability of the non-homogeneous step on the channel mapping (CPU, oracle only).

Two facts decide which configuration the LES workloads can run:

1. The reference's Chebyshev derivative (dct_diff, chebyshev.f90:204-270) scales
   the DCT coefficients 2..n-1 by the half interval but not the top one
   (chebyshev.f90:242).  Its derivative of T_N is therefore exact for xl = 1
   and exactly twice the true one for the channel's xl = 2.  ``dct_exact = 1``
   scales the top coefficient by 1/xl.  Pinned here against the closed-form
   derivatives of the Chebyshev polynomials (numpy's chebder).

2. The time step is linear in the perturbation when the base flow is zero and
   the amplitude is tiny (the nonlinear term is quadratic), and it acts on each
   (kx, ky) pencil separately.  The oracle's step map of one pencil is read off
   by stepping unit states; its spectral radius is compared with an
   independent dense restatement of navierstokes.f90:1282-1488 for one pencil
   (exact Dirichlet solves of the collocation Helmholtz problems in place of
   the Green's-function lifting, no oracle code).  With the reference operator
   the non-homogeneous path (bctype > 0) has a mode that grows by ~1.6 per step
   at the LES viscosities, whichever ceiling condition is used; with the exact
   operator, and on the homogeneous path with either operator, every mode decays.
"""

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from oracle import Oracle, OracleConfig


# ------------------------------------------------------------------- operator
def _cheb_derivative_values(xi, n, scale):
    """d/dz T_n(xi(z)) at the points, xi = 2 z / xl - 1 style maps with d(xi)/dz = scale."""
    c = np.zeros(n + 1)
    c[n] = 1.0
    return npcheb.chebval(xi, npcheb.chebder(c)) * scale


@pytest.mark.parametrize("mapping,kw", [(3, dict(zlen=2.0)), (0, dict(zlen=2.0)),
                                        (2, dict(zlen=33.3333, x0=1.0, xlen=10.0))])
def test_dct_operator_against_chebyshev_derivatives(mapping, kw):
    nz = 48
    o_ref = Oracle(OracleConfig(nx=8, ny=4, nz=nz, mapping=mapping, **kw))
    o_ex = Oracle(OracleConfig(nx=8, ny=4, nz=nz, mapping=mapping, dct_exact=1, **kw))
    n = o_ref.nzp
    # computational coordinate on [-1, 1] of the Gauss-Lobatto points (k = 0 is xi = +1)
    xi = np.cos(np.pi * np.arange(n) / nz)
    # d(xi)/dz: 2/xl times the mapping factor gf (gf = 1 for the channel mappings)
    scale = (2.0 / o_ref.xl) * o_ref.gf
    for deg in (1, 2, 7, nz - 1, nz):
        t = npcheb.chebval(xi, np.eye(n)[deg])
        exact = _cheb_derivative_values(xi, deg, scale)
        big = np.abs(exact).max()
        got_ex = o_ex.dct_diff(t, 1)
        assert np.abs(got_ex - exact).max() < 1e-11 * big, deg
        got_ref = o_ref.dct_diff(t, 1)
        if deg < nz or o_ref.xl == 1.0:
            assert np.abs(got_ref - exact).max() < 1e-11 * big, deg
        else:
            # the reference's top-mode scaling: exactly twice the derivative of T_N
            assert np.abs(got_ref - 2.0 * exact).max() < 1e-11 * big
    # second derivative of T_N with the exact operator
    t = npcheb.chebval(xi, np.eye(n)[nz])
    c = np.zeros(n)
    c[nz] = 1.0
    exact2 = npcheb.chebval(xi, npcheb.chebder(c, 2)) * scale ** 2 if mapping != 2 else None
    if exact2 is not None:
        assert np.abs(o_ex.dct_diff(t, 2) - exact2).max() < 1e-10 * np.abs(exact2).max()
    # with xl = 1 the two operators are the same
    if o_ref.xl == 1.0:
        y = np.random.default_rng(3).standard_normal((n, 3))
        assert np.array_equal(o_ref.dct_diff(y, 1), o_ex.dct_diff(y, 1))


# ------------------------------------------------------------ step map (oracle)
NU, DT = 4.0e-6, 5.0e-4
NZ = 36
KX_IDX, KY_IDX = 1, 1


def _oracle(dct_exact, top_bc, bctype):
    cfg = OracleConfig(nx=8, ny=4, nz=NZ, mapping=3, zlen=2.0, xnu=NU, dt=DT, bctype=bctype, samp=0.0,
                       dealiasflag=0, dct_exact=dct_exact, top_bc=top_bc)
    return Oracle(cfg)       # ubar = 0: the step is linear in a small perturbation


def oracle_step_map(o):
    """Real matrix of one steady BDF3 step on pencil (KX_IDX, KY_IDX): state
    (u, v, w at n, n-1, n-2; p at n, n-1) as real and imaginary parts."""
    nzp, ny, nx = o.nzp, o.ny, o.nx
    nblk = 11
    dim = 2 * nblk * nzp
    eps = 1e-100
    M = np.zeros((dim, dim))

    def put(arr, vec):
        arr[:, KY_IDX, 2 * KX_IDX] = vec[:nzp]
        arr[:, KY_IDX, 2 * KX_IDX + 1] = vec[nzp:]

    def get(arr):
        return np.concatenate([arr[:, KY_IDX, 2 * KX_IDX], arr[:, KY_IDX, 2 * KX_IDX + 1]])

    shape = (nzp, ny, nx)
    for col in range(dim):
        x = np.zeros(dim)
        x[col] = eps
        blocks = x.reshape(nblk, 2 * nzp)
        f = [np.zeros(shape) for _ in range(nblk)]
        for b in range(nblk):
            put(f[b], blocks[b])
        st = o.new_state(f[0], f[1], f[2])
        st["unif"] = np.stack(f[3:6])
        st["unmif"] = np.stack(f[6:9])
        st["pf"], st["pnf"] = f[9], f[10]
        st["istep"] = 10
        st["t"] = 1.0
        st["abfac"] = o.abcoef(0, 10)
        st["bdf"], st["bdfv"] = o.bdfcoef(0, 10, DT, DT, DT)
        o.navierstokes_update(st)
        out = [st["uf"], st["vf"], st["wf"], *st["unif"], *st["unmif"], st["pf"], st["pnf"]]
        M[:, col] = np.concatenate([get(a) for a in out]) / eps
    return M


# ------------------------------------------------- independent dense restatement
def _cgl(nz, zlen):
    """Gauss-Lobatto points (top first) and the exact first-derivative matrix in z."""
    n = nz
    x = np.cos(np.pi * np.arange(n + 1) / n)
    c = np.ones(n + 1)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(n + 1)
    dX = x[:, None] - x[None, :]
    D = np.outer(c, 1.0 / c) / (dX + np.eye(n + 1))
    D = D - np.diag(D.sum(axis=1))
    return D * (2.0 / zlen)


def dense_step_map(d1, kx, ky, top):
    """navierstokes.f90:1282-1488 for one pencil, steady BDF3, zero nonlinear term and
    zero boundary data, Helmholtz/Poisson problems solved directly with Dirichlet rows;
    d1 is the first-derivative operator of dfdz/d2fdz2, D^2 = (exact D)^2 as helmholtz."""
    n = NZ + 1
    De = _cgl(NZ, 2.0)
    D2 = De @ De
    I = np.eye(n)
    k2 = kx * kx + ky * ky
    bdf = (3.0, -1.5, 1.0 / 3.0)
    bdfv = 6.0 / 11.0
    s0 = 1.0 / (NU * DT)
    sig = s0 / bdfv

    def dsolve(A, rhs, top_v, bot_v):
        A = A.astype(complex).copy()
        r = rhs.astype(complex).copy()
        A[0, :] = 0.0
        A[0, 0] = 1.0
        r[0] = top_v
        A[-1, :] = 0.0
        A[-1, -1] = 1.0
        r[-1] = bot_v
        return np.linalg.solve(A, r)

    H = D2 - (k2 + sig) * I
    P = D2 - k2 * I
    dim = 11 * n
    M = np.zeros((dim, dim), complex)
    for col in range(dim):
        x = np.zeros(dim, complex)
        x[col] = 1.0
        U = [x[i * n:(i + 1) * n] for i in range(11)]
        un, un1, un2, p0, p1 = U[0:3], U[3:6], U[6:9], U[9], U[10]
        rhs = [-s0 * (bdf[0] * un[c] + bdf[1] * un1[c] + bdf[2] * un2[c]) for c in range(3)]
        ps = 2.0 * p0 - p1
        bu, bv = DT * bdfv * 1j * kx * ps[-1], DT * bdfv * 1j * ky * ps[-1]
        if top == 0:
            tu, tv = -rhs[0][0] / (k2 + sig), -rhs[1][0] / (k2 + sig)
        else:
            tu, tv = DT * bdfv * 1j * kx * ps[0], DT * bdfv * 1j * ky * ps[0]
        tw = NU * DT * bdfv * ((d1 @ d1) @ un[2])[0]
        us, vs, ws = dsolve(H, rhs[0], tu, bu), dsolve(H, rhs[1], tv, bv), dsolve(H, rhs[2], tw, 0.0)
        divh = 1j * kx * us + 1j * ky * vs
        w = dsolve(P, -d1 @ divh - k2 * ws, 0.0, 0.0)
        p = -(d1 @ w + divh) / (k2 * DT * bdfv)
        u, v = us - DT * bdfv * 1j * kx * p, vs - DT * bdfv * 1j * ky * p
        M[:, col] = np.concatenate([u, v, w, *un, *un1, p, p0])
    return M


def _radius(M):
    return float(np.max(np.abs(np.linalg.eigvals(M))))


def test_step_map_grows_with_reference_operator_and_boundary_data():
    """Reference operator, reference ceiling, bctype > 0: the oracle's step map and the
    dense restatement have the same growing mode."""
    o = _oracle(0, 0, bctype=3)
    rho_oracle = _radius(oracle_step_map(o))
    kx, ky = o.wavx[2 * KX_IDX], o.wavy[KY_IDX]
    d_ref = o.dct_diff(np.eye(o.nzp), 1)
    rho_dense = _radius(dense_step_map(d_ref, kx, ky, 0))
    rho_dirichlet = _radius(dense_step_map(d_ref, kx, ky, 1))
    print(f"spectral radius oracle {rho_oracle:.9f}, dense restatement {rho_dense:.9f}, "
          f"dense with a Dirichlet ceiling {rho_dirichlet:.6f}")
    assert rho_oracle > 1.2 and rho_dirichlet > 1.2
    assert abs(rho_oracle - rho_dense) < 1e-6 * rho_dense


def test_step_map_decays_with_exact_operator():
    o = _oracle(1, 0, bctype=3)
    rho_oracle = _radius(oracle_step_map(o))
    kx, ky = o.wavx[2 * KX_IDX], o.wavy[KY_IDX]
    rho_dense = _radius(dense_step_map(_cgl(NZ, 2.0), kx, ky, 0))
    rho_dirichlet = _radius(dense_step_map(_cgl(NZ, 2.0), kx, ky, 1))
    print(f"spectral radius oracle {rho_oracle:.12f}, dense restatement {rho_dense:.12f}, "
          f"dense with a Dirichlet ceiling {rho_dirichlet:.12f}")
    assert rho_oracle < 1.0 + 1e-9 and rho_dense < 1.0 + 1e-9 and rho_dirichlet < 1.0 + 1e-9


def test_step_map_homogeneous_path_decays():
    """bctype 0 with the reference operator: no boundary data, no growing mode."""
    rho = _radius(oracle_step_map(_oracle(0, 0, bctype=0)))
    print(f"bctype 0, reference operator: spectral radius {rho:.12f}")
    assert rho < 1.0 + 1e-9
