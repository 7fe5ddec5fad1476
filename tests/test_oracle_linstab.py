"""Pin the Orr-Sommerfeld restatement (orrsomm.f90:110-562) and the library's host solver
to the reference's linear-stability sweep (run/output_lintest.txt, run/energy.dat; fixtures
in tests/golden/reference_run/linstab.json).

The sweep's grid is not in the repository; nx = 64, nzp = 65 are recovered from the log,
and mapping 2, zlen = 33.3333, blfac = 8 (run/input.dat's values) reproduce the logged
Orr-Sommerfeld eigenvalues to about 1e-8 from the reference's own starting guess
alpha*cguess*u0 (the iglb = 0 iteration; the reference's global zggevx step first picks a
near-singular eigenvalue of the membrane rows and its iteration then lands on the same
mode)."""

import json
import math
import os

import numpy as np
import pytest

from oracle import Oracle, OracleConfig
import lsbflow_amd as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_run")
FIX = json.load(open(os.path.join(GOLD, "linstab.json")))
XNU, U0 = 1.0e-3, 1.0


def lintest_case(re_d, astar, ny=4):
    """init.f90:147-149 and set_linstab_param (orrsomm.f90:15-31)."""
    x0 = re_d ** 2 * XNU / (6.02 ** 2 * U0)
    xlen = 2 * math.pi * 6.02 * math.sqrt(XNU * x0 / U0) / astar
    o = Oracle(OracleConfig(nx=64, ny=ny, nz=64, xlen=xlen, ylen=xlen, zlen=33.3333, x0=x0, u0=U0, xnu=XNU,
                            mapping=2, blfac=8.0, bctype=0))
    o.setup_blasius()
    return o


# (Re_d, alpha*) whose logged value the restatement meets only loosely: 2555/0.5 went through
# 7 random restarts in the reference's run (run/output_lintest.txt), and the long-wave
# alpha* = 0.5 modes at higher Re (and 4805/0.675), near neutral, move by 1e-7..1e-5 with
# details of the sweep's grid that the repository does not hold.
LOOSE = {(2555.0, 0.5), (3305.0, 0.5), (4055.0, 0.5), (4805.0, 0.5), (4805.0, 0.675)}


def cguess(re_d):
    return 0.4 - (re_d - 1805.0) / 750.0 * 0.1 / 4


@pytest.mark.parametrize("case", FIX["orrsomm_omega"], ids=lambda c: f"{c['re_d']:.0f}-{c['alphastar']}")
def test_orrsomm_omega_output_lintest(case):
    """The 10-digit omega of run/output_lintest.txt (orrsomm.f90:443) from the oracle and from
    the library's host solver.  The iteration stops once the eigenvector changes by less than
    5e-5 in L1 (orrsomm.f90:394), which fixes omega, a two-sided Rayleigh quotient, to a few
    1e-9 (the rounding of the two LU factorisations decides where it stops) and the vector to
    about 1e-4."""
    o = lintest_case(case["re_d"], case["alphastar"])
    al = o.alpha
    g = al * case["cguess"] * U0
    u, v, w, om = o.solve_orrsomm(1e-10, al, 0.0, g, 0)
    want = complex(*case["omega"])
    tol = 5e-7 if (case["re_d"], case["alphastar"]) in LOOSE else 3e-8
    assert abs(om - want) < tol, (om, want)
    lu, lv, lw, lom = L.solve_orrsomm(o.d, o.d2, o.ubar, o.ubarp, o.ubarpp, o.zpts, XNU, 1e-10, al, 0.0, g, 0)
    assert abs(lom - want) < tol, (lom, want)
    assert abs(lom - om) < 1e-8
    # the mode's phase is whatever the iteration left (L1 normalisation only, orrsomm.f90:386-387)
    ph = np.vdot(u, lu)
    ph = ph / abs(ph)
    for a, b in ((lu, u), (lw, w)):
        assert np.max(np.abs(a - ph * b)) < 1e-3 * np.max(np.abs(b))
    assert not np.any(lv)


@pytest.mark.parametrize("row", FIX["energy_dat"], ids=lambda r: f"{r['re_d']:.0f}-{r['alphastar']}")
def test_orrsomm_omega_i_energy_dat(row):
    """omega_i of every (Re_d, alpha*) of run/energy.dat (printed to 6 digits)."""
    o = lintest_case(row["re_d"], row["alphastar"])
    al = o.alpha
    *_, om = o.solve_orrsomm(1e-10, al, 0.0, al * cguess(row["re_d"]) * U0, 0)
    # G14.6: half a unit of the sixth significant digit, plus 1.5e-7 for where the iteration
    # stops (a stationary Rayleigh quotient whose mode is converged to 5e-5 in L1 leaves omega
    # within ~1e-7 of the point either run reached)
    v = row["omega_i"]
    half_unit = 0.5e-6 * 10.0 ** (math.floor(math.log10(abs(v))) + 1)
    tol = 0.15 * abs(v) if (row["re_d"], row["alphastar"]) in LOOSE else half_unit + 1.5e-7
    assert abs(om.imag - v) < tol, (om, row)


def test_orrsomm_mode_satisfies_the_pencil():
    """The returned pair solves (A - omega B) phi = 0 to round-off and carries energy eps."""
    o = lintest_case(1805.0, 0.5)
    al = o.alpha
    u, v, w, om = o.solve_orrsomm(1e-10, al, 0.0, al * 0.4, 0)
    a, b = o.orrsomm_matrices(al, 0.0)
    # phi from w = -i alpha phi on the nzp rows; the two extra unknowns from the least squares
    phi = np.zeros(a.shape[0], dtype=complex)
    phi[: o.nzp] = w / (-1j * al)
    M = a - om * b
    rest = np.linalg.lstsq(M[:, o.nzp:], -M[:, : o.nzp] @ phi[: o.nzp], rcond=None)[0]
    phi[o.nzp:] = rest
    assert np.linalg.norm(M @ phi) < 1e-8 * np.linalg.norm(a) * np.linalg.norm(phi)
    def energy(q):
        e = 0.0
        for k in range(1, o.nzp - 1):
            e -= 0.5 * (abs(q[k + 1]) ** 2 + abs(q[k]) ** 2) * (o.zpts[k + 1] - o.zpts[k])
        return e * 2 / 7
    assert abs(energy(u) + energy(v) + energy(w) - 1e-10) < 1e-22
