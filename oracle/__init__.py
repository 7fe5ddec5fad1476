"""CPU oracle for the lsbflow time-step hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy FP64, the algorithm of the reference
Fortran routines the HIP library replaces (horfft.f90, derivatives.f90,
chebyshev.f90, helmholtz.f90, sgsterm.f90, filters.f90, dealiasing.f90 and the
step driver navierstokes.f90).  It exists only to check the product:

* only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
  ``cpu_baseline`` leg may import it, and only as the checker / the timed CPU
  baseline -- never as the thing measured or shipped;
* the product package ``lsbflow_amd`` never imports it and has no CPU
  fallback.

The reference itself cannot be built in this image (it is Fortran and needs
FFTW3, LAPACK/BLAS and MPI, none of which are installed), so there is no
``oracle/_ref``.  FFTW's r2c/c2r/c2c/REDFT00 transforms are restated with
numpy/scipy's pocketfft and LAPACK's dggev/dgetri with numpy.linalg; both differ
from the originals by rounding only.

Each function's header says which reference fixture, closed form or invariant
pins it (see ``tests/test_oracle_*.py``).
"""

from .lsb_oracle import Oracle, OracleConfig  # noqa: F401
