"""lsbflow_amd -- MI355X (gfx950) implementation of lsbflow's per-time-step hot path.

The numerical work runs in ``liblsbflow_amd.so`` (hand-written HIP kernels behind
the C-ABI of ``include/lsbflow.h``); this package is the host-side mirror of the
reference's module interfaces (``core``), the configurations of the
benchmarks (``cases``) and the build recipe (``build``).
"""

from .core import Config, Context, build_cheb_zder_matrix, dct_diff, diagonalize_matrix, setup_case, solve_orrsomm, sponge_lambda  # noqa: F401,E501
from . import cases  # noqa: F401

__all__ = ["Config", "Context", "build_cheb_zder_matrix", "dct_diff", "diagonalize_matrix", "setup_case",
           "solve_orrsomm", "sponge_lambda", "cases"]
