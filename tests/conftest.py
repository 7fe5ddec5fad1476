import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblsbflow_amd.so")


def rel_l2(a, b):
    """||a - b|| / ||b|| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    na = np.linalg.norm((a - b).ravel())
    if nb == 0.0:
        return na
    return na / nb


@pytest.fixture
def rng():
    return np.random.default_rng(20261202)


NOISE_KEYS = ("uf", "vf", "wf", "pf", "pnf", "unif", "unmif", "nlnif", "nlnmif",
              "tempf", "tnf", "tnmf", "nltnf", "nltnmf")


def rounding_noise(st, rng, eps=1e-16):
    """Multiply every prognostic array of an oracle state by (1 + eps N(0, 1)): noise of the
    size that a different but equivalent order of floating-point operations leaves in each
    step.  Applied after every step of a second oracle run, it measures how far two correct
    FP64 implementations of the same step can drift apart on the case."""
    for k in NOISE_KEYS:
        if isinstance(st.get(k), np.ndarray):
            st[k] = st[k] * (1.0 + eps * rng.standard_normal(st[k].shape))


NOISY_OPS = ("horfft_fwd", "horfft_inv", "dfdx", "dfdy", "dfdz", "d2fdz2", "dpdz", "divuv",
             "solve_helmholtz", "solve_nhom_helmholtz")


def _noisy(f, rng, eps):
    def g(*a, **k):
        r = f(*a, **k)
        if isinstance(r, np.ndarray) and r.dtype.kind == "f":
            return r * (1.0 + eps * rng.standard_normal(r.shape))
        return r
    return g


def wrap_noisy_ops(o, rng, eps=1e-16):
    """Give an oracle instance 1e-16 relative noise on the output of every transform,
    derivative, divergence and Helmholtz solve (the rounding sites a different
    implementation orders differently).  The instance is modified: use one made for it."""
    for name in NOISY_OPS:
        if hasattr(o, name):
            setattr(o, name, _noisy(getattr(o, name), rng, eps))
    return o


def noisy_oracle_run(o, fields, nsteps, seed=99, eps=1e-16, setup=None):
    """The oracle from ``fields`` for nsteps with wrap_noisy_ops and rounding_noise on the
    state after every step."""
    rng = np.random.default_rng(seed)
    wrap_noisy_ops(o, rng, eps)
    st = o.new_state(*fields)
    if setup is not None:
        setup(st)
    for _ in range(nsteps):
        o.timestep(st)
        rounding_noise(st, rng, eps)
    return st
