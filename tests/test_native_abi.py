"""The C-ABI library loads, exports every entry point of include/lsbflow.h and its
host-only setup routines agree with the oracle (no GPU calls)."""

import ctypes
import os
import re

import numpy as np

from oracle import Oracle, OracleConfig
import lsbflow_amd as L
from lsbflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lsbflow.h")).read()
    return sorted(set(re.findall(r"\b(lsb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 40
    for name in decl:
        assert hasattr(lib, name), name
    # the binding declares exactly the header's entry points
    assert sorted(_lib.exported_symbols()) == decl
    _lib.load()


def test_config_struct_layout():
    """lsb_config as seen by ctypes has the header's field order (17 ints, 4 bytes of
    alignment padding, then 17 doubles)."""
    names = [n for n, _ in _lib.LsbConfig._fields_]
    txt = open(os.path.join(ROOT, "include", "lsbflow.h")).read()
    body = txt.split("typedef struct lsb_config {")[1].split("} lsb_config;")[0]
    hdr = re.findall(r"\b([a-z_0-9]+)\s*[,;]", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert names == hdr
    assert ctypes.sizeof(_lib.LsbConfig) == 18 * 4 + 17 * 8


def test_host_setup_matches_oracle():
    o = Oracle(OracleConfig(nx=8, ny=4, nz=48, mapping=3, zlen=2.0))
    dor, d, d2 = L.build_cheb_zder_matrix(o.nzp, o.xl, o.gf)
    assert np.array_equal(dor, o.dor)
    assert np.max(np.abs(d2 - o.d2)) < 1e-14 * np.max(np.abs(o.d2))
    y = np.random.default_rng(0).standard_normal((o.nzp, 5))
    for p in (1, 2):
        assert np.max(np.abs(L.dct_diff(y, o.xl, o.gf, p) - o.dct_diff(y, p))) < 1e-12 * np.max(np.abs(o.dct_diff(y, p)))
    lam, e, ei = L.diagonalize_matrix(o.d2[1:o.nz, 1:o.nz])
    D = o.d2[1:o.nz, 1:o.nz]
    assert np.max(np.abs(e @ np.diag(lam) @ ei - D)) < 1e-13 * np.max(np.abs(D))
    assert np.max(np.abs(np.sort(lam) - np.sort(o.diag)) / np.abs(np.sort(o.diag))) < 1e-10


def test_diagonalize_rejects_complex_spectrum():
    rot = np.array([[0.0, -1.0], [1.0, 0.0]])
    try:
        L.diagonalize_matrix(rot)
    except _lib.LsbError as e:
        assert "complex" in str(e)
    else:
        raise AssertionError("complex eigenvalues must be rejected")
