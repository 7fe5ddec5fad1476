"""bench.py's host logic on the CPU: the mixed-vs-FP64 comparison of two states, and the
single-GPU refusal of --gpus above 1 (before anything touches a GPU)."""

import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_compare_forms_relative_errors_and_separation():
    rng = np.random.default_rng(3)
    ref = {k: rng.standard_normal((5, 4, 8)) for k in ("uf", "vf", "wf", "pf")}
    scale = {"uf": 1e-6, "vf": 2e-7, "wf": 0.0, "pf": 3e-5}
    mixed = {}
    for k, a in ref.items():
        d = rng.standard_normal(a.shape)
        mixed[k] = a + scale[k] * np.linalg.norm(a) / np.linalg.norm(d) * d
    x = np.linspace(0.0, 1.0, 16)
    cf0 = np.cos(2 * np.pi * x)                      # + to -, then - to +
    cf1 = cf0 * (1 + 1e-6)
    wall = {0: (cf0, 0.25, 0.75), 1: (cf1, 0.25 + 3e-9, np.nan)}
    errs, sep, finite = bench.compare_forms({0: ref, 1: mixed}, wall)
    for k in ("uf", "vf", "wf", "pf"):
        assert abs(errs[k] - scale[k]) <= 1e-3 * scale[k] + 1e-18, (k, errs[k])
    assert abs(errs["wall_shear"] - 1e-6) < 1e-9
    assert sep["x_sep_fp64"] == 0.25 and abs(sep["x_sep_abs_diff"] - 3e-9) < 1e-15
    assert sep["x_reatt_abs_diff"] is None           # no reattachment in one form: JSON null
    assert finite
    mixed["pf"] = mixed["pf"].copy()
    mixed["pf"][0, 0, 0] = np.nan
    assert not bench.compare_forms({0: ref, 1: mixed}, wall)[2]


def test_several_gpus_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
