"""Throughput of the device-resident lsbflow time step on one MI355X.

A "step" is one timestep() of the hot path (navierstokes_update + norm) over the
whole 512 x 193 x 256 (nx x nzp x ny) LES domain resident in HBM, started from
the seeded tanh-plus-perturbation initial condition with the ceiling
suction/blowing profile (lsbflow_amd.cases).  W untimed warm-up steps, then K
timed steps bracketed by device synchronisation; value = nx*nzp*ny*K / seconds.

The metric compares the two forms of the step: with --precision mixed (the
default) the FP64 form runs first from the same initial condition, timed the
same way, and the line reports the mixed form's throughput (value) and the FP64
form's.  The accuracy leg then integrates both forms for --accuracy-steps (100)
steps of the same workload (same ceiling, grid, model and time step) and
reports the relative L2 difference of the final states per field, of the wall
skin friction, and the separation and reattachment points.  A run whose state
is not finite prints a null value and exits with status 1.

Besides the metric the JSON line carries
  roofline      the dominant kernel class -- the largest device time in an untimed
                pass of K steps with every launch group bracketed by HIP events
                (kernel_classes_ms) -- measured live with HIP events on the launch
                stream around its launch groups during the timed steps (only those
                groups carry events there; algorithmic work per launch / device time),
                against the MI355X peak that bounds it;
  cpu_baseline  the CPU oracle (numpy/scipy FP64 restatement, 1 thread) timed on a
                bounded sample of the same workload (same nx, nzp and SGS model,
                spanwise extent reduced), in the same unit.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6        # MI355X FP64 vector = FP64 matrix
PEAK_FP32_TFLOPS = 157.3       # MI355X dense FP32 matrix


def gpu_run(case, steps, warmup, precision, ic, tune=True):
    import numpy as np
    from lsbflow_amd.core import setup_case

    ctx = setup_case(case, precision=precision)
    # launch shapes from the autotuner, timed on this grid before the warm-up (untimed)
    table = ctx.autotune() if tune else ctx.tune_table()
    ctx.state_set(*ic)
    if warmup > 0:
        ctx.timestep(warmup)
    ctx.synchronize()
    # an untimed pass of `steps` steps with every launch group bracketed: the per-class
    # device times and the dominant class
    ctx.profile_begin()
    ctx.timestep(steps)
    ctx.synchronize()
    prof_all = ctx.profile_end()
    dom = max(prof_all, key=lambda k: prof_all[k]["ms"])
    # the timed steps carry events only around the dominant class's launch groups
    ctx.profile_begin_classes([dom])
    t0 = time.perf_counter()
    ctx.timestep(steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    prof_dom = ctx.profile_end()
    g = ctx.state_get()
    finite = all(bool(np.all(np.isfinite(g[k]))) for k in ("uf", "vf", "wf", "pf"))
    ctx.close()
    return t1 - t0, (prof_all, prof_dom), finite, table


def accuracy_leg(case, steps, ic):
    """Relative L2 of the mixed form against the FP64 form after `steps` steps: the four
    state fields, the wall skin friction Cf(x) (spanwise mean), and the separation and
    reattachment points (first sign changes of Cf, reported as absolute differences in x)."""
    import numpy as np
    from lsbflow_amd.core import setup_case

    states, wall = {}, {}
    for prec in (0, 1):
        ctx = setup_case(case, precision=prec)
        ctx.state_set(*ic)
        ctx.timestep(steps)
        g = ctx.state_get()
        wall[prec] = ctx.wall_shear()
        ctx.close()
        states[prec] = g
    return compare_forms(states, wall)


def compare_forms(states, wall):
    """Per-field relative L2 of the mixed form's state (key 1) against the FP64 form's (key 0),
    the skin friction's, and the separation / reattachment differences."""
    import numpy as np

    def rel(a, b):
        nb = np.linalg.norm(b)
        return float(f"{np.linalg.norm(a - b) / nb:.3e}") if nb > 0 else float(np.linalg.norm(a - b))

    errs = {k: rel(states[1][k], states[0][k]) for k in ("uf", "vf", "wf", "pf")}
    errs["wall_shear"] = rel(wall[1][0], wall[0][0])
    xs = {p: (wall[p][1], wall[p][2]) for p in (0, 1)}
    finite = all(bool(np.all(np.isfinite(states[p][k]))) for p in (0, 1) for k in ("uf", "vf", "wf", "pf"))
    def num(x):            # no sign change of Cf: no such point (JSON null)
        return float(f"{x:.6g}") if np.isfinite(x) else None

    sep = {"x_sep_fp64": num(xs[0][0]), "x_reatt_fp64": num(xs[0][1]),
           "x_sep_abs_diff": num(abs(xs[1][0] - xs[0][0])), "x_reatt_abs_diff": num(abs(xs[1][1] - xs[0][1]))}
    return errs, sep, finite


COPY_BYTES = 1 << 30


def copy_bandwidth(reps=5, per=20):
    """The practical streaming ceiling of HBM on this GPU: a plain device-to-device copy
    (torch copy_) of 1 GiB and of 4 GiB, bytes read + written over its event-timed duration."""
    import torch

    out = {"copy": "torch copy_, device to device, float64", "unit": "TB/s (read + written)"}
    for size in (1 << 30, 1 << 32):
        src = torch.ones(size // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        rates = []
        for _ in range(reps):
            e0.record()
            for _ in range(per):
                dst.copy_(src)
            e1.record()
            torch.cuda.synchronize()
            rates.append(round(2.0 * size / (e0.elapsed_time(e1) / per * 1e-3) / 1e12, 3))
        out[f"{size >> 30}_GiB"] = rates
        del src, dst
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)
    return 0


def pmc_probe(case):
    """Known-size copy, then one step of each form (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)."""
    import torch
    import lsbflow_amd as L
    from lsbflow_amd.core import setup_case

    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    ic = L.cases.initial_condition(case)
    for prec in (0, 1):
        ctx = setup_case(case, precision=prec)
        ctx.state_set(*ic)
        ctx.timestep(2)
        ctx.synchronize()
        ctx.close()
    print(json.dumps({"pmc_probe": case.name, "copy_bytes": COPY_BYTES}), flush=True)
    return 0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c4_sigma_les_512x193x256_hbm_traffic.json")


def sources_digest():
    """SHA-256 over the library's kernel and host sources (the build the counters measured)."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "lsbflow_amd", "csrc")
    for f in sorted(os.listdir(d)):
        with open(os.path.join(d, f), "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def measured_traffic(case, cls, precision, p, steps):
    """HBM bytes per launch group of class `cls`, from the committed rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs (copy-calibrated), when they cover this workload and were
    measured on the current library sources (the table carries their digest)."""
    if case is None or case.name != "c4_sigma_les" or not os.path.exists(TRAFFIC_FILE) or p["n"] == 0:
        return None
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)
    if t.get("sources_sha") != sources_digest():
        return None
    d = t.get("classes", {}).get("mixed" if precision == 1 else "fp64", {}).get(cls)
    if d is None:
        return None
    return float(f"{d['bytes_per_step'] * steps / p['n']:.4g}")


def traffic_class(name):
    """Profiling class of a kernel of the step (the classes lsb_profile_end reports)."""
    if "k_yfwd_rhs" in name or "k_rhs_vel" in name:
        return "rhs"
    if "_fast<" in name:
        return "fft"
    if "k_zhelm32" in name:       # the mixed form's velocity solves: FP32 matrix work
        return "zcontract"
    if any(t in name for t in ("k_zcontract<1>", "k_zcontract<2>", "k_zsplit<1,", "k_zsplit<2,", "k_zhelm",
                               "k_zpress")):
        return "helmholtz"
    if "zcontract" in name or "zsplit" in name:
        return "zcontract"
    if "k_nonlinear" in name:
        return "nonlinear"
    if name.startswith("lsb::"):
        return "pointwise"
    return None


def pmc_summarize(fetch_csv, write_csv, out_path=TRAFFIC_FILE, steps_per_form=2):
    """HBM bytes per kernel and per class and step from the FETCH_SIZE and WRITE_SIZE runs of
    --pmc-probe: each counter scaled by what it reads for the 1 GiB copy; the FP64 form's
    kernels are those up to its last step's final k_norm_nyq, the mixed form's follow."""
    import csv

    def load(fn, counter):
        rows = [r for r in csv.DictReader(open(fn)) if r["Counter_Name"] == counter]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        return [(r["Kernel_Name"].split("(")[0].replace("void ", ""), float(r["Counter_Value"])) for r in rows]

    F, W = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    if len(F) != len(W) or any(a[0] != b[0] for a, b in zip(F, W)):
        raise RuntimeError("pmc_summarize: the two runs dispatched different kernels")
    ci = next(i for i, (n, v) in enumerate(F) if "copyBuffer" in n and v > 1e5)
    fetch_cal = COPY_BYTES / (F[ci][1] * 1024.0)
    write_cal = COPY_BYTES / (W[ci][1] * 1024.0)
    ends = [i for i, (n, _) in enumerate(F) if "k_norm_nyq" in n]
    last_fp64 = ends[steps_per_form - 1]
    out = {"sources_sha": sources_digest(),
           "calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": F[ci][1], "write_kib": W[ci][1],
                           "fetch_factor": fetch_cal, "write_factor": write_cal},
           "probe_steps_per_form": steps_per_form, "classes": {"fp64": {}, "mixed": {}}, "kernels": {}}
    for i, ((n, f), (_, w)) in enumerate(zip(F, W)):
        if i <= ci:
            continue
        form = "fp64" if i <= last_fp64 else "mixed"
        rb, wb = f * 1024 * fetch_cal, w * 1024 * write_cal
        c = traffic_class(n)
        if c:
            d = out["classes"][form].setdefault(c, {"read_bytes_per_step": 0.0, "write_bytes_per_step": 0.0})
            d["read_bytes_per_step"] += rb / steps_per_form
            d["write_bytes_per_step"] += wb / steps_per_form
        k = out["kernels"].setdefault(form + ":" + n, {"dispatches": 0, "read_bytes": 0.0, "write_bytes": 0.0})
        k["dispatches"] += 1
        k["read_bytes"] += rb
        k["write_bytes"] += wb
    for form in ("fp64", "mixed"):
        for d in out["classes"][form].values():
            d["bytes_per_step"] = d["read_bytes_per_step"] + d["write_bytes_per_step"]
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({"pmc_summary": out_path, "classes": {f: {c: f"{d['bytes_per_step']:.4g}" for c, d in
                                                                out["classes"][f].items()} for f in out["classes"]}}))
    return 0


def roofline(prof, precision, case=None, steps=1):
    """Dominant class by device time -> achieved vs peak (algorithmic work / measured time)."""
    cls = max(prof, key=lambda k: prof[k]["ms"])
    p = prof[cls]
    sec = p["ms"] * 1e-3
    traffic = measured_traffic(case, cls, precision, p, steps)
    if cls in ("zcontract", "helmholtz"):
        # the Helmholtz / Poisson projections of the FP64 form and the pressure step of both
        # forms run on the FP64 matrix cores; the mixed form's contraction class (the
        # derivative contractions of its FP32 chain and the FP32 correction of its velocity
        # solves) on the FP32 ones
        peak = PEAK_FP32_TFLOPS if (cls == "zcontract" and precision == 1) else PEAK_FP64_TFLOPS
        ach = p["flops"] / sec / 1e12 if sec > 0 else 0.0
        return dict(bound="mfma", kernel=cls, achieved=round(ach, 3), peak=peak, unit="TFLOP/s",
                    frac=round(ach / peak, 4), traffic=traffic, launches=p["n"], device_ms=round(p["ms"], 3))
    ach = p["bytes"] / sec / 1e9 if sec > 0 else 0.0
    return dict(bound="hbm", kernel=cls, achieved=round(ach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(ach / PEAK_HBM_GBS, 4), traffic=traffic,
                algorithmic_bytes_per_launch=float(f"{p['bytes'] / max(p['n'], 1):.4g}"),
                launches=p["n"], device_ms=round(p["ms"], 3))


def step_roofline(prof, precision, steps, ms_per_step):
    """Speed of light of the whole step: each class's algorithmic bytes at the HBM peak or
    its flops at the matrix peak of its precision, whichever takes longer, summed over the
    classes, against the measured time per step."""
    per, sol = {}, 0.0
    for cls, p in prof.items():
        if p["n"] == 0:
            continue
        t_hbm = p["bytes"] / (PEAK_HBM_GBS * 1e9)
        t_mfma = 0.0
        if cls in ("zcontract", "helmholtz"):
            peak = PEAK_FP32_TFLOPS if (cls == "zcontract" and precision == 1) else PEAK_FP64_TFLOPS
            t_mfma = p["flops"] / (peak * 1e12)
        t = max(t_hbm, t_mfma) / steps * 1e3
        sol += t
        per[cls] = {"sol_ms": round(t, 4), "device_ms": round(p["ms"] / steps, 4),
                    "bound": "mfma" if t_mfma > t_hbm else "hbm"}
    return {"sol_ms_per_step": round(sol, 4), "ms_per_step": round(ms_per_step, 4),
            "frac": round(sol / ms_per_step, 4) if ms_per_step > 0 else None, "classes": per}


def cpu_worker(case_name, sample_ny, sample_steps):
    """One oracle sample (one thread): prints its wall time as JSON.  Runs in a child process
    that never touches the GPU."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    import lsbflow_amd as L
    from oracle import Oracle, OracleConfig

    case = L.cases.CASES[case_name]
    sc = case.with_(ny=sample_ny)
    keys = OracleConfig.__dataclass_fields__
    with threadpool_limits(limits=1):
        o = Oracle(OracleConfig(**{k: v for k, v in sc.as_dict().items() if k in keys}))
        o.ubar, o.ubarp, o.ubarpp = L.cases.base_flow(sc)
        u, v, w = L.cases.initial_condition(sc)
        st = o.new_state(u, v, w)
        t0 = time.perf_counter()
        for _ in range(sample_steps):
            o.timestep(st)
        t1 = time.perf_counter()
    print(json.dumps({"seconds": t1 - t0, "points": sc.nx * sc.nzp * sc.ny,
                      "finite": bool(np.all(np.isfinite(st["uf"])))}), flush=True)
    return 0


def host_cores():
    """CPU threads this process may use, capped at 16 (one GPU's share of a GPU node)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(case, sample_ny, sample_steps):
    """The CPU oracle on every host core: one spanwise-reduced sample of the same workload per
    core, run concurrently in child processes (one thread each); aggregate grid-point-steps/s
    over the slowest worker's time."""
    import subprocess

    cores = host_cores()
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", case.name, str(sample_ny),
           str(sample_steps)]
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    t0 = time.perf_counter()
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
             for _ in range(cores)]
    res = []
    for p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"cpu worker failed: {err[-400:]}")
        res.append(json.loads(out.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    slowest = max(r["seconds"] for r in res)
    val = sum(r["points"] for r in res) * sample_steps / slowest
    return dict(value=float(f"{val:.4g}"), unit="grid-point-steps/s", cores=cores, kind="port",
                sample=f"{cores} concurrent processes, one per core, each {sample_steps} oracle steps of "
                       f"{case.nx}x{case.nzp}x{sample_ny} (nx x nzp x ny; spanwise extent {sample_ny}/{case.ny} "
                       f"of the workload), slowest {slowest:.1f} s, wall {wall:.1f} s, "
                       f"finite={all(r['finite'] for r in res)}")


def diagnostics(case, reps):
    """Wall time per call (host-synchronised; the records include their PCIe copy to
    host memory) of the per-step diagnostics, the I/O records and the two filters (mode-cut
    and TNS) on the FP64 resident state of `case`, after two steps; then one FP64 step
    without and with the on-the-fly time average (and its qavg record), and with the
    temperature / buoyancy path.  The kernels' own device times come from a
    rocprofv3 --kernel-trace run of this mode."""
    import numpy as np
    import lsbflow_amd as L
    ctx = L.setup_case(case, precision=0)
    ctx.state_set(*L.cases.initial_condition(case))
    ctx.timestep(2)
    ctx.synchronize()
    n = case.nx * case.ny * case.nzp
    snap = np.zeros(7 + 5 * n, dtype=np.float32)
    rst = np.zeros(9 + 20 * n, dtype=np.float32)
    calls = {
        "state_maxcalc": lambda: ctx.state_maxcalc(1.0, 0.0),
        "state_nancheck": ctx.state_nancheck,
        "state_wall_shear": lambda: ctx.wall_shear(),
        "state_snapshot_record": lambda: ctx._L.lsb_state_snapshot(ctx.h, 0, None, ctx._fp(snap)),
        "write_restart_record": lambda: ctx._L.lsb_write_restart(ctx.h, None, ctx._fp(rst)),
        "read_restart_record": lambda: ctx._L.lsb_read_restart(ctx.h, None, ctx._fp(rst)),
        "state_filterall3pt": lambda: (ctx.state_filterall3pt(), ctx.synchronize()),
        "state_tns_filter": ctx.state_tns_filter,
    }

    def step():
        ctx.timestep(1)
        ctx.synchronize()

    def temperature_on():
        z = ctx.grid()["zpts"]
        th = np.tanh(z / 0.2)
        ctx.set_temperature(1.0 + 0.1 * th, 0.5 * (1 - th ** 2), -2 * 0.5 / 0.2 * th * (1 - th ** 2), 1.0, 0.0, 0.0,
                            -1.0)
        ctx.state_set_temperature(0.5 * ctx.state_get()["uf"])

    calls["timestep_fp64"] = step
    calls["time_average_on"] = lambda: ctx.set_time_average(1, 0, 5)
    calls["timestep_fp64_with_time_average"] = step
    calls["state_snapshot_average_record"] = lambda: ctx._L.lsb_state_snapshot_average(ctx.h, None, ctx._fp(snap))
    calls["time_average_off"] = lambda: ctx.set_time_average(0, 0, 1)
    calls["temperature_on"] = temperature_on
    calls["timestep_fp64_with_temperature"] = step
    out = {}
    for name, fn in calls.items():
        if name.endswith("_on") or name.endswith("_off"):   # set-up, not timed
            fn()
            continue
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            rc = fn()
            if isinstance(rc, int) and rc != 0:
                raise RuntimeError(f"{name}: status {rc}")
        out[name] = round(1e3 * (time.perf_counter() - t0) / reps, 3)
        print(json.dumps({"diagnostics_step": name, "ms": out[name]}), flush=True)
    ctx.close()
    print(json.dumps({"diagnostics": case.name, "points": n, "reps": reps, "wall_ms_per_call": out}), flush=True)
    return 0


def c3_helmholtz(reps):
    """Batched Helmholtz solve alone: 256 x 256 pencils x 4 fields on 193 points (the
    512 x 193 x 256 spectral grid) per call, FP64 / FP32 / FP32 + one FP64 refinement / the
    FP64 leading term with the FP32 correction (the mixed form's velocity solve; the FP64
    solve at sigma = 0), for sigma = c / (nu dt), c in {1, 1.5}, and sigma = 0.  Device-resident arrays; wall time
    around `reps` calls with the stream synchronised on both sides."""
    import ctypes as C
    import numpy as np
    import lsbflow_amd as L
    from lsbflow_amd import _lib
    from lsbflow_amd.core import setup_case

    case = L.cases.CASES["c4_sigma_les"]
    ctx = setup_case(case)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    nf, nzp, P = 4, ctx.nzp, ctx.nx * ctx.ny
    rng = np.random.default_rng(20261203)
    n = np.arange(nzp)
    T = np.polynomial.chebyshev.chebvander(np.cos(np.pi * n / (nzp - 1)), nzp - 1)
    rhs = np.stack([T @ (rng.standard_normal((nzp, P)) * np.exp(-0.1 * n)[:, None]) for _ in range(nf)])
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), rhs.nbytes) == 0
    lib = _lib.lib()
    nzm = nzp - 2
    out = {"c3": "batched Helmholtz, 256x256 pencils x 4 fields x 193 points per call (nx=512, ny=256, nzp=193)",
           "systems_per_call": (P // 2) * nf, "reps": reps, "sigma": {}}
    try:
        for cval in (1.0, 1.5, 0.0):
            sigma = cval / (case.xnu * case.dt)
            row = {}
            for mode, name in ((0, "fp64"), (1, "fp32"), (2, "fp32_refined"), (3, "fp64_lead_fp32_corr")):
                assert hip.hipMemcpy(d, rhs.ctypes.data_as(C.c_void_p), rhs.nbytes, 1) == 0
                lib.lsb_solve_helmholtz_mixed_dev(ctx.h, d, sigma, nf, mode, None)     # warm-up
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    rc = lib.lsb_solve_helmholtz_mixed_dev(ctx.h, d, sigma, nf, mode, None)
                    if rc != 0:
                        raise RuntimeError(lib.lsb_last_error())
                ctx.synchronize()
                ms = 1e3 * (time.perf_counter() - t0) / reps
                flops = 2 * 2.0 * nzm * nzm * P * nf * (2 if mode == 2 else 1)   # dense count of the projections
                row[name] = {"ms_per_call": round(ms, 4), "systems_per_s": float(f"{(P // 2) * nf / (ms * 1e-3):.4g}"),
                             "dense_equiv_tflops": round(flops / (ms * 1e-3) / 1e12, 2)}
            out["sigma"][f"c={cval}"] = row
    finally:
        hip.hipFree(d)
        ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def dyn_companion(steps, warmup, accuracy_steps, tune):
    """The dynamic-Smagorinsky LES at the same 512 x 193 x 256 size (c4_dyn_les): both forms
    timed as the headline workload is, and the mixed form against the FP64 form after
    `accuracy_steps` steps from the same seed."""
    import lsbflow_amd as L
    case = L.cases.CASES["c4_dyn_les"]
    ic = L.cases.initial_condition(case)
    s64, _, f64, _ = gpu_run(case, steps, warmup, 0, ic, tune)
    s32, _, f32, _ = gpu_run(case, steps, warmup, 1, ic, tune)
    errs, sep, fin = accuracy_leg(case, accuracy_steps, ic)
    pts = case.points
    return {"workload": f"c4_dyn_les: {case.nx}x{case.nzp}x{case.ny}, dynamic Smagorinsky, dt={case.dt}",
            "value": float(f"{pts * steps / s32:.6g}"), "ms_per_step": round(1e3 * s32 / steps, 4),
            "fp64": {"value": float(f"{pts * steps / s64:.6g}"), "ms_per_step": round(1e3 * s64 / steps, 4)},
            "speedup_vs_fp64": round(s64 / s32, 3),
            "mixed_vs_fp64": {"steps": accuracy_steps, "rel_l2": errs, "separation": sep,
                              "all_within_1e-5": bool(max(errs.values()) <= 1e-5)},
            "finite": bool(f64 and f32 and fin)}


def c2_dynamic(steps, warmup):
    """C2: the plain FP64 dynamic-Smagorinsky LES, 256 x 129 x 128, Re 5e4, dt 5e-4, with its
    configured ceiling, `steps` timed steps after `warmup`; the state finite and the discrete
    divergence after the run relative to the velocity scale."""
    import numpy as np
    import lsbflow_amd as L
    from lsbflow_amd.core import setup_case
    case = L.cases.CASES["c2_dyn_les"]
    ic = L.cases.initial_condition(case)
    out = {"c2": f"{case.name}: {case.nx}x{case.nzp}x{case.ny} (nx x nzp x ny), dynamic Smagorinsky, "
                 f"Re 5e4, dt {case.dt}", "steps": steps, "warmup": warmup, "unit": "grid-point-steps/s"}
    states, wall = {}, {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        ctx = setup_case(case, precision=prec)
        ctx.autotune()
        ctx.state_set(*ic)
        ctx.timestep(warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.timestep(steps)
        ctx.synchronize()
        t1 = time.perf_counter()
        g = ctx.state_get()
        finite = all(bool(np.all(np.isfinite(g[k]))) for k in ("uf", "vf", "wf", "pf"))
        div = ctx.divcheck(g["uf"], g["vf"], g["wf"])
        scale = float(np.sqrt(np.mean(ctx.inv(g["uf"])[..., : ctx.nx] ** 2)))
        states[prec], wall[prec] = g, ctx.wall_shear()
        ctx.close()
        out[name] = {"value": float(f"{case.points * steps / (t1 - t0):.6g}"),
                     "ms_per_step": round(1e3 * (t1 - t0) / steps, 4), "finite_after_run": finite,
                     "mean_abs_divergence_over_u_rms": float(f"{div / max(scale, 1e-300):.3g}")}
    out["speedup_vs_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    # both forms took warmup + steps steps from the same seeded state (the launch shapes
    # the autotuner picks give the same bits)
    errs, sep, _ = compare_forms(states, wall)
    out["mixed_vs_fp64"] = {"steps": warmup + steps, "rel_l2": errs, "separation": sep,
                            "all_within_1e-5": all(v <= 1e-5 for v in errs.values())}
    print(json.dumps(out), flush=True)
    return 0 if out["fp64"]["finite_after_run"] and out["mixed"]["finite_after_run"] else 1


def c5_ensemble(steps, warmup, precision=1):
    """The uneven ensemble: 16 static-Smagorinsky LES members of different grids stepped
    together from captured graphs (one graph launch per step for all members), against
    the same members stepped eagerly one after another on their own streams.  Every member
    runs the mixed form (the FP64 form with --precision fp64); members whose nx/2 is not a
    register-resident FFT length (nx 96, 192) take the FP32 Stockham transforms."""
    import numpy as np
    import lsbflow_amd as L
    from lsbflow_amd.core import Ensemble, setup_case

    members = L.cases.ensemble_members()
    ctxs, prec = [], []
    for m in members:
        p = precision
        c = setup_case(m, precision=p)
        c.autotune()
        c.state_set(*L.cases.initial_condition(m))
        c.timestep(3)
        ctxs.append(c)
        prec.append(p)
    pts = sum(m.points for m in members)
    for c in ctxs:
        c.synchronize()
    # eager: every member's steps enqueued in turn, each on its own stream
    for _ in range(warmup):
        for c in ctxs:
            c.timestep(1)
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in ctxs:
            c.timestep(1)
    for c in ctxs:
        c.synchronize()
    t_eager = time.perf_counter() - t0
    ens = Ensemble(ctxs)
    try:
        ens.timestep(warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.timestep(steps)
        ens.synchronize()
        t_graph = time.perf_counter() - t0
        info = ens.info()
        finite = all(bool(np.isfinite(c.state_get()["uf"]).all()) for c in ctxs)
    finally:
        ens.close()
        for c in ctxs:
            c.close()
    out = {"c5": "uneven ensemble of 16 static-Smagorinsky LES members (Cs 0.1), one graph per step",
           "members": [f"{m.nx}x{m.nzp}x{m.ny} Re {0.2 / m.xnu:.3g} samp {m.samp:.3f} "
                       f"{'mixed' if p else 'fp64'}" for m, p in zip(members, prec)],
           "points": pts, "steps": steps, "graphs": info["graphs"],
           "graph": {"value": float(f"{pts * steps / t_graph:.4g}"), "ms_per_step": round(1e3 * t_graph / steps, 4)},
           "eager": {"value": float(f"{pts * steps / t_eager:.4g}"), "ms_per_step": round(1e3 * t_eager / steps, 4)},
           "unit": "grid-point-steps/s (all members)", "finite_after_run": finite}
    print(json.dumps(out), flush=True)
    return 0 if finite else 1


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="c4_sigma_les")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="built-in launch shapes (no lsb_autotune)")
    ap.add_argument("--c5", action="store_true",
                    help="time the uneven 16-member ensemble stepped from captured graphs (and eagerly)")
    ap.add_argument("--c2", action="store_true",
                    help="the 256x129x128 dynamic LES (C2) line instead of the headline")
    ap.add_argument("--no-dyn", action="store_true",
                    help="skip the 512x193x256 dynamic-Smagorinsky companion of the headline line")
    ap.add_argument("--c3", action="store_true",
                    help="time the batched Helmholtz solve alone (FP64, FP32, FP32 + one FP64 refinement)")
    ap.add_argument("--accuracy-steps", type=int, default=100)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="the workload of the HBM-traffic counter runs: a 1 GiB device copy (calibration), "
                         "then one step of each form after one warm-up step")
    ap.add_argument("--diagnostics", action="store_true",
                    help="time the snapshot / restart records and the per-step diagnostics (maxcalc, NaN "
                         "scan, wall shear) on the resident state of the workload instead of the step")
    ap.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the HBM-traffic table bench.py reads (profiles/..._hbm_traffic.json) from the "
                         "counter_collection.csv files of the FETCH_SIZE and WRITE_SIZE runs of --pmc-probe")
    ap.add_argument("--copy-bw", action="store_true",
                    help="the bandwidth of a plain 1 GiB and 4 GiB device copy (the practical HBM ceiling)")
    ap.add_argument("--cpu-sample-ny", type=int, default=16)
    ap.add_argument("--cpu-sample-steps", type=int, default=3)
    ap.add_argument("--cpu-worker", nargs=3, metavar=("CASE", "NY", "STEPS"), help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.cpu_worker:
        return cpu_worker(args.cpu_worker[0], int(args.cpu_worker[1]), int(args.cpu_worker[2]))
    if args.pmc_summarize:
        return pmc_summarize(*args.pmc_summarize)
    if args.gpus != 1:
        print("bench.py: the library runs the whole domain on one GPU; runs on several GPUs are not "
              "supported (use --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_bw:
        return copy_bandwidth()
    import numpy as np
    import lsbflow_amd as L
    case = L.cases.CASES[args.case]
    if args.pmc_probe:
        return pmc_probe(case)
    if args.diagnostics:
        return diagnostics(case, args.steps)
    if args.c2:
        return c2_dynamic(100 if args.steps == 10 else args.steps, args.warmup)
    if args.c3:
        return c3_helmholtz(args.steps)
    if args.c5:
        return c5_ensemble(args.steps, args.warmup, 1 if args.precision == "mixed" else 0)
    ic = L.cases.initial_condition(case)
    pts = case.points
    tune = not args.no_autotune
    secs64, (prof64, dom64), finite64, tune64 = gpu_run(case, args.steps, args.warmup, 0, ic, tune)
    fp64 = {"value": float(f"{pts * args.steps / secs64:.6g}"), "ms_per_step": round(1e3 * secs64 / args.steps, 4),
            "kernel_classes_ms": {k: round(v["ms"], 3) for k, v in prof64.items()}}
    table = tune64
    if args.precision == "mixed":
        secs, (prof, dom), finite, table = gpu_run(case, args.steps, args.warmup, 1, ic, tune)
        dtype = ("mixed: f32 nonlinear chain (FFTs, products/SGS storage, its two Chebyshev contractions on "
                 "f32 MFMA), f64 state, right-hand sides and pressure step; velocity solves as an f64 leading "
                 "term (-rhs/(k^2+sigma)) plus an f32-MFMA eigen-projection correction")
        precision = 1
        acc_errs, acc_sep, acc_finite = accuracy_leg(case, args.accuracy_steps, ic)
        accuracy = {"workload": f"{case.name}, the timed workload, {args.accuracy_steps} steps from the same seed",
                    "rel_l2": acc_errs, "separation": acc_sep, "finite": acc_finite,
                    "fields_within_1e-5": {k: bool(v <= 1e-5) for k, v in acc_errs.items()},
                    "all_within_1e-5": bool(max(acc_errs.values()) <= 1e-5)}
    else:
        secs, prof, dom, finite = secs64, prof64, dom64, finite64
        dtype = "f64"
        precision = 0
        accuracy = None
    value = pts * args.steps / secs
    out = {
        "metric": "grid-point-steps/s @512x193x256 LES, mixed vs FP64 GPU at rel-L2<=1e-5",
        "value": float(f"{value:.6g}"),
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * secs / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: seeded tanh base flow + curl of a random vector potential (rms 1e-3), "
                "ceiling suction/blowing profile (amplitude 0.3)",
        "config": {"workload": f"{case.name}: {case.nx}x{case.nzp}x{case.ny} (nx x nzp x ny), sgsmodel="
                               f"{case.sgsmodel if case.sgsflag else 'off'}, dt={case.dt}, xnu={case.xnu:.3g}",
                   "nx": case.nx, "nzp": case.nzp, "ny": case.ny, "precision": args.precision,
                   "ceiling": f"bctype {case.bctype}, top_bc {case.top_bc}, dct_exact {case.dct_exact}",
                   "finite_after_run": finite},
        "fp64": fp64,
        "speedup_vs_fp64": round(secs64 / secs, 3),
        "mixed_vs_fp64": accuracy,
        # the dominant class of the untimed all-class pass, measured again in the timed steps
        "roofline": roofline(dom, precision, case, args.steps),
        "step_roofline": step_roofline(prof, precision, args.steps, 1e3 * secs / args.steps),
        "kernel_classes_ms": {k: round(v["ms"], 3) for k, v in prof.items()},
        "autotune": table,
    }
    if args.precision == "mixed" and case.name == "c4_sigma_les" and not args.no_dyn:
        out["c4_dyn_les"] = dyn_companion(args.steps, args.warmup, args.accuracy_steps, tune)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_sample_ny, args.cpu_sample_steps)
    ok = finite and finite64 and (accuracy is None or accuracy["finite"])
    if "c4_dyn_les" in out:
        ok = ok and out["c4_dyn_les"]["finite"]
    if not ok:
        out["value"] = None
        out["error"] = "non-finite state after the timed steps or the accuracy leg"
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
