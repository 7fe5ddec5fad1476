"""lsb_horfft_dev_noalloc against lsb_horfft_dev through the C-ABI (device arrays).

Device memory for the caller's arrays comes from the HIP runtime through ctypes.  The
allocation-free entry point runs the same kernels as lsb_horfft_dev with the staging
buffer taken from the library's scratch array, so the two agree bit for bit.  Two calls
in a row (three fields, then one) check that the second call, on the array the first one
left behind, is unaffected; a context whose single field needs more staging than the
array holds is refused on the host before any launch.
"""

import ctypes as C

import numpy as np
import pytest

import lsbflow_amd as L
from lsbflow_amd import _lib
from lsbflow_amd.core import setup_case

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, arr):
        a = np.ascontiguousarray(arr)
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, src, shape):
        a = np.empty(shape)
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, D2H) == 0
        return a

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.lib.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value


def test_noalloc_matches_horfft_dev_two_calls():
    hip = Hip()
    lib = _lib.lib()
    case = L.cases.CASES["c1_dns"]
    ctx = setup_case(case)
    nzp, ny, nx = ctx.nzp, ctx.ny, ctx.nx
    phys_shape, spec_shape = (nzp, ny, nx + 2), (nzp, ny, nx)
    rng = np.random.default_rng(7)
    bufs = []
    try:
        for nf in (3, 1):          # the second call with a smaller input
            q = rng.standard_normal((nf,) + phys_shape)
            q[..., nx:] = 0.0
            pa, pb = hip.alloc(q.nbytes), hip.alloc(q.nbytes)
            sa, sb = hip.alloc(nf * nzp * ny * nx * 8), hip.alloc(nf * nzp * ny * nx * 8)
            bufs += [pa, pb, sa, sb]
            hip.put(pa, q)
            hip.put(pb, q)
            # forward (no staging) and inverse (staging in the scratch array)
            assert lib.lsb_horfft_dev(ctx.h, pa, sa, -1, nf, None) == 0
            assert lib.lsb_horfft_dev_noalloc(ctx.h, pb, sb, -1, nf, None) == 0
            assert lib.lsb_synchronize(ctx.h) == 0
            fa, fb = hip.get(sa, (nf,) + spec_shape), hip.get(sb, (nf,) + spec_shape)
            assert np.array_equal(fa, fb), f"forward, {nf} fields"
            assert lib.lsb_horfft_dev(ctx.h, pa, sa, 1, nf, None) == 0
            assert lib.lsb_horfft_dev_noalloc(ctx.h, pb, sb, 1, nf, None) == 0
            assert lib.lsb_synchronize(ctx.h) == 0
            ia, ib = hip.get(pa, (nf,) + phys_shape), hip.get(pb, (nf,) + phys_shape)
            assert np.array_equal(ia[..., :nx], ib[..., :nx]), f"inverse, {nf} fields"
            # a second round trip returns the first one's result (the transform pair drops the
            # kx-Nyquist mode and the ky-Nyquist row, so it is a projection: idempotent)
            assert lib.lsb_horfft_dev_noalloc(ctx.h, pb, sb, -1, nf, None) == 0
            assert lib.lsb_horfft_dev_noalloc(ctx.h, pb, sb, 1, nf, None) == 0
            assert lib.lsb_synchronize(ctx.h) == 0
            jb = hip.get(pb, (nf,) + phys_shape)
            err = np.abs(jb[..., :nx] - ib[..., :nx]).max() / np.abs(ib[..., :nx]).max()
            print(f"{nf} fields: noalloc == horfft_dev bit for bit, second round trip {err:.1e}", flush=True)
            assert err < 1e-13
    finally:
        for p in bufs:
            hip.lib.hipFree(p)
        ctx.close()


def test_noalloc_makes_no_device_allocation():
    """After a warm-up call (the library's code objects and scratch array are resident),
    further calls leave the device's free memory unchanged (printed beside it: what
    lsb_horfft_dev's first inverse call on the context takes for its staging buffer)."""
    hip = Hip()
    lib = _lib.lib()
    case = L.cases.CASES["c1_dns"].with_(nx=128, ny=64)
    ctx = setup_case(case)
    nzp, ny, nx = ctx.nzp, ctx.ny, ctx.nx
    nf = 4
    q = np.random.default_rng(3).standard_normal((nf, nzp, ny, nx + 2))
    q[..., nx:] = 0.0
    p, s = hip.alloc(q.nbytes), hip.alloc(nf * nzp * ny * nx * 8)
    try:
        hip.put(p, q)
        assert lib.lsb_horfft_dev_noalloc(ctx.h, p, s, -1, nf, None) == 0
        assert lib.lsb_horfft_dev_noalloc(ctx.h, p, s, 1, nf, None) == 0
        hip.lib.hipDeviceSynchronize()
        before = hip.free_bytes()
        for _ in range(3):
            assert lib.lsb_horfft_dev_noalloc(ctx.h, p, s, -1, nf, None) == 0
            assert lib.lsb_horfft_dev_noalloc(ctx.h, p, s, 1, nf, None) == 0
        hip.lib.hipDeviceSynchronize()
        after_noalloc = hip.free_bytes()
        assert lib.lsb_horfft_dev(ctx.h, p, s, 1, nf, None) == 0
        hip.lib.hipDeviceSynchronize()
        after_dev = hip.free_bytes()
        staging = nf * nzp * ny * nx * 8
        print(f"free memory: {before} -> {after_noalloc} (noalloc) -> {after_dev} (horfft_dev, "
              f"staging {staging} bytes)", flush=True)
        assert after_noalloc == before
    finally:
        hip.lib.hipFree(p)
        hip.lib.hipFree(s)
        ctx.close()


def test_noalloc_refuses_fields_beyond_the_scratch():
    """512 x 257 x 256: one field's inverse staging (257 * 256 * 512 = 33,685,504 doubles)
    exceeds the 2^25-double array; refused on the host before any launch, with the
    library's message.  The arrays are allocated at full size all the same, so that the
    call stays in bounds whatever it does."""
    hip = Hip()
    lib = _lib.lib()
    case = L.cases.CASES["c1_dns"].with_(nx=512, ny=256, nz=256)
    nzp, ny, nx = case.nz + 1, case.ny, case.nx
    # preconditions, on the host: both lengths take the staged (register-resident) path,
    # and one field's staging is larger than the array
    assert (nx // 2) in (16, 32, 64, 128, 256) and ny in (16, 32, 64, 128, 256)
    assert nzp * ny * nx > (1 << 25)
    ctx = setup_case(case)
    phys = hip.alloc(nzp * ny * (nx + 2) * 8)
    spec = hip.alloc(nzp * ny * nx * 8)
    try:
        rc = lib.lsb_horfft_dev_noalloc(ctx.h, phys, spec, 1, 1, None)
        msg = lib.lsb_last_error().decode()
        print("refused:", rc, msg, flush=True)
        assert rc == -1 and "exceeds the scratch maximum" in msg
    finally:
        hip.lib.hipFree(phys)
        hip.lib.hipFree(spec)
        ctx.close()
