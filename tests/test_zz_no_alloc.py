"""swe_hip_run_noalloc -- the device loop on caller-owned arrays with its
working memory in a static scratch array -- against swe_hip_run on a block
with the same state: identical states and step sizes, on two calls in a row
(the second with a smaller block, so stale scratch from the first call must
not leak into it), and a block too large for the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from helpers import same_bits

pytestmark = pytest.mark.gpu


def _bind(swe):
    lib = swe.lib
    f = lib.swe_hip_run_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_int)] + [C.c_void_p] * 7 + [C.c_int, C.c_void_p, C.c_void_p]
    return f


def _reference(swe, state, solver, bnd, nsteps):
    h, hu, hv, b, dx, dy = state
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    with swe.SweBlock(nx, ny, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        _, dts = blk.run(nsteps)
        return blk.download(), np.asarray(dts, np.float32)


class _Hip:
    """The HIP runtime the library itself uses, for the caller's device
    buffers (the test plays the caller that owns its memory)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")  # the soname the library links
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.rt, name).restype = C.c_int
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out


def _noalloc(swe, f, state, solver, bnd, nsteps):
    h, hu, hv, b, dx, dy = state
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    hip = _Hip()
    dev = [hip.put(a) for a in (h, hu, hv, b)]
    second = [hip.alloc(h.nbytes) for _ in range(3)]
    log = hip.alloc(4 * nsteps)
    bn = (C.c_int * 4)(*bnd)
    try:
        rc = f(nx, ny, dx, dy, solver, bn, *dev, *second, nsteps, log, None)
        assert rc == 0, swe.lib.swe_hip_last_error().decode()
        assert hip.rt.hipDeviceSynchronize() == 0
        return [hip.get(p, h.shape) for p in dev[:3]], hip.get(log, (nsteps,))
    finally:
        for p in dev + second + [log]:
            hip.rt.hipFree(p)


def test_noalloc_matches_run_on_two_calls(swe):
    from swe_ppm_amd import scenarios

    f = _bind(swe)
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=700, seed=21)
    big = tuple(np.ascontiguousarray(a[:, :502]) for a in (h, hu, hv, b)) + (dx, dy)  # 700 x 500
    h, hu, hv, b, dx, dy, _ = scenarios.dam_break_islands(n=300, seed=22)
    wall = [swe.WALL, swe.OUTFLOW, swe.WALL, swe.WALL]
    small = tuple(np.ascontiguousarray(a[:, :262]) for a in (h, hu, hv, b))  # 300 x 260
    scenarios.apply_bc(small[0], small[1], small[2], wall)
    scenarios.apply_bc_bathymetry(small[3], wall)
    small = small + (dx, dy)
    for state, solver, b4, nsteps in ((big, swe.AUGRIE, [swe.OUTFLOW] * 4, 120), (small, swe.HYBRID, wall, 77)):
        ref, rdts = _reference(swe, state, solver, b4, nsteps)
        got, gdts = _noalloc(swe, f, state, solver, b4, nsteps)
        for name, a, c in zip(("h", "hu", "hv"), got, ref):
            assert same_bits(a, c), (state[0].shape, name)
        assert same_bits(gdts, rdts)


def test_noalloc_refuses_blocks_beyond_its_scratch(swe):
    f = _bind(swe)
    bn = (C.c_int * 4)(0, 0, 0, 0)
    dummy = C.c_void_p(1)
    rc = f(4000, 4000, 1.0, 1.0, 2, bn, *[dummy] * 7, 10, None, None)
    assert rc == -1
    assert b"SWE_HIP_NOALLOC_SCRATCH_BYTES" in swe.lib.swe_hip_last_error()
