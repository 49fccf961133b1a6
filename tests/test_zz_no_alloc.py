"""fh_sweep_device_noalloc (no device allocation in the call, per-block
partials in the library's fixed scratch) gives fh_sweep_device's results bit
for bit, on two calls in a row, the second on a smaller input."""
import ctypes as C

import numpy as np
import pytest

from streamsim_amd import FacetEngine, _lib, scenes

pytestmark = pytest.mark.gpu


class _Device:
    """hipMalloc / hipMemcpy / hipFree through ctypes for the device arrays of the test."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        self.held.append(p)
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def _sweep_pair(lib, eng, dev, S, nb, precision, seed):
    """One sweep through each entry point on the same device inputs."""
    ldb = int(lib.fh_band_stride(nb))
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 50.0, (S, ldb))
    r = rng.uniform(0.0, 0.5, (S, ldb))
    t = rng.uniform(0.0, 0.4, (S, ldb))
    B = rng.uniform(0.0, 80.0, (S, ldb))
    dc, dr, dt, dB = dev.put(c), dev.put(r), dev.put(t), dev.put(B)
    d_fixed, d_ref = dev.put(np.full((S, ldb), -1.0)), dev.put(np.full((S, ldb), -2.0))
    # the no-allocation call first: it needs nothing the other one sets up
    assert lib.fh_sweep_device_noalloc(eng.handle, nb, dc, dr, dt, dB, d_fixed, precision) == 0, \
        lib.fh_last_error(eng.handle)
    assert lib.fh_sweep_device(eng.handle, nb, dc, dr, dt, dB, d_ref, precision) == 0
    return dev.get(d_fixed, (S, ldb)), dev.get(d_ref, (S, ldb))


def test_noalloc_sweep_matches_sweep_device():
    print("no-alloc sweep", flush=True)
    lib = _lib.load()
    eng = FacetEngine(rays_per_side=64, sun_rays_per_side=16, seed=3)
    dev = _Device()
    try:
        big = scenes.row_crop(leaves_per_row=60, cells=8)
        eng.setGeometry(big.vertices)
        eng.buildVisibilityGraph()
        for nb, prec, seed in ((47, _lib.FH_PRECISION_FP32, 1), (24, _lib.FH_PRECISION_BF16_DC, 2)):
            fixed, ref = _sweep_pair(lib, eng, dev, big.surface_count, nb, prec, seed)
            assert np.all(np.isfinite(ref[:, :nb])) and np.any(ref[:, :nb] != 0.0)
            np.testing.assert_array_equal(fixed[:, :nb], ref[:, :nb])
        dev.free()
        # second call on a smaller scene (fewer blocks, the entry-parallel kernels)
        small = scenes.row_crop(leaves_per_row=12, cells=3)
        eng.setGeometry(small.vertices)
        eng.buildVisibilityGraph()
        fixed, ref = _sweep_pair(lib, eng, dev, small.surface_count, 3, _lib.FH_PRECISION_FP32, 3)
        assert np.any(ref[:, :3] != 0.0)
        np.testing.assert_array_equal(fixed[:, :3], ref[:, :3])
        # invalid arguments are rejected as by fh_sweep_device
        assert lib.fh_sweep_device_noalloc(eng.handle, 0, None, None, None, None, None, 0) == _lib.FH_INVALID_ARGUMENT
    finally:
        dev.free()
        eng.destroy()
