"""The entry point that allocates nothing (fs_step_noalloc) against the one it
varies (fs_create_from_h + fs_step), through the C ABI.

fs_step_noalloc keeps every work array in one device array built into the
library, shared by all calls on a GPU, so these checks run several calls in
a row with different grids (the second smaller than the first, then a larger
one again, on the null stream and on a caller's stream): each must give
exactly the bits of a freshly created model, which proves that what one call
leaves in the shared array does not leak into the next.  The cases cover the
tile path with an odd number of steps (the result comes back from the
second buffer), the 32-cell halo and the level path (the reference preset
passes 32 levels within a few tens of steps; fs_step_noalloc then takes the
level path, the model it is compared with the tile-graph path), and n = 2.
"""
import ctypes
import sys

import numpy as np
import pytest

import fastscape_amd as fa

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipStreamCreate, h.hipStreamDestroy):
            f.restype = ctypes.c_int
        _HIP = h
    return _HIP


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def noalloc_steps(h0, W, H, params, iterations, stream=None):
    """Copy h0 to a device buffer (the caller's allocation), step it with
    fs_step_noalloc, copy it back."""
    dev = ctypes.c_void_p()
    nbytes = W * H * 8
    assert hip().hipMalloc(ctypes.byref(dev), nbytes) == 0
    try:
        src = np.ascontiguousarray(h0, dtype=np.float64)
        assert hip().hipMemcpy(dev, src.ctypes.data, nbytes, 1) == 0        # host to device
        st = fa.library().fs_step_noalloc(dev, W, H, ctypes.byref(params), iterations, stream)
        assert st == 0, fa.FastScapeError(st, "fs_step_noalloc")
        out = np.empty(W * H, dtype=np.float64)
        assert hip().hipMemcpy(out.ctypes.data, dev, nbytes, 2) == 0        # device to host
        return out
    finally:
        hip().hipFree(dev)


def fs_step_reference(h0, W, H, params, iterations):
    with fa.FastScape(W, H, params, h=h0) as m:
        m.step(iterations)
        return m.getH().ravel(), m.path_stats(), m.tile_stats()


# (W, H, preset, terrain seed, steps, expect drainage deeper than the halos)
CASES = [
    (700, 600, fa.PRESET_BENCHMARK, 0, 31, False),
    (300, 200, fa.PRESET_REFERENCE, 4, 40, True),
    (1024, 1024, fa.PRESET_BENCHMARK, 9, 12, False),
]


def test_noalloc_equals_fs_step_on_consecutive_calls():
    stream = ctypes.c_void_p()
    assert hip().hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for k, (W, H, pre, seed, n, deep) in enumerate(CASES):
            note("fs_step_noalloc call %d: %dx%d preset %d, %d steps" % (k, W, H, pre, n))
            p = fa.preset(pre)
            h0 = fa.generate_terrain(W, H, seed)
            want, paths, tiles = fs_step_reference(h0, W, H, p, n)
            if deep:
                assert paths["graph"] > 0 and tiles["halo32"] > 0, (paths, tiles)
            got = noalloc_steps(h0, W, H, p, n, stream if k % 2 else None)
            np.testing.assert_array_equal(got, want, err_msg="call %d %dx%d" % (k, W, H))
        # the small grid once more, right after the large one
        W, H, pre, seed, n, _ = CASES[1]
        p = fa.preset(pre)
        h0 = fa.generate_terrain(W, H, seed)
        want, _, _ = fs_step_reference(h0, W, H, p, n)
        np.testing.assert_array_equal(noalloc_steps(h0, W, H, p, n), want, err_msg="repeat")
    finally:
        hip().hipStreamDestroy(stream)


def test_noalloc_rejects_grids_beyond_its_array():
    lib = fa.library()
    p = fa.preset(fa.PRESET_BENCHMARK)
    dev = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(dev), 8 * 64) == 0
    try:
        assert lib.fs_step_noalloc(dev, 1025, 1024, ctypes.byref(p), 1, None) == 2
        assert lib.fs_step_noalloc(None, 64, 64, ctypes.byref(p), 1, None) == 1
        assert lib.fs_step_noalloc(dev, 8, 8, ctypes.byref(p), -1, None) == 1
        assert lib.fs_step_noalloc(dev, 8, 8, ctypes.byref(p), 0, None) == 0
    finally:
        hip().hipFree(dev)
