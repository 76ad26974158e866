"""GPU: zo_spacing_static, the spacing entry point that allocates nothing, against
zo_spacing on the same positions: the nearest orbit point's rank of every point, its SO(3)
angle and the separation, bit-identical, on two calls in a row (the second smaller, so it
runs over what the first left in the shared scratch array), and its errors.

The test is the caller: it owns the device buffers (hipMalloc through the HIP runtime the
library is linked against) and passes them in; the library allocates nothing.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import zxtalori_amd as zo
from zxtalori_amd.api import _group_id, _library

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def note(*args):
    print(*args, file=sys.stderr, flush=True)


class Hip:
    """The few HIP runtime calls a caller makes to own device buffers."""

    def __init__(self):
        _library()  # libzxtalori.so first: then this name resolves to the runtime it loaded
        self.rt = C.CDLL("libamdhip64.so.7")
        self.held = []

    def check(self, code, what):
        assert code == 0, f"{what}: hip error {code}"

    def alloc(self, nbytes):
        ptr = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)), "hipMalloc")
        self.held.append(ptr)
        return ptr

    def put(self, array):
        array = np.ascontiguousarray(array)
        ptr = self.alloc(array.nbytes)
        self.check(self.rt.hipMemcpy(ptr, array.ctypes.data_as(C.c_void_p), C.c_size_t(array.nbytes), H2D), "upload")
        return ptr

    def get(self, ptr, count, dtype):
        out = np.empty(count, dtype)
        self.check(self.rt.hipDeviceSynchronize(), "synchronize")
        self.check(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr, C.c_size_t(out.nbytes), D2H), "download")
        return out

    def free(self):
        for ptr in self.held:
            self.rt.hipFree(ptr)
        self.held.clear()


def spacing_static(lib, hip, group, q):
    n = len(q)
    q_dev = hip.put(q.astype(np.float32))
    rank_dev, deg_dev, least_dev = hip.alloc(4 * n), hip.alloc(4 * n), hip.alloc(256)
    code = lib.zo_spacing_static(_group_id(group), q_dev, n, rank_dev, deg_dev, least_dev, None)
    assert code == 0, lib.zo_create_error().decode()
    return (hip.get(rank_dev, n, np.uint32), hip.get(deg_dev, n, np.float32),
            float(hip.get(least_dev, 1, np.float32)[0]))


def engine_spacing(group, q):
    with zo.from_quaternions(q, group) as grid:
        stats = grid.spacing()
        src, op, sign = grid.nearest_indices()
        order = stats["order"]
        rank = (src * 2 * order + op * 2 + sign).astype(np.uint32)
        return rank, grid.point_values("separation"), stats["separation_deg"], grid.quaternions


def test_spacing_static_matches_zo_spacing():
    lib = _library()
    lib.zo_spacing_static.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    lib.zo_spacing_static.restype = C.c_int32
    hip = Hip()
    try:
        # Two calls in a row, the second smaller, over the same scratch.
        for group, count, seed in (("I", 16_000, 3), ("O", 2_500, 8), ("C2", 20_000, 1)):
            note("static spacing", group, count)
            with zo.haar(group, count, seed) as grid:
                q = grid.quaternions
            # The engine's positions (imported, so folded once more) are the static call's input.
            rank_ref, deg_ref, least_ref, q = engine_spacing(group, q)
            rank, deg, least = spacing_static(lib, hip, group, q)
            np.testing.assert_array_equal(rank, rank_ref)
            np.testing.assert_array_equal(deg, deg_ref)
            assert least == least_ref
            hip.free()
        # Sizes beyond the scratch array, and bad arguments, fail with a message.
        q_dev, out = hip.alloc(16 * 20_001), hip.alloc(4 * 20_001)
        assert lib.zo_spacing_static(_group_id("I"), q_dev, 20_001, out, out, out, None) == -1
        assert "CountExceedsStaticScratch" in lib.zo_create_error().decode()
        assert lib.zo_spacing_static(_group_id("I"), q_dev, 1, out, out, out, None) == -1
        assert "TooFewOrientations" in lib.zo_create_error().decode()
        assert lib.zo_spacing_static(12, q_dev, 100, out, out, out, None) == -1
        assert "UnknownGroup" in lib.zo_create_error().decode()
        assert lib.zo_spacing_static(_group_id("I"), None, 100, out, out, out, None) == -1
        assert "NullPointer" in lib.zo_create_error().decode()
    finally:
        hip.free()
