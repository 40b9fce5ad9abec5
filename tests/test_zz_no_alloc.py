"""Allocation-free dense k=2 search (orb_knn2_device_noalloc) against
orb_knn2_device on the same device buffers: identical neighbour indices,
distances and ratio flags, for two calls in a row of the allocation-free
entry point, the second smaller than the first (the fixed scratch array
still holds the first call's partial keys when the second starts).  Device
buffers come from the HIP runtime through ctypes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0      # device to host
        return out

    def free(self, *ps):
        for p in ps:
            self.lib.hipFree(p)


def _knn(orb, ctx, hip, fn, dq, nq, dm, nm, method):
    idx = hip.put(np.full((nq, 2), -7, np.int32))
    dist = hip.put(np.full((nq, 2), -7, np.int32))
    acc = hip.put(np.full(nq, 9, np.uint8))
    rc = getattr(orb.lib(), fn)(ctx.h, dq, nq, dm, nm, method, 0.7, idx, dist, acc)
    assert rc == 0, (fn, rc)
    ctx.synchronize()
    out = (hip.get(idx, (nq, 2), np.int32), hip.get(dist, (nq, 2), np.int32), hip.get(acc, nq, np.uint8))
    hip.free(idx, dist, acc)
    return out


@pytest.mark.parametrize("method", [0, 1])
def test_noalloc_equals_allocating_entry_point(orb, method):
    from orb_slam2_detailed_comments_amd import synth
    ctx = orb.Context.default()
    hip = _Hip()
    cases = []
    for nq, nm, seed in [(2000, 300_000, 5), (700, 30_000, 6)]:
        q, m, _ = synth.descriptor_sets(nq, nm, seed)
        dq, dm = hip.put(q), hip.put(m)
        cases.append((dq, nq, dm, nm, _knn(orb, ctx, hip, "orb_knn2_device", dq, nq, dm, nm, method)))
    got = [_knn(orb, ctx, hip, "orb_knn2_device_noalloc", dq, nq, dm, nm, method) for dq, nq, dm, nm, _ in cases]
    for (dq, nq, dm, nm, want), g in zip(cases, got):
        for a, b, name in zip(g, want, ("idx", "dist", "accept")):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} nq={nq}")
        assert (g[0][:, 0] >= 0).all()
        hip.free(dq, dm)


def test_noalloc_rejects_sizes_beyond_its_array(orb):
    """2000 queries x 3,000,000 rows needs 2000 * 367 partial keys > 2^19:
    rejected before any launch."""
    ctx = orb.Context.default()
    hip = _Hip()
    d = hip.alloc(64)
    rc = orb.lib().orb_knn2_device_noalloc(ctx.h, d, 2000, d, 3_000_000, 1, 0.7, d, d, d)
    hip.free(d)
    assert rc == orb.ORB_E_CAPACITY
