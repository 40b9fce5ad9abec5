"""GPU tier: a context outlives its handle while extractors, batches and
servers made on it are alive (each holds a reference), so the order in which
a caller (or Python's finaliser) destroys them does not matter."""
import ctypes

import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import KEYPOINT_DTYPE, synth

pytestmark = pytest.mark.gpu


def test_context_destroyed_before_its_extractor(orb, oracle):
    L = orb.lib()
    ctx, ex = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.orb_context_create(0, ctypes.byref(ctx)) == 0
    assert L.orb_extractor_create(ctx, 1000, ctypes.c_float(1.2), 8, 20, 7, ctypes.byref(ex)) == 0
    L.orb_context_destroy(ctx)                 # the extractor still holds the context
    img = np.ascontiguousarray(synth.shapes_frame(640, 480, 3), np.uint8)
    cap = L.orb_extractor_capacity(ex, 640, 480)
    kps = np.zeros(cap, KEYPOINT_DTYPE)
    desc = np.zeros((cap, 32), np.uint8)
    n = ctypes.c_int()
    rc = L.orb_extract(ex, img.ctypes.data_as(ctypes.c_void_p), 640, 480, ctypes.c_size_t(640),
                       kps.ctypes.data_as(ctypes.c_void_p), cap, desc.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(n))
    assert rc == 0
    ok, od = oracle.OracleExtractor(1000, 1.2, 8, 20, 7)(img)
    assert n.value == len(ok) and kps[:n.value].tobytes() == ok.tobytes()
    np.testing.assert_array_equal(desc[:n.value], od)
    L.orb_extractor_destroy(ex)                # the last reference frees the context


def test_python_objects_closed_in_any_order(orb):
    ctx = orb.Context(0)
    ex = orb.ORBextractor(500, 1.2, 8, 20, 7, context=ctx)
    batch = orb.Batch([(640, 480)] * 2, 500, context=ctx)
    ctx.close()
    img = synth.shapes_frame(640, 480, 4)
    k, _ = ex(img)
    batch.upload(0, img)
    batch.upload(1, img)
    batch.run(1)
    batch.synchronize()
    assert len(batch.results(0)[0]) == len(k) > 0
    ex.close()
    batch.close()
