#!/usr/bin/env python3
"""Benchmark: ORB extract + stereo match frames/s at 1241x376, 2000 keypoints.

A step is one pass of the hot path over one batch of synthetic stereo frames
resident in HBM: extraction of both views (pyramid, FAST cells, quadtree,
orientation, rotated BRIEF) and Frame::ComputeStereoMatches for every frame
of the batch, replayed from one captured hipGraph.  value = frames/s where a
frame is one stereo pair.

Also reported:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the library's stream), against
                MI355X HBM peak 8 TB/s
  cpu_baseline  the CPU oracle (a restatement of the reference path) timed on
                this host's cores on a bounded sample of the same frames:
                floor(cores/2) independent Frame pipelines, each extracting
                L/R on two threads then matching (src/Frame.cc:146-169)
One GPU only: --gpus N>1 exits with a message.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s ORB extract+match @1241x376/2000kp; p99 ms at 16 streams"
W, H, NF, NLEV = 1241, 376, 2000, 8
BF, FX = 386.1, 718.856
HBM_PEAK_GBS = 8000.0


def make_frames(n, seed0=1000):
    from orb_slam2_detailed_comments_amd import synth
    return [synth.stereo_pair(W, H, seed0 + i)[:2] for i in range(n)]


def level_sizes(w, h, nlevels=NLEV, scale=1.2):
    inv = [1.0]
    s = np.float32(1.0)
    for _ in range(1, nlevels):
        s = np.float32(np.float64(s) * np.float64(np.float32(scale)))
        inv.append(np.float32(1.0) / s)
    return [(int(np.rint(np.float32(w) * np.float32(i))), int(np.rint(np.float32(h) * np.float32(i)))) for i in inv]


def stage_bytes(stage, n_images, kps_per_image, cands_per_image):
    """Algorithmic HBM bytes of one launch of a stage over n_images images."""
    sizes = level_sizes(W, H)
    if stage.startswith("pyramid_l"):
        l = int(stage[len("pyramid_l"):])
        w, h = sizes[l]
        src = W * H if l == 0 else sizes[l - 1][0] * sizes[l - 1][1]
        # read the source level; write the level (l>0) and its blur
        return n_images * (src + (2 if l > 0 else 1) * w * h)
    if stage == "octree":
        return n_images * (8 * cands_per_image + 8 * kps_per_image)
    if stage == "describe":
        # 721-px orientation disc + 512 BRIEF samples per keypoint, 60 B out
        return n_images * kps_per_image * (721 + 512 + 60)
    if stage == "stereo_match":
        return (n_images // 2) * kps_per_image * (28 + 32 + 8 + 2 * 121 * 2)
    return None


def knn_extra(orb, ctx, steps):
    """Dense k=2 search, 2000 queries x 1,000,000 map rows resident in HBM."""
    import ctypes
    from orb_slam2_detailed_comments_amd import synth
    Q, M = 2000, 1_000_000
    q, m, _ = synth.descriptor_sets(Q, M, 3)
    out = {"workload": "k=2 Hamming search, 2000 queries x 1,000,000 map rows (10% planted), ratio 0.7",
           "int8_ops": 2 * Q * M * 256}
    for name, meth in (("popcount", 0), ("mfma", 1)):
        ms_s, ms_m = ctypes.c_float(), ctypes.c_float()
        rc = orb.lib().orb_knn2_time(ctx.h, ctypes.c_void_p(q.ctypes.data), Q, ctypes.c_void_p(m.ctypes.data), M,
                                     meth, 0.7, steps, ctypes.byref(ms_s), ctypes.byref(ms_m))
        if rc != 0:
            raise RuntimeError(f"orb_knn2_time {rc}")
        out[f"{name}_search_ms"] = round(ms_s.value, 4)
        out[f"{name}_merge_ms"] = round(ms_m.value, 4)
        out[f"{name}_searches_per_s"] = round(1e3 / (ms_s.value + ms_m.value), 1)
    tops = out["int8_ops"] / (out["mfma_search_ms"] * 1e-3) / 1e12
    out["mfma_roofline"] = {"bound": "mfma", "achieved": round(tops, 1), "peak": 5000.0, "unit": "TOP/s",
                            "frac": round(tops / 5000.0, 4)}
    # the popcount form does 8 XOR + 8 v_bcnt + 3 select ops per pair
    gops = Q * M * 19 / (out["popcount_search_ms"] * 1e-3) / 1e12
    out["popcount_valu_tops"] = round(gops, 1)
    return out


def bow_extra(orb, ctx, steps, cpu_seconds):
    """DBoW2 transform (Frame::ComputeBoW, levelsup 4) of 64 images x 2000
    descriptors against a k=10, L=6 tree (1,111,111 nodes, ORBvoc.txt's
    shape), descriptors resident in HBM; the oracle on one host core beside."""
    import ctypes
    from orb_slam2_detailed_comments_amd import synth
    voc = synth.vocabulary(10, 6, 41)
    n_img, per = 64, 2000
    q = synth.vocabulary_queries(voc, n_img * per, 43, flip_ands=3)
    offs = np.arange(0, n_img * per + 1, per, dtype=np.int32)
    v = orb.Vocabulary(context=ctx, **voc)
    ms = ctypes.c_float()
    rc = orb.lib().orb_vocabulary_bow_time(v.h, n_img, ctypes.c_void_p(q.ctypes.data),
                                           ctypes.c_void_p(offs.ctypes.data), 4, steps, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"orb_vocabulary_bow_time {rc}")
    out = {"workload": "DBoW2 transform + BowVector + FeatureVector, 64 images x 2000 descriptors, k=10 L=6 "
                       "synthetic tree (1,111,111 nodes), levelsup 4",
           "ms_per_batch": round(ms.value, 4), "images_per_s": round(n_img / (ms.value * 1e-3), 1),
           "descriptors_per_s": round(n_img * per / (ms.value * 1e-3), 1)}
    v.close()
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.OracleVocabulary(**voc)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            o.bow(q[(n % n_img) * per:(n % n_img + 1) * per], 4)
            n += 1
        cpu = n / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(cpu, 2), "unit": "images/s", "cores": 1, "kind": "port",
                               "sample": f"{n} images of 2000 descriptors through the oracle in ~{cpu_seconds:.0f}s"}
    return out


def mappoint_extra(orb, ctx, steps, cpu_seconds):
    """MapPoint::ComputeDistinctiveDescriptors for 20,000 map points with
    2..40 observations each (clustered descriptors), rows resident in HBM."""
    import ctypes
    rng = np.random.default_rng(77)
    sizes = rng.integers(2, 41, size=20000)
    pts = []
    for n in sizes:
        base = rng.integers(0, 256, size=32, dtype=np.uint8)
        pts.append(base ^ (rng.integers(0, 256, size=(n, 32), dtype=np.uint8) &
                           rng.integers(0, 256, size=(n, 32), dtype=np.uint8) &
                           rng.integers(0, 256, size=(n, 32), dtype=np.uint8)))
    offs = np.zeros(len(pts) + 1, np.int32)
    offs[1:] = np.cumsum(sizes)
    desc = np.concatenate(pts)
    ms = ctypes.c_float()
    rc = orb.lib().orb_distinctive_descriptors_time(ctx.h, len(pts), ctypes.c_void_p(desc.ctypes.data),
                                                    ctypes.c_void_p(offs.ctypes.data), steps, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"orb_distinctive_descriptors_time {rc}")
    out = {"workload": "ComputeDistinctiveDescriptors, 20,000 map points x U[2,40] observations",
           "ms_per_batch": round(ms.value, 4), "points_per_s": round(len(pts) / (ms.value * 1e-3), 1)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        n, t0 = 0, time.perf_counter()
        chunk = 500
        while time.perf_counter() - t0 < cpu_seconds:
            k = (n // chunk) % (len(pts) // chunk)
            pyoracle.distinctive_descriptors(pts[k * chunk:(k + 1) * chunk])
            n += chunk
        cpu = n / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(cpu, 1), "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"{n} points through the oracle in ~{cpu_seconds:.0f}s"}
    return out


def frame_ops_extra(orb, ctx, steps, cpu_seconds):
    """The RGB-D Frame constructor's per-keypoint steps after the extraction:
    UndistortKeyPoints (TUM fr1's 5-coefficient model) and
    ComputeStereoFromRGBD on a 640x480 float depth map, 2000 keypoints,
    host arrays in and out per call."""
    from orb_slam2_detailed_comments_amd import KEYPOINT_DTYPE
    rng = np.random.default_rng(5)
    n, w, h = 2000, 640, 480
    kps = np.zeros(n, KEYPOINT_DTYPE)
    kps["x"] = (rng.random(n) * w).astype(np.float32)
    kps["y"] = (rng.random(n) * h).astype(np.float32)
    depth = (rng.random((h, w)) * 4 + 0.3).astype(np.float32)
    K = (517.306408, 516.469215, 318.643040, 255.313989)
    D = (0.262383, -0.953104, -0.005358, 0.002628, 1.163314)
    for _ in range(3):
        kun = orb.undistort_keypoints(kps, K, D, context=ctx)
        orb.rgbd_depth(kps, kun, depth, 40.0, context=ctx)
    t0 = time.perf_counter()
    for _ in range(steps):
        kun = orb.undistort_keypoints(kps, K, D, context=ctx)
        orb.rgbd_depth(kps, kun, depth, 40.0, context=ctx)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = {"workload": "UndistortKeyPoints + ComputeStereoFromRGBD, 2000 keypoints, 640x480 depth, host arrays",
           "ms_per_frame": round(ms, 4), "frames_per_s": round(1e3 / ms, 1)}
    # The whole RGB-D Frame hot path in one call (orb_rgbd_frame): a 640x480
    # frame extracted with 1000 features, then the two steps on the device.
    from orb_slam2_detailed_comments_amd import synth
    img = synth.shapes_frame(640, 480, 8)
    ex = orb.ORBextractor(1000, 1.2, 8, 20, 7, context=ctx)
    for _ in range(3):
        orb.rgbd_frame(ex, img, depth, K, D, 40.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        orb.rgbd_frame(ex, img, depth, K, D, 40.0)
    ms_f = (time.perf_counter() - t0) * 1e3 / steps
    out["rgbd_frame"] = {"workload": "orb_rgbd_frame: 640x480 extraction (1000 features) + UndistortKeyPoints + "
                                     "ComputeStereoFromRGBD, host image and depth in, results out",
                         "ms_per_frame": round(ms_f, 4), "frames_per_s": round(1e3 / ms_f, 1)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        xy = np.stack([kps["x"], kps["y"]], 1)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            un = pyoracle.undistort_points(xy, K, D)
            pyoracle.rgbd_depth(xy, un[:, 0], depth, 40.0)
            k += 1
        cpu = k / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(cpu, 1), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"{k} frames through the oracle in ~{cpu_seconds:.0f}s"}
        oex = pyoracle.OracleExtractor(1000, 1.2, 8, 20, 7)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            okps, _ = oex(img)
            oxy = np.stack([okps["x"], okps["y"]], 1)
            oun = pyoracle.undistort_points(oxy, K, D)
            pyoracle.rgbd_depth(oxy, oun[:, 0], depth, 40.0)
            k += 1
        cpu = k / (time.perf_counter() - t0)
        out["rgbd_frame"]["cpu_baseline"] = {"value": round(cpu, 1), "unit": "frames/s", "cores": 1, "kind": "port",
                                             "sample": f"{k} frames through the oracle in ~{cpu_seconds:.0f}s"}
    return out


def input_extra(orb, ctx, steps, cpu_seconds):
    """Input-side image ops in front of the pyramid: cvtColor to gray of a
    1241x376 RGB frame and EuRoC cam0 rectification (initUndistortRectifyMap
    maps in HBM, remap INTER_LINEAR) of a 752x480 frame, each alone (host
    image in, host image out) and fused in front of the extraction (against
    the plain extraction of the same gray frame), per frame; the oracle on
    one core beside the standalone ops."""
    from orb_slam2_detailed_comments_amd import synth
    gray = synth.shapes_frame(1241, 376, 9)
    rgb = np.stack([gray, np.roll(gray, 3, 1), np.roll(gray, 5, 0)], 2)
    raw = synth.shapes_frame(752, 480, 10)
    K = [[458.654, 0, 367.215], [0, 457.296, 248.375], [0, 0, 1]]
    D = [-0.28340811, 0.07395907, 0.00019359, 1.76187114e-05, 0.0]
    R = [[0.999966347530033, -0.001422739138722922, 0.008079580483432283],
         [0.001365741834644127, 0.9999741760894847, 0.007055629199258132],
         [-0.008089410156878961, -0.007044357138517677, 0.9999424675829176]]
    P = [[435.2046959714599, 0, 367.4517211914062, 0], [0, 435.2046959714599, 252.2008514404297, 0], [0, 0, 1, 0]]
    rect = orb.Rectifier(K, D, R, P, 752, 480, context=ctx)
    ex_k = orb.ORBextractor(2000, 1.2, 8, 20, 7, context=ctx)
    ex_e = orb.ORBextractor(1000, 1.2, 8, 20, 7, context=ctx)

    def per_frame(fn):
        for _ in range(3):
            fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        return round((time.perf_counter() - t0) * 1e3 / steps, 4)

    out = {"workload": "cvtColor RGB->gray 1241x376; EuRoC cam0 remap 752x480 (maps resident); host images",
           "gray_ms": per_frame(lambda: orb.cvt_color_gray(rgb, True, context=ctx)),
           "remap_ms": per_frame(lambda: rect.remap(raw)),
           "extract_gray_ms": per_frame(lambda: ex_k(gray)),
           "extract_color_ms": per_frame(lambda: ex_k.extract_color(rgb, True)),
           "extract_752x480_ms": per_frame(lambda: ex_e(raw)),
           "extract_rectified_ms": per_frame(lambda: rect.extract(ex_e, raw))}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        mx, my = rect.maps()
        for name, fn in (("gray", lambda: pyoracle.cvt_gray(rgb, True)),
                         ("remap", lambda: pyoracle.remap_linear(raw, mx, my))):
            k, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < cpu_seconds / 2:
                fn()
                k += 1
            out[f"{name}_cpu_baseline"] = {"value": round(k / (time.perf_counter() - t0), 1), "unit": "frames/s",
                                           "cores": 1, "kind": "port"}
    return out


def winmatch_extra(orb, ctx, cpu_seconds):
    """SearchForInitialization (window 100, ratio 0.9, orientation check) for
    32 frame pairs of 640x480 with the initializer's 2000 features, one call:
    host-to-host time, and the device work alone with the inputs in HBM; the
    oracle beside it on one core."""
    from orb_slam2_detailed_comments_amd import synth
    ex = orb.ORBextractor(2000, 1.2, 8, 20, 7, context=ctx)
    pairs = []
    for i in range(32):
        a, b = synth.frame_pair(640, 480, 900 + i, angle_deg=2.0 + 0.1 * i, shift=(5.0, 3.0))
        k1, d1 = ex(a)
        k2, d2 = ex(b)
        pairs.append((k1, d1, k2, d2, (0.0, 640.0, 0.0, 480.0), np.stack([k1["x"], k1["y"]], 1)))
    orb.search_for_initialization(pairs[:2], context=ctx)           # warm-up
    reps, t0 = 5, time.perf_counter()
    for _ in range(reps):
        res = orb.search_for_initialization(pairs, context=ctx)
    py_ms = (time.perf_counter() - t0) * 1e3 / reps
    # the C ABI call on flat host arrays, as a C++ caller makes it (the
    # arrays built once; vbPrevMatched restored between calls, untimed)
    a = orb._sfi_arrays(pairs)
    prev0, ms = a["prev"].copy(), 0.0
    for _ in range(reps):
        a["prev"][:] = prev0
        t0 = time.perf_counter()
        orb._sfi_call(ctx, a, 100, 0.9, True)
        ms += (time.perf_counter() - t0) * 1e3 / reps
    _, dev_ms = orb.search_for_initialization(pairs, context=ctx, timing_steps=10)
    out = {"workload": "SearchForInitialization over the 64x48 grid, 32 pairs of 640x480 (2000 features each), "
                       "window 100, ratio 0.9, orientation check; ms_per_call: the C ABI call on flat host arrays in "
                       "and out; python_ms_per_call adds the wrapper's concatenation of 32 pairs",
           "ms_per_call": round(ms, 3), "pairs_per_s": round(32 / (ms * 1e-3), 1),
           "python_ms_per_call": round(py_ms, 3),
           "device_resident": {"ms_per_call": round(dev_ms, 3), "pairs_per_s": round(32 / (dev_ms * 1e-3), 1),
                               "includes": "the call's device work with its inputs already in HBM: grid, candidate "
                                           "lists, the parallel rounds and their flag reads, the finish"},
           "mean_matches": round(float(np.mean([r[0] for r in res])), 1)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            pyoracle.search_for_initialization(*pairs[n % 32])
            n += 1
        out["cpu_baseline"] = {"value": round(n / (time.perf_counter() - t0), 1), "unit": "pairs/s", "cores": 1,
                               "kind": "port", "sample": f"{n} pairs through the oracle in ~{cpu_seconds:.0f}s"}
    return out


def projection_extra(orb, ctx, cpu_seconds):
    """ORBmatcher::SearchByProjection, both per-frame overloads, for 32
    frames of 2000 keypoints and 1800 projected points each in one call
    (host arrays in and out); the oracle on one core beside."""
    from orb_slam2_detailed_comments_amd import synth
    scale = np.float32(1.2) ** np.arange(8, dtype=np.float32)
    out = {"workload": "SearchByProjection, 32 frames x (2000 keypoints, 1800 points) per call; local map (th 3, "
                       "ratio 0.8) and last frame (th 15, rotation check); ms_per_call: the C ABI call on flat host "
                       "arrays in and out; python_ms_per_call adds the wrapper's concatenation of 32 frames"}
    for name, last in (("local_map", False), ("last_frame", True)):
        frames = [synth.projection_frame(500 + i, last=last) for i in range(32)]
        fn = (lambda **kw: orb.search_by_projection_last(frames, scale, 15.0, False, True, context=ctx, **kw)) \
            if last else (lambda **kw: orb.search_by_projection(frames, scale, 3.0, 0.8, context=ctx, **kw))
        fn()
        reps, t0 = 5, time.perf_counter()
        for _ in range(reps):
            res = fn()
        py_ms = (time.perf_counter() - t0) * 1e3 / reps
        # the C ABI call on flat host arrays, as a C++ caller makes it
        a = orb._projection_frames(frames, orb.LAST_POINT_DTYPE if last else orb.PROJ_POINT_DTYPE)
        cam = np.ascontiguousarray([f[7] for f in frames], np.float32).reshape(-1)
        P = orb._ptr
        common = (ctx.h, 32, P(a["kp"]), P(a["desc"]), P(a["uright"]), P(a["occupied"]), P(a["ko"]), P(a["bounds"]),
                  P(scale), len(scale))
        call = (lambda: orb.lib().orb_search_by_projection_last(
                    *common, P(cam), 0, P(a["pts"]), P(a["pdesc"]), P(a["po"]), 15.0, 1, P(a["pm"]), P(a["kpm"]),
                    P(a["nm"]))) if last else \
               (lambda: orb.lib().orb_search_by_projection(
                    *common, P(a["pts"]), P(a["pdesc"]), P(a["po"]), 3.0, 0.8, P(a["pm"]), P(a["kpm"]), P(a["nm"])))
        t0 = time.perf_counter()
        for _ in range(reps):
            assert call() == 0
        ms = (time.perf_counter() - t0) * 1e3 / reps
        _, dev_ms = fn(timing_steps=10)
        r = {"ms_per_call": round(ms, 3), "frames_per_s": round(32 / (ms * 1e-3), 1), "python_ms_per_call": round(py_ms, 3),
             "device_resident": {"ms_per_call": round(dev_ms, 3), "frames_per_s": round(32 / (dev_ms * 1e-3), 1)},
             "mean_matches": round(float(np.mean([x[0] for x in res])), 1)}
        if cpu_seconds > 0:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < cpu_seconds / 2:
                kp, d, ur, occ, b, pts, pd, cam = frames[n % 32]
                if last:
                    pyoracle.search_by_projection_last(kp, d, ur, occ, b, scale, cam, False, pts, pd, 15.0, True)
                else:
                    pyoracle.search_by_projection(kp, d, ur, occ, b, scale, pts, pd, 3.0, 0.8)
                n += 1
            r["cpu_baseline"] = {"value": round(n / (time.perf_counter() - t0), 1), "unit": "frames/s", "cores": 1,
                                 "kind": "port", "sample": f"{n} frames through the oracle"}
        out[name] = r
    return out


def batch_extra(orb, ctx, steps):
    """Mixed-size batch: 64 frames, 16 each of 640x480, 752x480 (nFeatures
    1000), 1241x376 and 1280x720 (nFeatures 2000), extracted by one
    graph-replayed launch per stage against per-frame launches of the same
    kernels."""
    from orb_slam2_detailed_comments_amd import synth
    sizes = [(640, 480, 1000), (752, 480, 1000), (1241, 376, 2000), (1280, 720, 2000)]
    slots = [sizes[i // 16] for i in range(64)]
    imgs = [synth.shapes_frame(w, h, 3000 + i) for i, (w, h, _) in enumerate(slots)]
    batch = orb.Batch([(w, h) for w, h, _ in slots], [nf for _, _, nf in slots], context=ctx)
    for i, im in enumerate(imgs):
        batch.upload(i, im)
    th, _, _ = batch.autotune_tiles(steps=5)
    best, tms = batch.autotune(steps=5)
    out = {"frames": 64, "graph_branches": best, "tile_heights": th}
    for name, mode in (("graph", orb.Batch.GRAPH), ("per_frame_launches", orb.Batch.PER_IMAGE)):
        for _ in range(3):
            batch.run(mode)
        batch.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run(mode)
        batch.synchronize()
        out[f"{name}_frames_per_s"] = round(64 * steps / (time.perf_counter() - t0), 1)
    out["graph_over_per_frame"] = round(out["graph_frames_per_s"] / out["per_frame_launches_frames_per_s"], 2)
    return out


def serving_extra(orb, ctx):
    """16 camera streams (4 each of 640x480, 752x480, 1241x376, 1280x720),
    Poisson arrivals; extract + match against the stream's resident previous
    descriptors.  Also the plain per-frame GPU path on the same frames."""
    from orb_slam2_detailed_comments_amd import serving
    streams = serving.MIXED_16
    frames = serving.make_stream_frames(streams, 8, seed=300)
    srv = orb.Server(streams, max_inflight=3, context=ctx)
    out = {"workload": "16 streams (4x 640x480, 4x 752x480, 4x 1241x376, 4x 1280x720), Poisson arrivals, "
                       "extract + match vs the stream's previous frame (TH_LOW 50, ratio 0.6); 8 frames per stream "
                       "replayed back and forth, so each match pairs sequence neighbours"}
    recs, _ = serving.run(srv, frames, fps=30.0, duration=3.0, seed=1)
    out["at_30fps_per_stream"] = serving.summarize(recs)
    print(f"[bench] serving 30 fps/stream: {out['at_30fps_per_stream']}", file=sys.stderr, flush=True)
    sweep = []
    for fps in (120.0, 240.0, 480.0, 960.0, 1920.0, 3840.0):
        recs, _ = serving.run(srv, frames, fps=fps, duration=1.0, seed=2)
        r = serving.summarize(recs)
        r["offered_fps_per_stream"] = fps
        sweep.append(r)
        print(f"[bench] serving {fps} fps/stream: {r}", file=sys.stderr, flush=True)
    out["sweep"] = sweep
    ok = [r for r in sweep if r["p99_ms"] <= 10.0 and r["errors"] == 0 and r["refused"] == 0]
    out["max_frames_per_s_at_p99_le_10ms"] = max((r["frames_per_s"] for r in ok), default=None)
    # The same schedule shape from native threads (one submitter per stream,
    # orb_server_replay), without the Python driver's per-frame cost.
    native = []
    for fps in (960.0, 1440.0, 1920.0, 2880.0):
        r = srv.replay(frames, fps, 1.0, seed=4)
        r["offered_fps_per_stream"] = fps
        native.append(r)
        print(f"[bench] serving, native driver, {fps} fps/stream: {r}", file=sys.stderr, flush=True)
    out["native_driver_sweep"] = native
    okn = [r for r in native if r["p99_ms"] <= 10.0 and r["errors"] == 0 and r["refused"] == 0]
    out["native_max_frames_per_s_at_p99_le_10ms"] = max((r["frames_per_s"] for r in okn), default=None)
    out["server_stats"] = srv.stats()
    srv.close()
    # plain per-frame GPU path: one ORBextractor call and one match per frame
    exs = [orb.ORBextractor(nf, 1.2, 8, 20, 7, context=ctx) for (_, _, nf) in streams]
    prev = [None] * len(streams)
    for i, f in enumerate(frames):       # warm-up: one frame per stream
        prev[i] = exs[i](f[0])[1]
    n, t0 = 0, time.perf_counter()
    for k in range(1, len(frames[0])):
        for i, f in enumerate(frames):
            _, d = exs[i](f[k])
            orb.match_bf(d, prev[i], 50, 0.6, context=ctx)
            prev[i] = d
            n += 1
    plain = n / (time.perf_counter() - t0)
    out["plain_per_frame_frames_per_s"] = round(plain, 1)
    if out["max_frames_per_s_at_p99_le_10ms"]:
        out["served_over_plain"] = round(out["max_frames_per_s_at_p99_le_10ms"] / plain, 2)
    return out


# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz; a SIMD issues one wave64 VALU
# instruction every 4 cycles (16 lanes wide), so the VALU issue ceiling is
# 256 * 4 * 2.4e9 / 4 wave-instructions per second.
VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4


def committed_counters(stage, stage_ms, hbm_frac, n_images):
    """Counter-derived fields for the dominant stage, from the summaries the
    rocprofv3 runs left in profiles/ (counters are not collected while
    timing): L2-to-fabric traffic (FETCH_SIZE + WRITE_SIZE, which also counts
    Infinity Cache hits, so it bounds HBM bytes from above), and the VALU
    issue fraction = SQ_INSTS_VALU per launch x 4 cycles over this run's
    launch time on 1024 SIMDs at 2.4 GHz, beside the profiler's VALUBusy.
    Both counts are per-tile sums over launches of images_per_launch images
    of this size (the file says how many), so they are scaled to this run's
    n_images."""
    out = {}
    tfile = os.path.join(ROOT, "profiles", f"traffic_{stage}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        scale = n_images / t.get("images_per_launch", 64)
        out["traffic"] = int(round(t.get("bytes_per_launch") * scale))
        n_prof = t.get("images_per_launch", 64)
        out["traffic_kind"] = (f"L2 egress bytes (FETCH_SIZE+WRITE_SIZE, calibrated; an upper bound on HBM bytes), "
                               f"profiles/traffic_{stage}.json, its {n_prof}-image launch"
                               + ("" if n_prof == n_images else f" scaled to {n_images} images"))
    vfile = os.path.join(ROOT, "profiles", f"valu_{stage}.json")
    if os.path.exists(vfile):
        with open(vfile) as f:
            v = json.load(f)
        insts = int(round(v["valu_insts_per_launch"] * n_images / v.get("images_per_launch", 64)))
        frac = insts / (stage_ms * 1e-3) / VALU_WAVE_INSTS_PER_S
        out["valu"] = {"issue_frac": round(frac, 4), "valu_insts_per_launch": insts,
                       "cycles_per_inst": 4, "profiled_valu_busy": v.get("valu_busy"),
                       "source": f"profiles/valu_{stage}.json"}
        if hbm_frac is not None and frac > hbm_frac:
            out["bound"] = "valu"
    return out


def parity_sample(pyoracle, frames, gpu):
    """The oracle as checker on a sample of the timed batch: for each sampled
    pair p, gpu[p] = (kl, dl, kr, dr, uRight, depth) read back after the
    timed loop; keypoints, descriptors, uRight and depth must be bit-exact."""
    b = float(np.float32(BF) / np.float32(FX))
    L = pyoracle.OracleExtractor(NF, 1.2, NLEV, 20, 7)
    R = pyoracle.OracleExtractor(NF, 1.2, NLEV, 20, 7)
    bad, nkp, ndesc, amax = [], 0, 0, 0.0
    for p, (kl, dl, kr, dr, uR, dep) in sorted(gpu.items()):
        okl, odl, okr, odr, ouR, odep = pyoracle.stereo_frame(L, R, frames[p][0], frames[p][1], BF, b)
        same = (len(kl) == len(okl) and len(kr) == len(okr) and
                kl.tobytes() == okl.tobytes() and kr.tobytes() == okr.tobytes() and
                np.array_equal(dl, odl) and np.array_equal(dr, odr) and
                uR.tobytes() == ouR.tobytes() and dep.tobytes() == odep.tobytes())
        nkp += len(okl) + len(okr)
        for k, ok, d, od in ((kl, okl, dl, odl), (kr, okr, dr, odr)):
            if len(k) == len(ok):
                # descriptors that differ (the straddle count: rows whose
                # float angle falls on the other side of a pattern rounding
                # boundary) and the largest angle difference
                ndesc += int(np.count_nonzero(np.any(d != od, axis=1)))
                if len(k):
                    amax = max(amax, float(np.max(np.abs(k["angle"].astype(np.float64) - ok["angle"]))))
        if not same:
            bad.append(p)
    return {"pairs": sorted(gpu), "images": 2 * len(gpu), "keypoints": nkp,
            "fields": "keypoints (all 28 bytes), descriptors, uRight, depth", "bit_exact": not bad,
            "mismatched_pairs": bad, "descriptor_straddles": ndesc, "angle_max_abs_diff_deg": amax}


def cpu_baseline(frames, seconds, gpu_sample=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = len(os.sched_getaffinity(0))
    b = np.float32(BF) / np.float32(FX)
    parity = parity_sample(pyoracle, frames, gpu_sample) if gpu_sample else None

    def pipeline(stop, counter, idx0):
        L = pyoracle.OracleExtractor(NF, 1.2, NLEV, 20, 7)
        R = pyoracle.OracleExtractor(NF, 1.2, NLEV, 20, 7)
        i = idx0
        while not stop.is_set():
            imL, imR = frames[i % min(len(frames), 8)]
            pyoracle.stereo_frame(L, R, imL, imR, BF, float(b))
            counter.append(1)
            i += 1

    def run(npipe, secs):
        stop, counter = threading.Event(), []
        th = [threading.Thread(target=pipeline, args=(stop, counter, k)) for k in range(npipe)]
        for t in th:
            t.start()
        time.sleep(0.5)  # warm-up
        c0, t0 = len(counter), time.perf_counter()
        time.sleep(secs)
        c1, t1 = len(counter), time.perf_counter()
        stop.set()
        for t in th:
            t.join()
        return (c1 - c0) / (t1 - t0)

    single = run(1, seconds * 0.3)
    npipe = max(1, cores // 2)
    allc = run(npipe, seconds * 0.7)
    return {"value": round(allc, 2), "unit": "frames/s", "cores": 2 * npipe, "kind": "port",
            "sample": f"{min(len(frames), 8)} synthetic 1241x376 stereo frames cycled for ~{seconds:.0f}s: "
                      f"{npipe} independent Frame pipelines (L/R extraction on 2 threads + stereo match) on "
                      f"{cores} host cores",
            "reference_structure_frames_per_s": round(single, 2), "host_cores": cores,
            "cpu_model": cpu_model(), "build": "oracle/Makefile: g++ -O3 -march=x86-64-v3 -ffp-contract=off"}, parity


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64, help="stereo frames per step")
    ap.add_argument("--mode", choices=["graph", "stages", "per-image", "streams"], default="graph")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip every extra workload")
    ap.add_argument("--extras", default="all",
                    help="comma-separated extras to run: knn, batch, serving, bow, mappoint, input, frame, "
                         "winmatch, projection")
    ap.add_argument("--tiles", default="auto",
                    help="level kernel tile heights: auto (the tile autotuner), 32, 16, or 8 comma-separated values")
    ap.add_argument("--branches", type=int, default=0,
                    help="independent hipGraph branches (mode graph); 0: pick by the batch's autotuner")
    ap.add_argument("--pipeline", type=int, default=0,
                    help="batches in flight: step i runs batch i mod P, each batch on its own stream, so one "
                         "step's last launches overlap the next step's first ones; 0: pick P (1-3) together with "
                         "the graph branch count by timing them")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (--gpus 1): frames and streams shard across GPUs with no collective, "
              "so the single-GPU figure is the whole measurement.", file=sys.stderr)
        sys.exit(2)

    import orb_slam2_detailed_comments_amd as orb
    mode = {"graph": orb.Batch.GRAPH, "stages": orb.Batch.STAGE_LAUNCHES, "per-image": orb.Batch.PER_IMAGE,
            "streams": orb.Batch.STREAMS}[args.mode]
    B = args.batch
    auto_pipe = args.pipeline <= 0 and mode == orb.Batch.GRAPH
    P = 3 if auto_pipe else max(1, args.pipeline)
    frames = make_frames(B * P)
    print(f"[bench] {B * P} stereo frames {W}x{H} generated", file=sys.stderr, flush=True)
    b = float(np.float32(BF) / np.float32(FX))
    ctxs, batches = [], []
    for k in range(P):
        # one context (and so one stream) per batch in flight
        ctxs.append(orb.Context(0))
        bk = orb.Batch([(W, H)] * (2 * B), NF, 1.2, NLEV, 20, 7, context=ctxs[-1])
        for i, (l, r) in enumerate(frames[k * B:(k + 1) * B]):
            bk.upload(2 * i, l)
            bk.upload(2 * i + 1, r)
        bk.set_stereo(list(range(0, 2 * B, 2)), list(range(1, 2 * B, 2)), BF, b)
        batches.append(bk)
    ctx, batch = ctxs[0], batches[0]
    tiles = None
    if args.tiles == "auto":
        th, m32, m16 = batch.autotune_tiles(steps=10)
        tiles = {"tile_heights": th, "ms_64x32": {str(k): round(v, 4) for k, v in m32.items()},
                 "ms_64x16": {str(k): round(v, 4) for k, v in m16.items()}}
        print(f"[bench] tile autotune: {tiles}", file=sys.stderr, flush=True)
    else:
        th = [int(v) for v in args.tiles.split(",")] if "," in args.tiles else [int(args.tiles)] * NLEV
        batch.set_tile_heights(th)
        tiles = {"tile_heights": th}
    tuned = None
    for bk in batches[1:]:
        bk.set_tile_heights(th)
    if auto_pipe:
        # Batches in flight and graph branches picked together: 20 steps of
        # each (P, branches) after a warm-up, the fastest kept (the smaller
        # P, then the fewer branches, on a tie within 1 %).
        cand = [(p_, br) for p_ in (1, 2, 3) for br in ((1, 2, 3) if args.branches <= 0 else (args.branches,))]
        ms = {}
        for p_, br in cand:
            for bk in batches[:p_]:
                bk.set_branches(br)
            for i in range(2 * p_):
                batches[i % p_].run(mode)
            for bk in batches[:p_]:
                bk.synchronize()
            t0 = time.perf_counter()
            for i in range(20):
                batches[i % p_].run(mode)
            for bk in batches[:p_]:
                bk.synchronize()
            ms[(p_, br)] = (time.perf_counter() - t0) * 1e3 / 20
        fastest = min(ms.values())
        P, args.branches = min((k for k in cand if ms[k] <= fastest * 1.01), key=lambda k: (k[0], k[1]))
        for bk in batches[:P]:
            bk.set_branches(args.branches)
        tuned = {"chosen": {"batches_in_flight": P, "branches": args.branches},
                 "ms_per_step": {f"{p_}x{br}": round(v, 4) for (p_, br), v in ms.items()}}
        print(f"[bench] autotuned pipeline: {tuned}", file=sys.stderr, flush=True)
    elif args.branches > 0:
        for bk in batches:
            bk.set_branches(args.branches)
    elif mode == orb.Batch.GRAPH:
        best, tms = batch.autotune(steps=10)
        tuned = {"chosen": best, "ms_per_replay": {str(k): round(v, 4) for k, v in tms.items()}}
        args.branches = best
        for bk in batches[1:]:
            bk.set_branches(best)
        print(f"[bench] autotuned graph branches: {best} ({tuned['ms_per_replay']})", file=sys.stderr, flush=True)

    for i in range(args.warmup * P):
        batches[i % P].run(mode)
    for bk in batches:
        bk.synchronize()
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    for i in range(args.steps):
        batches[i % P].run(mode)
    for bk in batches:
        bk.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # Result sizes of the last step (keypoints and FAST candidates per image),
    # and the outputs of six pairs spread over the batch (and so over the
    # graph's branches), kept for the oracle's check after the CPU timing.
    kps = [len(batch.results(s)[0]) for s in range(2 * B)]
    matched = [int((batch.stereo_results(p)[0] >= 0).sum()) for p in range(B)]
    sample = sorted({round(i * (B - 1) / 5) for i in range(6)})
    gpu_sample = {p: batch.results(2 * p) + batch.results(2 * p + 1) + batch.stereo_results(p) for p in sample}
    for k in range(1, P):           # two pairs of every other batch in flight
        for p in (1, B - 2):
            bk = batches[k]
            gpu_sample[k * B + p] = bk.results(2 * p) + bk.results(2 * p + 1) + bk.stereo_results(p)

    # Host to host: the same step with the 64 images copied in from host
    # memory and every result read back (one step at a time, no overlap of
    # copies with compute), outside the headline figure.
    imgs = [im for pair in frames[:B] for im in pair]
    res = None
    for _ in range(2):
        batch.upload_all(imgs)
        batch.run(mode)
        res = batch.read_all(out=res)
    h2h_steps = max(5, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(h2h_steps):
        batch.upload_all(imgs)
        batch.run(mode)
        res = batch.read_all(out=res)
    h2h_s = time.perf_counter() - t0
    host_to_host = {"frames_per_s": round(B * h2h_steps / h2h_s, 1), "ms_per_step": round(h2h_s * 1e3 / h2h_steps, 3),
                    "steps": h2h_steps,
                    "includes": "copies of the 64 host images (pageable memory), graph replay, read-back of every "
                                "slot's keypoints + descriptors and every pair's uRight + depth; no copy/compute "
                                "overlap"}
    print(f"[bench] host to host: {host_to_host}", file=sys.stderr, flush=True)

    # Host to host through the pinned transfers, with every batch created
    # above in flight (the copies, not the kernels, set this rate, so it
    # gains from more batches than the headline picks): the frames sit in
    # each batch's pinned image buffers (where a capture or decode step would
    # write them), and every step enqueues one upload, the graph and one
    # gathered read of all results on its batch's stream; a batch is waited
    # for (its previous step's results in pinned memory) before its next step
    # is enqueued.
    PH = len(batches)
    for k in range(PH):
        views = batches[k].pinned_images()
        for i, (l, r) in enumerate(frames[k * B:(k + 1) * B]):
            views[2 * i][:] = l
            views[2 * i + 1][:] = r

    def pinned_steps(n):
        for i in range(n):
            bk = batches[i % PH]
            if i >= PH:
                bk.synchronize()
            bk.upload_pinned()
            bk.run(mode)
            bk.read_async()
        for bk in batches:
            bk.synchronize()

    pinned_steps(2 * PH)
    pin_steps = max(20, args.steps)
    t0 = time.perf_counter()
    pinned_steps(pin_steps)
    pin_s = time.perf_counter() - t0
    # the pinned copies of the sampled pairs equal the device results read
    # above (which the oracle checks below)
    pin_equal = True
    for key, ref in gpu_sample.items():
        bk, p = batches[key // B], key % B
        got = bk.pinned_results(2 * p) + bk.pinned_results(2 * p + 1) + bk.pinned_stereo(p)
        pin_equal &= all(np.array_equal(np.asarray(a).view(np.uint8), np.asarray(g).view(np.uint8))
                         for a, g in zip(ref, got))
    host_to_host["pinned_pipelined"] = {
        "frames_per_s": round(B * pin_steps / pin_s, 1), "ms_per_step": round(pin_s * 1e3 / pin_steps, 3),
        "steps": pin_steps, "batches_in_flight": PH, "sampled_pairs_equal_device_results": bool(pin_equal),
        "includes": "host-to-device copy of the 64 frames from pinned buffers, graph replay, gather of every "
                    "slot's keypoints + descriptors and every pair's uRight + depth, one device-to-host copy "
                    "into pinned memory; copies overlap other batches' compute"}
    print(f"[bench] host to host, pinned: {host_to_host['pinned_pipelined']}", file=sys.stderr, flush=True)

    stages = batch.stage_times(args.steps)
    # The dominant kernel is the one with the largest share of the step: the
    # level kernel's eight launches against the other stages' one each.  For
    # the level kernel the roofline is that of its largest launch, level 0.
    fam = {"pyramid": sum(v for k, v in stages.items() if k.startswith("pyramid_"))}
    fam.update({k: v for k, v in stages.items() if not k.startswith("pyramid_") and k != "stereo_median"})
    top = max(fam, key=fam.get)
    dom = "pyramid_l0" if top == "pyramid" else top
    cands = None
    bytes_dom = stage_bytes(dom, 2 * B, float(np.mean(kps)), 10 * NF if cands is None else cands)
    achieved = bytes_dom / (stages[dom] * 1e-3) / 1e9 if bytes_dom else None
    roofline = {"bound": "hbm", "kernel": dom, "kernel_share_of_stage_time": round(fam[top] / sum(fam.values()), 3),
                "achieved": round(achieved, 2) if achieved else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
                "traffic": None, "avg_launch_ms": round(stages[dom], 4),
                "algorithmic_bytes_per_launch": bytes_dom,
                "stage_ms": {k: round(v, 4) for k, v in stages.items()}}
    roofline.update(committed_counters(dom, stages[dom], roofline["frac"], 2 * B))
    out = {"metric": METRIC, "value": round(value, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8",
           "data": "synthetic (seeded shapes+noise frames; right view = left through a 12-plane disparity map)",
           "config": {"workload": f"stereo extract+match, {W}x{H}, nFeatures {NF}, 8 levels, scale 1.2, FAST 20/7, "
                                  f"bf {BF}, fx {FX}", "frames_per_step": B, "images_per_step": 2 * B,
                      "run_mode": args.mode, "graph_branches": args.branches, "pipeline_autotune": tuned,
                      "batches_in_flight": P,
                      "tile_autotune": tiles,
                      "mean_keypoints_per_image": round(float(np.mean(kps)), 1),
                      "mean_stereo_matches": round(float(np.mean(matched)), 1)},
           "roofline": roofline, "host_to_host": host_to_host}
    want = set() if args.no_extras else (
        {"knn", "batch", "serving", "bow", "mappoint", "input", "frame", "winmatch", "projection", "hbm"}
        if args.extras == "all"
        else set(args.extras.split(",")))
    cpu_s = 0.0 if args.no_cpu else min(args.cpu_seconds, 3.0)
    if "hbm" in want:
        import ctypes
        ms = ctypes.c_float()
        nbytes = 1 << 30
        if orb.lib().orb_hbm_copy_time(ctx.h, nbytes, 10, ctypes.byref(ms)) == 0:
            out["hbm_copy_microbench"] = {"bytes_moved": 2 * nbytes, "ms": round(ms.value, 4),
                                          "GB_per_s": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}
    if "knn" in want:
        print("[bench] dense k=2 search", file=sys.stderr, flush=True)
        out["knn2"] = knn_extra(orb, ctx, max(10, args.steps))
    if "batch" in want:
        print("[bench] mixed-size batch", file=sys.stderr, flush=True)
        out["mixed_batch"] = batch_extra(orb, ctx, max(5, args.steps // 2))
        print(f"[bench] mixed-size batch: {out['mixed_batch']}", file=sys.stderr, flush=True)
    if "serving" in want:
        print("[bench] serving", file=sys.stderr, flush=True)
        out["serving"] = serving_extra(orb, ctx)
    if "bow" in want:
        print("[bench] vocabulary transform", file=sys.stderr, flush=True)
        out["bow"] = bow_extra(orb, ctx, max(args.steps, 5), 0.0 if args.no_cpu else min(args.cpu_seconds, 5.0))
    if "mappoint" in want:
        print("[bench] map-point descriptors", file=sys.stderr, flush=True)
        out["distinctive_descriptors"] = mappoint_extra(orb, ctx, max(args.steps, 5), cpu_s)
    if "input" in want:
        print("[bench] input-side image ops", file=sys.stderr, flush=True)
        out["input_ops"] = input_extra(orb, ctx, max(args.steps, 20), min(cpu_s, 3.0))
    if "frame" in want:
        print("[bench] frame keypoint ops", file=sys.stderr, flush=True)
        out["frame_ops"] = frame_ops_extra(orb, ctx, max(args.steps, 20), min(cpu_s, 3.0))
    if "winmatch" in want:
        print("[bench] windowed matching", file=sys.stderr, flush=True)
        out["search_for_initialization"] = winmatch_extra(orb, ctx, cpu_s)
    if "projection" in want:
        print("[bench] search by projection", file=sys.stderr, flush=True)
        out["search_by_projection"] = projection_extra(orb, ctx, cpu_s)
        print(f"[bench] search by projection: {out['search_by_projection']}", file=sys.stderr, flush=True)
    out["parity_checked"] = None
    if not args.no_cpu:
        print("[bench] checking sampled pairs against the oracle, then timing it", file=sys.stderr, flush=True)
        cpu, parity = cpu_baseline(frames, args.cpu_seconds, gpu_sample)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1) if cpu["value"] else None
        out["parity_checked"] = parity
        print(f"[bench] parity: {parity}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
