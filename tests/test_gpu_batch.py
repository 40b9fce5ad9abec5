"""Batched extraction: every slot of a mixed-size batch equals the oracle, in
every run mode (stage launches, hipGraph replay, per-image launches, and
the graph's branches launched directly on their streams)."""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(640, 480, 1000), (752, 480, 1000), (1241, 376, 2000), (1280, 720, 2000)]


@pytest.fixture(scope="module")
def mixed(orb, oracle):
    slots = [SIZES[i % 4] for i in range(8)]
    imgs = [synth.shapes_frame(w, h, 500 + i) for i, (w, h, _) in enumerate(slots)]
    batch = orb.Batch([(w, h) for w, h, _ in slots], [nf for _, _, nf in slots])
    for i, im in enumerate(imgs):
        batch.upload(i, im)
    ref = []
    for (w, h, nf), im in zip(slots, imgs):
        o = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)
        ref.append(o(im))
    return batch, ref


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_batch_modes_match_oracle(mixed, mode):
    batch, ref = mixed
    for _ in range(2):   # the graph is replayed, not only captured
        batch.run(mode)
    batch.synchronize()
    for s, (ok, od) in enumerate(ref):
        gk, gd = batch.results(s)
        assert len(gk) == len(ok), (s, len(gk), len(ok))
        for f in ("x", "y", "size", "angle", "response", "octave", "class_id"):
            np.testing.assert_array_equal(gk[f], ok[f], err_msg=f"slot {s} {f}")
        np.testing.assert_array_equal(gd, od, err_msg=f"slot {s} descriptors")


def test_batch_stereo_pairs(orb, oracle):
    bf, fx = 386.1, 718.856
    b = float(np.float32(bf) / np.float32(fx))
    pairs = [synth.stereo_pair(1241, 376, 40 + i)[:2] for i in range(3)]
    batch = orb.Batch([(1241, 376)] * 6, 2000)
    for i, (l, r) in enumerate(pairs):
        batch.upload(2 * i, l)
        batch.upload(2 * i + 1, r)
    batch.set_stereo([0, 2, 4], [1, 3, 5], bf, b)
    for _ in range(3):
        batch.run(1)
    batch.synchronize()
    for i, (l, r) in enumerate(pairs):
        ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
        orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
        okl, odl, okr, odr, ouR, odep = oracle.stereo_frame(ol, orr, l, r, bf, b)
        uR, dep = batch.stereo_results(i)
        np.testing.assert_array_equal(uR.view(np.uint32), ouR.view(np.uint32))
        np.testing.assert_array_equal(dep.view(np.uint32), odep.view(np.uint32))


def test_stage_times_cover_pipeline(orb):
    batch = orb.Batch([(640, 480)] * 2, 1000)
    batch.upload(0, synth.shapes_frame(640, 480, 1))
    batch.upload(1, synth.shapes_frame(640, 480, 2))
    batch.set_stereo([0], [1], 386.1, 0.537)
    t = batch.stage_times(3)
    assert set(t) >= {"pyramid_l0", "octree", "describe", "stereo_match"}
    assert all(v >= 0 for v in t.values())


def test_batch_graph_branches_match_oracle(mixed):
    """Branches as graph replays (mode 1) and as direct launches on their own
    streams (mode 3), alternating on one batch."""
    batch, ref = mixed
    for k in (3, 8, 1, 4):
        batch.set_branches(k)
        for mode in (1, 3, 1):
            for _ in range(2):
                batch.run(mode)
            batch.synchronize()
            for s, (ok, od) in enumerate(ref):
                gk, gd = batch.results(s)
                assert len(gk) == len(ok), (k, mode, s, len(gk), len(ok))
                np.testing.assert_array_equal(gk.view(np.uint8), ok.view(np.uint8),
                                              err_msg=f"branches {k} mode {mode} slot {s}")
                np.testing.assert_array_equal(gd, od, err_msg=f"branches {k} mode {mode} slot {s} descriptors")


def test_batch_autotune_keeps_results(mixed):
    """The branch autotuner times every candidate count, picks one of them and
    leaves a graph that still gives the oracle's keypoints and descriptors."""
    batch, ref = mixed
    best, ms = batch.autotune(steps=3)
    assert sorted(ms) == [1, 2, 3, 4, 6, 8] and all(v > 0 for v in ms.values())
    assert best in ms and ms[best] <= min(ms.values()) / 0.99 + 1e-6
    batch.run(1)
    batch.synchronize()
    for s, (ok, od) in enumerate(ref):
        gk, gd = batch.results(s)
        np.testing.assert_array_equal(gk.view(np.uint8), ok.view(np.uint8), err_msg=f"autotuned slot {s}")
        np.testing.assert_array_equal(gd, od, err_msg=f"autotuned slot {s} descriptors")


@pytest.mark.parametrize("left,right", [([0, 2, 4], [1, 3, 5]), ([0, 1, 2], [5, 4, 3])])
def test_batch_stereo_branches(orb, oracle, left, right):
    """Pairs stay whole across branches; a nesting with no valid cut keeps
    one branch.  Either way the results are the single-branch ones."""
    bf, fx = 386.1, 718.856
    b = float(np.float32(bf) / np.float32(fx))
    pairs = [synth.stereo_pair(1241, 376, 60 + i)[:2] for i in range(3)]
    batch = orb.Batch([(1241, 376)] * 6, 2000)
    for i, (l, r) in enumerate(pairs):
        batch.upload(left[i], l)
        batch.upload(right[i], r)
    batch.set_stereo(left, right, bf, b)
    batch.run(0)
    batch.synchronize()
    want = [batch.stereo_results(i) for i in range(3)]
    batch.set_branches(2)
    for mode in (1, 3):
        for _ in range(3):
            batch.run(mode)
        batch.synchronize()
        for i in range(3):
            uR, dep = batch.stereo_results(i)
            np.testing.assert_array_equal(uR.view(np.uint32), want[i][0].view(np.uint32))
            np.testing.assert_array_equal(dep.view(np.uint32), want[i][1].view(np.uint32))
    l, r = pairs[0]
    ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    okl, odl, okr, odr, ouR, odep = oracle.stereo_frame(ol, orr, l, r, bf, b)
    np.testing.assert_array_equal(want[0][0].view(np.uint32), ouR.view(np.uint32))


def test_mixed_batch_64_frames_graph_autotuned(orb, oracle):
    """The mixed-size batch at its full size: 64 frames, 16 each of 640x480,
    752x480 (1000 features), 1241x376 and 1280x720 (2000 features), replayed
    from the graph with the autotuner's branch count; every slot bit-exact."""
    slots = [SIZES[i // 16] for i in range(64)]
    imgs = [synth.shapes_frame(w, h, 3000 + i) for i, (w, h, _) in enumerate(slots)]
    batch = orb.Batch([(w, h) for w, h, _ in slots], [nf for _, _, nf in slots])
    for i, im in enumerate(imgs):
        batch.upload(i, im)
    best, _ = batch.autotune(steps=3)
    for _ in range(3):
        batch.run(1)
    batch.synchronize()
    oracles = {nf: oracle.OracleExtractor(nf, 1.2, 8, 20, 7) for nf in (1000, 2000)}
    for s, ((w, h, nf), im) in enumerate(zip(slots, imgs)):
        ok, od = oracles[nf](im)
        gk, gd = batch.results(s)
        assert len(gk) == len(ok), (best, s, len(gk), len(ok))
        np.testing.assert_array_equal(gk.view(np.uint8), ok.view(np.uint8), err_msg=f"branches {best} slot {s}")
        np.testing.assert_array_equal(gd, od, err_msg=f"branches {best} slot {s} descriptors")


def test_stereo_batch_32_pairs_graph_autotuned(orb, oracle):
    """The timed stereo batch at its full size: 32 pairs of 1241x376 (the
    frames the benchmark generates), graph replay with the autotuned branch
    count; every pair's keypoints, descriptors, uRight and depth bit-exact."""
    bf, fx = 386.1, 718.856
    b = float(np.float32(bf) / np.float32(fx))
    frames = [synth.stereo_pair(1241, 376, 1000 + i)[:2] for i in range(32)]
    batch = orb.Batch([(1241, 376)] * 64, 2000)
    for i, (l, r) in enumerate(frames):
        batch.upload(2 * i, l)
        batch.upload(2 * i + 1, r)
    batch.set_stereo(list(range(0, 64, 2)), list(range(1, 64, 2)), bf, b)
    best, _ = batch.autotune(steps=3)
    for _ in range(3):
        batch.run(1)
    batch.synchronize()
    ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    for p, (l, r) in enumerate(frames):
        okl, odl, okr, odr, ouR, odep = oracle.stereo_frame(ol, orr, l, r, bf, b)
        gkl, gdl = batch.results(2 * p)
        gkr, gdr = batch.results(2 * p + 1)
        uR, dep = batch.stereo_results(p)
        assert gkl.tobytes() == okl.tobytes() and gkr.tobytes() == okr.tobytes(), (best, p)
        assert np.array_equal(gdl, odl) and np.array_equal(gdr, odr), (best, p)
        np.testing.assert_array_equal(uR.view(np.uint32), ouR.view(np.uint32), err_msg=f"pair {p}")
        np.testing.assert_array_equal(dep.view(np.uint32), odep.view(np.uint32), err_msg=f"pair {p}")


def test_whole_batch_transfers_equal_per_slot_calls(orb, oracle):
    """upload_all / read_all move the same bytes as per-slot upload / results."""
    bf, fx = 386.1, 718.856
    b = float(np.float32(bf) / np.float32(fx))
    pairs = [synth.stereo_pair(1241, 376, 80 + i)[:2] for i in range(3)]
    batch = orb.Batch([(1241, 376)] * 6, 2000)
    batch.set_stereo([0, 2, 4], [1, 3, 5], bf, b)
    batch.upload_all([im for p in pairs for im in p])
    batch.run(1)
    out = batch.read_all()
    for s in range(6):
        k, d = batch.results(s)
        n = out["counts"][s]
        assert n == len(k)
        assert out["kps"][s, :n].tobytes() == k.tobytes() and np.array_equal(out["desc"][s, :n], d)
    for p in range(3):
        uR, dep = batch.stereo_results(p)
        n = len(uR)
        np.testing.assert_array_equal(out["uR"][p, :n].view(np.uint32), uR.view(np.uint32))
        np.testing.assert_array_equal(out["depth"][p, :n].view(np.uint32), dep.view(np.uint32))
    ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    okl, _, _, _, ouR, _ = oracle.stereo_frame(ol, orr, pairs[1][0], pairs[1][1], bf, b)
    assert out["kps"][2, :len(okl)].tobytes() == okl.tobytes()
    np.testing.assert_array_equal(out["uR"][1, :len(ouR)].view(np.uint32), ouR.view(np.uint32))
