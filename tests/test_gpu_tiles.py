"""Tile shapes of the level kernel: 64 x 16 tiles give the same pyramid,
blur, candidates, keypoints and descriptors as 64 x 32 tiles, at every level
and size, alone, mixed per level, in batches, and after the autotuner."""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,nf", [(640, 480, 1000), (1241, 376, 2000), (752, 480, 1000), (1280, 720, 2000)])
def test_tile_16_equals_oracle(orb, oracle, w, h, nf):
    img = synth.shapes_frame(w, h, 70)
    g = orb.ORBextractor(nf, 1.2, 8, 20, 7)
    g.set_tile_heights([16] * 8)
    gk, gd = g(img)
    o = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)
    ok, od = o(img)
    for level in range(8):
        np.testing.assert_array_equal(g.pyramid_level(level), o.level_image(level), err_msg=f"level {level}")
        np.testing.assert_array_equal(g.debug_image(level, 1), o.level_image(level, blurred=True),
                                      err_msg=f"blur {level}")
        np.testing.assert_array_equal(g.debug_candidates(level), o.candidates(level), err_msg=f"cands {level}")
    assert gk.tobytes() == ok.tobytes() and np.array_equal(gd, od)


def test_tile_heights_mixed_per_level_and_scales(orb, oracle):
    img = synth.shapes_frame(1241, 376, 71)
    for sf, nl, th in [(1.2, 8, [32, 16, 32, 16, 16, 32, 16, 16]), (2.0, 3, [16, 32, 16]), (2.5, 2, [16, 16])]:
        nf = {1.2: 2000, 2.0: 1500, 2.5: 1200}[sf]
        g = orb.ORBextractor(nf, sf, nl, 20, 7)
        g.set_tile_heights(th)
        gk, gd = g(img)
        o = oracle.OracleExtractor(nf, sf, nl, 20, 7)
        ok, od = o(img)
        assert gk.tobytes() == ok.tobytes() and np.array_equal(gd, od), (sf, nl)


def test_stereo_batch_tile_autotune_keeps_results(orb, oracle):
    bf, fx = 386.1, 718.856
    b = float(np.float32(bf) / np.float32(fx))
    frames = [synth.stereo_pair(1241, 376, 1100 + i)[:2] for i in range(8)]
    batch = orb.Batch([(1241, 376)] * 16, 2000)
    for i, (l, r) in enumerate(frames):
        batch.upload(2 * i, l)
        batch.upload(2 * i + 1, r)
    batch.set_stereo(list(range(0, 16, 2)), list(range(1, 16, 2)), bf, b)
    th, ms32, ms16 = batch.autotune_tiles(steps=3)
    assert set(th) <= {16, 32} and len(th) == 8
    batch.autotune(steps=3)
    for _ in range(2):
        batch.run(1)
    batch.synchronize()
    ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    for p in (0, 5):
        okl, odl, okr, odr, ouR, odep = oracle.stereo_frame(ol, orr, frames[p][0], frames[p][1], bf, b)
        gkl, gdl = batch.results(2 * p)
        uR, dep = batch.stereo_results(p)
        assert gkl.tobytes() == okl.tobytes() and np.array_equal(gdl, odl)
        np.testing.assert_array_equal(uR.view(np.uint32), ouR.view(np.uint32))
    # every level at 16, then every level at 32 again: identical results
    batch.set_tile_heights([16] * 8)
    batch.run(1)
    a = batch.read_all()
    batch.set_tile_heights([32] * 8)
    batch.run(1)
    c = batch.read_all()
    assert a["kps"].tobytes() == c["kps"].tobytes() and np.array_equal(a["desc"], c["desc"])
    np.testing.assert_array_equal(a["uR"].view(np.uint32), c["uR"].view(np.uint32))


def test_stereo_frame_needs_matching_tile_heights(orb):
    L, R, _ = synth.stereo_pair(640, 480, 3)
    gl = orb.ORBextractor(1000, 1.2, 8, 20, 7)
    gr = orb.ORBextractor(1000, 1.2, 8, 20, 7)
    gl.set_tile_heights([16] * 8)
    with pytest.raises(orb.OrbError):
        orb.stereo_frame(gl, gr, L, R, 386.1, 0.537)
    gr.set_tile_heights([16] * 8)
    kl, dl, kr, dr, uR, dep = orb.stereo_frame(gl, gr, L, R, 386.1, 0.537)
    assert len(kl) > 0
