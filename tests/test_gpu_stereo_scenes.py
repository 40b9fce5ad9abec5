"""GPU ComputeStereoMatches on the hand-built scenes of stereo_scenes.py.

Each scene puts keypoints where one rule of Frame::ComputeStereoMatches
(Frame.cc:831-1082) decides the outcome: the row band and truncated row, the
octave and uR windows, the Hamming acceptance and tie order, this fork's
iniu/endu bounds, rejection at +-L, the disparity-0 step and the median cut
with a SAD exactly at the threshold.  uRight and depth must equal the
oracle's, and the independent statement's, bit for bit.
"""
import numpy as np
import pytest

import stereo_scenes as ss

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", [n for n, _ in ss.SCENES])
def test_stereo_scene_bit_exact(orb, oracle, name):
    scene = dict(ss.SCENES)[name]()
    levels = 3
    gl = orb.ORBextractor(500, 1.2, levels, 20, 7)
    gr = orb.ORBextractor(500, 1.2, levels, 20, 7)
    gl(scene.left)
    gr(scene.right)
    ol = oracle.OracleExtractor(500, 1.2, levels, 20, 7)
    orr = oracle.OracleExtractor(500, 1.2, levels, 20, 7)
    ol(scene.left)
    orr(scene.right)
    kl, dl, kr, dr = scene.keypoints(orb.KEYPOINT_DTYPE)
    uR, dep = orb.compute_stereo_matches(gl, gr, kl, dl, kr, dr, scene.bf, scene.b)
    ouR, odep = oracle.stereo_match(ol, orr, kl, dl, kr, dr, scene.bf, scene.b)
    t = ol.tables()
    wu, wd = ss.compute_stereo_matches([ol.level_image(l) for l in range(levels)],
                                       [orr.level_image(l) for l in range(levels)],
                                       t["scale"], t["inv_scale"], kl, dl, kr, dr, scene.bf, scene.b)
    np.testing.assert_array_equal(ouR.view(np.uint32), wu.view(np.uint32))
    np.testing.assert_array_equal(uR.view(np.uint32), ouR.view(np.uint32), err_msg="uRight")
    np.testing.assert_array_equal(dep.view(np.uint32), odep.view(np.uint32), err_msg="depth")
    assert (uR >= 0).any() or name == "median_zero"
