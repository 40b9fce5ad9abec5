"""GPU colour input: the device cvtColor equals the oracle's for 3/4
channels in both orders, and extracting a colour frame equals extracting the
oracle's gray conversion of it, bit for bit."""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,c,rgb", [(641, 480, 3, True), (752, 480, 4, False), (1241, 376, 3, False),
                                       (33, 7, 4, True)])
def test_cvt_color_matches_oracle(orb, oracle, w, h, c, rgb):
    rng = np.random.default_rng(w + c)
    im = rng.integers(0, 256, size=(h, w, c), dtype=np.uint8)
    np.testing.assert_array_equal(orb.cvt_color_gray(im, rgb), oracle.cvt_gray(im, rgb))


def test_extract_color_equals_gray_path(orb, oracle):
    gray = synth.shapes_frame(1241, 376, 5)
    rng = np.random.default_rng(3)
    # a colour frame whose conversion is not simply `gray`: per-channel offsets
    im = np.stack([gray, np.clip(gray.astype(int) + rng.integers(-20, 21, gray.shape), 0, 255),
                   255 - gray], axis=2).astype(np.uint8)
    ex = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    gk, gd = ex.extract_color(im, rgb=False)
    ref = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    rk, rd = ref(oracle.cvt_gray(im, rgb=False))
    np.testing.assert_array_equal(gk.view(np.uint8), rk.view(np.uint8))
    np.testing.assert_array_equal(gd, rd)
    ok, od = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)(oracle.cvt_gray(im, rgb=False))
    np.testing.assert_array_equal(gk.view(np.uint8), ok.view(np.uint8))
