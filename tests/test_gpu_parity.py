"""GPU parity: every stage of the HIP path against the CPU oracle.

Bar (from the north star): pyramid, keypoints (x, y, octave, response),
Hamming distances and match indices bit-exact; angles within 1e-3 deg (here
they are compared bit-exactly too, since the kernel evaluates the same float
statements); descriptors bit-exact.
"""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import synth

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "size", "response", "octave", "class_id"]


def _pair(orb, oracle, nf, img, levels=8):
    g = orb.ORBextractor(nf, 1.2, levels, 20, 7)
    o = oracle.OracleExtractor(nf, 1.2, levels, 20, 7)
    gk, gd = g(img)
    ok, od = o(img)
    return g, o, gk, gd, ok, od


def _assert_same_keypoints(gk, gd, ok, od):
    assert len(gk) == len(ok), (len(gk), len(ok))
    for f in FIELDS:
        np.testing.assert_array_equal(gk[f], ok[f], err_msg=f)
    np.testing.assert_array_equal(gk["angle"].view(np.uint32), ok["angle"].view(np.uint32), err_msg="angle")
    assert np.max(np.abs(gk["angle"] - ok["angle"]), initial=0.0) <= 1e-3
    np.testing.assert_array_equal(gd, od)


@pytest.mark.parametrize("w,h,nf,seed", [(640, 480, 1000, 1), (1241, 376, 2000, 5), (752, 480, 1000, 9),
                                         (1280, 720, 2000, 13)])
def test_pyramid_blur_candidates(orb, oracle, w, h, nf, seed):
    img = synth.shapes_frame(w, h, seed)
    g, o, gk, gd, ok, od = _pair(orb, oracle, nf, img)
    for level in range(8):
        assert g.pyramid_level(level).shape == o.level_image(level).shape
        np.testing.assert_array_equal(g.pyramid_level(level), o.level_image(level), err_msg=f"level {level}")
        np.testing.assert_array_equal(g.debug_image(level, 1), o.level_image(level, blurred=True),
                                      err_msg=f"blur {level}")
        np.testing.assert_array_equal(g.debug_candidates(level), o.candidates(level), err_msg=f"cands {level}")


@pytest.mark.parametrize("w,h,nf,seed", [(640, 480, 1000, 1), (640, 480, 1000, 2), (1241, 376, 2000, 5),
                                         (752, 480, 1000, 9), (1280, 720, 2000, 13)])
def test_extract_bit_exact(orb, oracle, w, h, nf, seed):
    img = synth.shapes_frame(w, h, seed)
    _, _, gk, gd, ok, od = _pair(orb, oracle, nf, img)
    _assert_same_keypoints(gk, gd, ok, od)


def test_extract_warped_second_frame(orb, oracle):
    _, b = synth.frame_pair(640, 480, 3)
    _, _, gk, gd, ok, od = _pair(orb, oracle, 1000, b)
    _assert_same_keypoints(gk, gd, ok, od)


def test_extract_repeat_is_identical(orb):
    img = synth.shapes_frame(640, 480, 4)
    g = orb.ORBextractor(1000, 1.2, 8, 20, 7)
    k1, d1 = g(img)
    k2, d2 = g(img)
    np.testing.assert_array_equal(k1.view(np.uint8), k2.view(np.uint8))
    np.testing.assert_array_equal(d1, d2)


def test_flat_and_empty_images(orb, oracle):
    g = orb.ORBextractor(1000, 1.2, 8, 20, 7)
    k, d = g(np.full((480, 640), 77, np.uint8))
    assert len(k) == 0 and d.shape == (0, 32)
    k, d = g(np.zeros((0, 0), np.uint8))
    assert len(k) == 0
    o = oracle.OracleExtractor(1000, 1.2, 8, 20, 7)
    ok, _ = o(np.full((480, 640), 77, np.uint8))
    assert len(ok) == 0


def test_size_change_between_calls(orb, oracle):
    g = orb.ORBextractor(1000, 1.2, 8, 20, 7)
    o = oracle.OracleExtractor(1000, 1.2, 8, 20, 7)
    for (w, h, s) in [(640, 480, 21), (1280, 720, 22), (752, 480, 23)]:
        img = synth.shapes_frame(w, h, s)
        gk, gd = g(img)
        ok, od = o(img)
        _assert_same_keypoints(gk, gd, ok, od)


def test_match_bf_config0(orb, oracle):
    a, b = synth.frame_pair(640, 480, 1)
    ex = oracle.OracleExtractor(1000, 1.2, 8, 20, 7)
    _, da = ex(a)
    _, db = ex(b)
    gi, gdist = orb.match_bf(da, db, 50, 0.6)
    oi, odist = oracle.match_bf(da, db, 50, 0.6)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gdist, odist)
    assert (gi >= 0).sum() > 50


def test_match_bf_ties_and_edges(orb, oracle):
    rng = np.random.default_rng(0)
    q = rng.integers(0, 256, (300, 32), dtype=np.uint8)
    t = np.concatenate([q[:100], q[:100], rng.integers(0, 256, (157, 32), dtype=np.uint8)])
    for th, ratio in [(50, 0.6), (256, 1.0), (0, 0.9)]:
        gi, gd = orb.match_bf(q, t, th, ratio)
        oi, od = oracle.match_bf(q, t, th, ratio)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
    gi, gd = orb.match_bf(q, t[:1], 256, 1.0)
    oi, od = oracle.match_bf(q, t[:1], 256, 1.0)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


@pytest.mark.parametrize("seed", [5, 6])
def test_stereo_frame_config1(orb, oracle, seed):
    L, R, disp = synth.stereo_pair(1241, 376, seed)
    bf, fx = 386.1, 718.856
    b = np.float32(bf) / np.float32(fx)
    gl = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    gr = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    kl, dl, kr, dr, uR, dep = orb.stereo_frame(gl, gr, L, R, bf, float(b))
    ol = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    orr = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    okl, odl, okr, odr, ouR, odep = oracle.stereo_frame(ol, orr, L, R, bf, float(b))
    _assert_same_keypoints(kl, dl, okl, odl)
    _assert_same_keypoints(kr, dr, okr, odr)
    np.testing.assert_array_equal(uR.view(np.uint32), ouR.view(np.uint32))
    np.testing.assert_array_equal(dep.view(np.uint32), odep.view(np.uint32))
    assert (uR >= 0).sum() > 300
    # The separate entry point on the same inputs gives the same answer.
    uR2, dep2 = orb.compute_stereo_matches(gl, gr, kl, dl, kr, dr, bf, float(b))
    np.testing.assert_array_equal(uR2.view(np.uint32), uR.view(np.uint32))
    np.testing.assert_array_equal(dep2.view(np.uint32), dep.view(np.uint32))


def test_stereo_no_matches(orb):
    L = synth.shapes_frame(1241, 376, 31)
    R = np.full_like(L, 128)
    gl = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    gr = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    kl, dl, kr, dr, uR, dep = orb.stereo_frame(gl, gr, L, R, 386.1, 386.1 / 718.856)
    assert len(kr) == 0 and len(kl) > 0
    assert np.all(uR == -1) and np.all(dep == -1)


# Parameter sets beyond the KITTI defaults.  Scale factors 1.5, 2.0 and 2.5
# drive the level kernel's three resize forms (narrow staged window, the
# global-tap form above ~1.9, the wide staged window), at sizes whose every
# level still has a FAST cell row (the reference divides by zero otherwise,
# ORBextractor.cc:1067-1071) and whose largest level quota fits the quadtree
# list (quota + 3 <= 1024).
PARAMS = [
    (640, 480, 1000, 1.5, 4, 20, 7), (1241, 376, 2000, 1.5, 4, 20, 7),
    (640, 480, 1000, 2.0, 3, 20, 7), (1241, 376, 1500, 2.0, 3, 20, 7),
    (640, 480, 1000, 2.5, 3, 20, 7), (1241, 376, 1200, 2.5, 2, 20, 7),
    (640, 480, 1000, 1.2, 1, 20, 7), (1241, 376, 1000, 1.2, 1, 20, 7),
    (640, 480, 1500, 1.2, 12, 20, 7), (1241, 376, 2000, 1.2, 10, 20, 7),
    (1241, 376, 2000, 1.2, 8, 30, 10), (640, 480, 1000, 1.2, 8, 12, 5),
]


@pytest.mark.parametrize("w,h,nf,sf,nl,ini,mn", PARAMS)
def test_extract_parameter_sets(orb, oracle, w, h, nf, sf, nl, ini, mn):
    img = synth.shapes_frame(w, h, 40 + nl)
    g = orb.ORBextractor(nf, sf, nl, ini, mn)
    o = oracle.OracleExtractor(nf, sf, nl, ini, mn)
    gk, gd = g(img)
    ok, od = o(img)
    for level in range(nl):
        np.testing.assert_array_equal(g.pyramid_level(level), o.level_image(level), err_msg=f"level {level}")
        np.testing.assert_array_equal(g.debug_image(level, 1), o.level_image(level, blurred=True),
                                      err_msg=f"blur {level}")
    _assert_same_keypoints(gk, gd, ok, od)
    assert len(gk) > 0


def test_extract_monocular_init_4000(orb, oracle):
    """Tracking's monocular initialiser runs ORBextractor(2*nFeatures, ...)
    (Tracking.cc:187-188): 4000 features on the KITTI yaml's 2000, a level-0
    quota of 869."""
    img = synth.shapes_frame(1241, 376, 17)
    g = orb.ORBextractor(4000, 1.2, 8, 20, 7)
    assert g.quotas()[0] == 869
    o = oracle.OracleExtractor(4000, 1.2, 8, 20, 7)
    gk, gd = g(img)
    ok, od = o(img)
    _assert_same_keypoints(gk, gd, ok, od)
    assert len(gk) > 3000


def test_nfeatures_bound(orb):
    """The quadtree list holds 1024 nodes per level: quota + 3 <= 1024, i.e.
    nfeatures up to 4703 at scale 1.2 over 8 levels; above, ORB_E_ARG."""
    assert orb.ORBextractor(4703, 1.2, 8, 20, 7).capacity(1241, 376) > 0
    with pytest.raises(orb.OrbError):
        orb.ORBextractor(4704, 1.2, 8, 20, 7).capacity(1241, 376)


@pytest.mark.parametrize("w,h,nf", [(640, 480, 1000), (1241, 376, 2000), (1280, 720, 2000)])
def test_generic_scalar_opencv_rounding(orb, oracle, w, h, nf):
    """opencv_sse2=False: resize and the Gaussian rounded as OpenCV's generic
    scalar code, against the oracle in the same mode; the default (SSE2)
    results differ from these."""
    img = synth.shapes_frame(w, h, 61)
    g = orb.ORBextractor(nf, 1.2, 8, 20, 7, opencv_sse2=False)
    gk, gd = g(img)
    with oracle.opencv_simd(0):      # the oracle blurs a level when asked for it: stay in the mode
        o = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)
        ok, od = o(img)
        for level in range(8):
            np.testing.assert_array_equal(g.pyramid_level(level), o.level_image(level), err_msg=f"level {level}")
            np.testing.assert_array_equal(g.debug_image(level, 1), o.level_image(level, blurred=True),
                                          err_msg=f"blur {level}")
    _assert_same_keypoints(gk, gd, ok, od)
    d = orb.ORBextractor(nf, 1.2, 8, 20, 7)
    d(img)
    assert not np.array_equal(d.pyramid_level(1), g.pyramid_level(1))


def test_pyramid_read_back_packed_and_strided(orb, oracle):
    """mvImagePyramid read-back (ORBextractor.h:162) into a packed buffer
    (rows packed on the device, one linear copy) and into rows wider than the
    level (one pitched copy): the same pixels, the oracle's, and nothing
    written past each row's width."""
    img = synth.shapes_frame(1241, 376, 5)
    g = orb.ORBextractor(2000, 1.2, 8, 20, 7)
    o = oracle.OracleExtractor(2000, 1.2, 8, 20, 7)
    g(img)
    o(img)
    for level in (0, 3, 7):
        for kind in (0, 1):
            packed = g.debug_image(level, kind)
            np.testing.assert_array_equal(packed, o.level_image(level, blurred=bool(kind)))
            h, w = packed.shape
            buf = np.full((h, w + 45), 7, np.uint8)
            assert orb.lib().orb_extractor_debug_image(g.h, level, kind, buf.ctypes.data, w + 45) == 0
            np.testing.assert_array_equal(buf[:, :w], packed)
            assert np.all(buf[:, w:] == 7)
    assert orb.lib().orb_extractor_debug_image(g.h, 0, 0, buf.ctypes.data, 10) != 0   # stride below the width
