"""CPU pins of the oracle's cvtColor to gray (8-bit, OpenCV's 14-bit fixed
point, Tracking.cc:250-276): OpenCV's well-known outputs for the primaries,
gray inputs unchanged, within one level of the float BT.601 weights, channel
order and alpha handling."""
import numpy as np
import pytest

oracle = pytest.importorskip("pyoracle")


def test_primaries_and_gray_levels():
    prim = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]], np.uint8)
    assert oracle.cvt_gray(prim, rgb=True).tolist() == [[76, 150, 29, 255, 0]]
    assert oracle.cvt_gray(prim, rgb=False).tolist() == [[29, 150, 76, 255, 0]]
    ramp = np.repeat(np.arange(256, dtype=np.uint8)[None, :, None], 3, axis=2)
    assert (oracle.cvt_gray(ramp)[0] == np.arange(256)).all()


def test_near_float_weights_and_order():
    rng = np.random.default_rng(1)
    im = rng.integers(0, 256, size=(37, 53, 3), dtype=np.uint8)
    g = oracle.cvt_gray(im, rgb=True).astype(np.float64)
    f = 0.299 * im[..., 0] + 0.587 * im[..., 1] + 0.114 * im[..., 2]
    assert np.abs(g - f).max() <= 1.0
    assert (oracle.cvt_gray(im[..., ::-1].copy(), rgb=False) == oracle.cvt_gray(im, rgb=True)).all()
    rgba = np.concatenate([im, rng.integers(0, 256, size=(37, 53, 1), dtype=np.uint8)], axis=2)
    assert (oracle.cvt_gray(rgba, rgb=True) == oracle.cvt_gray(im, rgb=True)).all()
