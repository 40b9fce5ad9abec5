"""CPU pins of the oracle's initUndistortRectifyMap (CV_32FC1) and remap
(INTER_LINEAR, zero border) as Examples/Stereo/stereo_euroc.cc:124-188 calls
them: identity and half-pixel maps with known outputs, the distortion model
against an independent float64 statement of the radial-tangential model, and
remap within one level of float bilinear interpolation at the 1/32-quantised
source position."""
import numpy as np
import pytest

oracle = pytest.importorskip("pyoracle")

# EuRoC left camera (Examples/Stereo/EuRoC.yaml)
K_L = [458.654, 0.0, 367.215, 0.0, 457.296, 248.375, 0.0, 0.0, 1.0]
D_L = [-0.28340811, 0.07395907, 0.00019359, 1.76187114e-05, 0.0]
R_L = [0.999966347530033, -0.001422739138722922, 0.008079580483432283, 0.001365741834644127, 0.9999741760894847,
       0.007055629199258132, -0.008089410156878961, -0.007044357138835809, 0.9999424675829176]
P_L = [435.2046959714599, 0, 367.4517211914062, 0, 0, 435.2046959714599, 252.2008514404297, 0, 0, 0, 1, 0]


def test_identity_map_and_copy():
    K = [400.0, 0, 320.5, 0, 410.0, 240.25, 0, 0, 1]
    mx, my = oracle.init_rectify_map(K, None, None, K, 64, 48)
    jj, ii = np.meshgrid(np.arange(64), np.arange(48))
    # the ray round trip leaves double-rounding residue (e.g. -2.8e-14 at 0)
    assert np.abs(mx - jj).max() < 1e-5 and np.abs(my - ii).max() < 1e-5
    img = np.random.default_rng(0).integers(0, 256, size=(48, 64), dtype=np.uint8)
    assert (oracle.remap_linear(img, mx, my) == img).all()


def test_half_pixel_shift_and_border():
    img = np.random.default_rng(1).integers(0, 256, size=(10, 12), dtype=np.uint8)
    jj, ii = np.meshgrid(np.arange(12, dtype=np.float32), np.arange(10, dtype=np.float32))
    out = oracle.remap_linear(img, jj + 0.5, ii)
    a = img.astype(int)
    right = np.concatenate([a[:, 1:], np.zeros((10, 1), int)], axis=1)   # zero border past the last column
    assert (out == (a + right + 1) >> 1).all()
    assert (oracle.remap_linear(img, jj - 20, ii) == 0).all()            # wholly outside


def _brown(K, D, R, P, w, h):
    K = np.array(K, float).reshape(3, 3)
    P = np.array(P, float).reshape(3, -1)[:, :3]
    R = np.eye(3) if R is None else np.array(R, float).reshape(3, 3)
    k1, k2, p1, p2, k3 = (list(D) + [0] * 5)[:5]
    jj, ii = np.meshgrid(np.arange(w, dtype=float), np.arange(h, dtype=float))
    ray = np.linalg.solve(P @ R, np.stack([jj.ravel(), ii.ravel(), np.ones(w * h)]))
    x, y = ray[0] / ray[2], ray[1] / ray[2]
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return (K[0, 0] * xd + K[0, 2]).reshape(h, w), (K[1, 1] * yd + K[1, 2]).reshape(h, w)


def test_distortion_model():
    mx, my = oracle.init_rectify_map(K_L, D_L, R_L, P_L, 752, 480)
    bx, by = _brown(K_L, D_L, R_L, P_L, 752, 480)
    assert np.abs(mx - bx).max() < 2e-3 and np.abs(my - by).max() < 2e-3
    # and it really distorts: far from the identity near the corners
    assert abs(mx[0, 0] - 0) > 5


def test_remap_close_to_float_bilinear():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, size=(40, 50), dtype=np.uint8)
    mx = rng.uniform(0, 48.9, size=(30, 35)).astype(np.float32)
    my = rng.uniform(0, 38.9, size=(30, 35)).astype(np.float32)
    out = oracle.remap_linear(img, mx, my).astype(float)
    X = np.rint(mx.astype(np.float64) * 32) / 32
    Y = np.rint(my.astype(np.float64) * 32) / 32
    x0, y0 = np.floor(X).astype(int), np.floor(Y).astype(int)
    fx, fy = X - x0, Y - y0
    a = img.astype(float)
    ref = (a[y0, x0] * (1 - fx) * (1 - fy) + a[y0, x0 + 1] * fx * (1 - fy) + a[y0 + 1, x0] * (1 - fx) * fy +
           a[y0 + 1, x0 + 1] * fx * fy)
    assert np.abs(out - ref).max() <= 0.5 + 1e-9
