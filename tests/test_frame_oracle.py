"""Oracle pins for the Frame keypoint ops: Frame::UndistortKeyPoints
(src/Frame.cc:724-776, cv::undistortPoints with R = I, P = K) and
Frame::ComputeStereoFromRGBD (src/Frame.cc:1085-1123).

OpenCV is absent, so the undistortion is pinned by the mathematics it solves:
points pushed through an independent float64 statement of the forward
Brown-Conrady model come back where they started, to within what five
fixed-point iterations reach; a radial-only point on the x axis follows the
scalar iteration x <- x_d / (1 + k1 x^2 + k2 x^4 + k3 x^6); k1 = 0 is a copy.
A vectorised float64 statement of cvUndistortPoints' operation order (the
reciprocal 1/fx multiplied, not a division) is bit-equal."""
import numpy as np
import pytest

TUM1_K = (517.306408, 516.469215, 318.643040, 255.313989)
TUM1_D = (0.262383, -0.953104, -0.005358, 0.002628, 1.163314)
EUROC_K = (458.654, 457.296, 367.215, 248.375)
EUROC_D = (-0.28340811, 0.07395907, 0.00019359, 1.76187114e-05)


def _distort(und_uv, K, D):
    """Forward model in float64: undistorted pixels -> distorted pixels."""
    fx, fy, cx, cy = K
    k1, k2, p1, p2 = D[:4]
    k3 = D[4] if len(D) > 4 else 0.0
    x = (und_uv[:, 0] - cx) / fx
    y = (und_uv[:, 1] - cy) / fy
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 * r2 + k3 * r2 * r2 * r2
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return np.stack([fx * xd + cx, fy * yd + cy], 1)


def _statement(xy, K, D):
    """cvUndistortPoints' order of operations, vectorised in float64."""
    K = np.float32(K).astype(np.float64)
    k = np.zeros(5)
    k[:len(D)] = np.float32(D).astype(np.float64)
    fx, fy, cx, cy = K
    ifx, ify = 1.0 / fx, 1.0 / fy
    x = (xy[:, 0].astype(np.float64) - cx) * ifx
    y = (xy[:, 1].astype(np.float64) - cy) * ify
    x0, y0 = x.copy(), y.copy()
    for _ in range(5):
        r2 = x * x + y * y
        icdist = 1.0 / (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2)
        dx = 2 * k[2] * x * y + k[3] * (r2 + 2 * x * x)
        dy = k[2] * (r2 + 2 * y * y) + 2 * k[3] * x * y
        x = (x0 - dx) * icdist
        y = (y0 - dy) * icdist
    return np.stack([(fx * x + cx).astype(np.float32), (fy * y + cy).astype(np.float32)], 1)


def _grid(w, h, step):
    u, v = np.meshgrid(np.arange(0, w, step, dtype=np.float64), np.arange(0, h, step, dtype=np.float64))
    return np.stack([u.ravel(), v.ravel()], 1)


@pytest.mark.parametrize("K,D,w,h,tol", [(EUROC_K, EUROC_D, 752, 480, 0.005), (TUM1_K, TUM1_D, 640, 480, 0.005)])
def test_undistort_inverts_the_distortion(oracle, K, D, w, h, tol):
    und = _grid(w, h, 8.0)
    dist = _distort(und, K, D).astype(np.float32)
    # where the image lies (distorted points inside the frame) and not in the
    # outer corners, where five iterations have not converged
    r = np.hypot((und[:, 0] - K[2]) / K[0], (und[:, 1] - K[3]) / K[1])
    inside = (dist[:, 0] >= 0) & (dist[:, 0] < w) & (dist[:, 1] >= 0) & (dist[:, 1] < h) & (r < 0.55)
    back = oracle.undistort_points(dist[inside], K, D)
    err = np.abs(back.astype(np.float64) - und[inside])
    assert inside.sum() > 2000
    assert err.max() < tol, err.max()
    # the model moves these points by pixels, so a dropped term cannot pass
    assert np.abs(dist[inside] - und[inside]).max() > 5.0


def test_undistort_radial_axis_known_answer(oracle):
    K = (500.0, 500.0, 320.0, 240.0)
    D = (-0.2, 0.05, 0.0, 0.0, 0.01)
    ud = np.array([[320.0 + 500.0 * xd, 240.0] for xd in (0.05, 0.2, 0.4, -0.3)], np.float32)
    got = oracle.undistort_points(ud, K, D)
    for (u, _), (gu, gv) in zip(ud, got):
        xd = (np.float64(u) - 320.0) * (1.0 / 500.0)
        x = xd
        for _ in range(5):
            r2 = x * x
            x = xd / (1 + ((0.01 * r2 + 0.05) * r2 - 0.2) * r2)
        assert gu == np.float32(500.0 * x + 320.0)
        assert gv == np.float32(240.0)


@pytest.mark.parametrize("K,D", [(EUROC_K, EUROC_D), (TUM1_K, TUM1_D), (TUM1_K, TUM1_D[:4])])
def test_undistort_matches_the_statement(oracle, K, D):
    rng = np.random.default_rng(3)
    xy = (rng.random((3000, 2)) * [752, 480]).astype(np.float32)
    np.testing.assert_array_equal(oracle.undistort_points(xy, K, D), _statement(xy, K, D))


def test_undistort_k1_zero_is_a_copy(oracle):
    xy = np.array([[10.25, 20.5], [700.0, 1.0]], np.float32)
    np.testing.assert_array_equal(oracle.undistort_points(xy, TUM1_K, (0.0, 0.3, 0.01, 0.01, 0.2)), xy)
    with pytest.raises(ValueError):
        oracle.undistort_points(xy, TUM1_K, (0.1, 0.2, 0.3))


def test_rgbd_depth_known_answers(oracle):
    h, w = 30, 40
    depth = (np.arange(h)[:, None] * 100 + np.arange(w)[None, :] + 0.5).astype(np.float32)
    depth[5, 7] = 0.0
    depth[6, 8] = -2.0
    xy = np.array([[10.7, 20.2], [0.0, 0.0], [39.99, 29.99], [7.5, 5.9], [8.0, 6.0], [40.0, 5.0], [3.0, 30.0]],
                  np.float32)
    xun = xy[:, 0] + np.float32(0.25)
    bf = np.float32(40.0)
    uR, dep = oracle.rgbd_depth(xy, xun, depth, bf)
    want_d = np.array([2010.5, 0.5, 2939.5, -1, -1, -1, -1], np.float32)
    np.testing.assert_array_equal(dep, want_d)
    for i in range(3):
        assert uR[i] == np.float32(xun[i] - np.float32(bf / want_d[i]))
    assert (uR[3:] == -1).all()
