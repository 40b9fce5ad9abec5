"""GPU SearchByProjection, both per-frame overloads (ORBmatcher.cc:72-183 and
:1710-1889), against the oracle: point matches, the keypoint -> point map
and nmatches must be identical, on the hand-built scenes of
projection_scenes.py (one frame per call, and all of them in one call) and
on dense random frames batched many per call."""
import numpy as np
import pytest

import projection_scenes as ps

pytestmark = pytest.mark.gpu


def _frame_tuple(s):
    kp, desc, ur, occ, b, sc, pts, pd = s.arrays()
    return (kp, desc, ur, occ, b, pts, pd, s.cam), sc


def _oracle_local(oracle, f, sc, th):
    kp, desc, ur, occ, b, pts, pd, _ = f
    return oracle.search_by_projection(kp, desc, ur, occ, b, sc, pts, pd, th, 0.8)


def _oracle_last(oracle, f, sc, th, mono):
    kp, desc, ur, occ, b, pts, pd, cam = f
    return oracle.search_by_projection_last(kp, desc, ur, occ, b, sc, cam, mono, pts, pd, th, True)


def _assert_same(got, want, tag):
    assert got[0] == want[0], (tag, got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1], err_msg=f"{tag} point_match")
    np.testing.assert_array_equal(got[2], want[2], err_msg=f"{tag} key_point")


def test_local_scenes(orb, oracle):
    for name, build, th in ps.LOCAL_SCENES:
        f, sc = _frame_tuple(build())
        got = orb.search_by_projection([f], sc, th, 0.8)[0]
        _assert_same(got, _oracle_local(oracle, f, sc, th), name)


def test_last_scenes(orb, oracle):
    for name, build, th, mono in ps.LAST_SCENES:
        f, sc = _frame_tuple(build())
        got = orb.search_by_projection_last([f], sc, th, mono, True)[0]
        _assert_same(got, _oracle_last(oracle, f, sc, th, mono), name)


def test_scenes_in_one_call(orb, oracle):
    """Frames of one call are independent: every scene at once (th = 1, so
    the th scene runs unscaled here)."""
    fs = [_frame_tuple(b())[0] for _, b, _ in ps.LOCAL_SCENES]
    got = orb.search_by_projection(fs, ps.SCALE, 1.0, 0.8)
    for (name, _, _), f, g in zip(ps.LOCAL_SCENES, fs, got):
        _assert_same(g, _oracle_local(oracle, f, ps.SCALE, 1.0), name)


@pytest.mark.parametrize("th", [1.0, 3.0])
def test_local_random_frames_batched(orb, oracle, th):
    scenes = [ps.random_frame(100 + i, n_kp=2000, n_pts=1800) for i in range(8)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got = orb.search_by_projection(fs, ps.SCALE, th, 0.8)
    for i, (f, g) in enumerate(zip(fs, got)):
        want = _oracle_local(oracle, f, ps.SCALE, th)
        _assert_same(g, want, f"frame {i}")
        assert want[0] > 300


@pytest.mark.parametrize("th,mono", [(15.0, False), (7.0, True)])
def test_last_random_frames_batched(orb, oracle, th, mono):
    scenes = [ps.random_frame(200 + i, n_kp=2000, n_pts=1800, last=True) for i in range(8)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got = orb.search_by_projection_last(fs, ps.SCALE, th, mono, True)
    for i, (f, g) in enumerate(zip(fs, got)):
        want = _oracle_last(oracle, f, ps.SCALE, th, mono)
        _assert_same(g, want, f"frame {i}")
        assert want[0] > 300


def test_contested_keypoints(orb, oracle):
    """Many points over few keypoints: long chains of claims that the rounds
    must resolve exactly as the in-order pass does."""
    s = ps.random_frame(300, n_kp=150, n_pts=2500)
    f, sc = _frame_tuple(s)
    got = orb.search_by_projection([f], sc, 3.0, 0.8)[0]
    _assert_same(got, _oracle_local(oracle, f, sc, 3.0), "contested local")
    s = ps.random_frame(301, n_kp=150, n_pts=2500, last=True)
    f, sc = _frame_tuple(s)
    got = orb.search_by_projection_last([f], sc, 15.0, False, True)[0]
    _assert_same(got, _oracle_last(oracle, f, sc, 15.0, False), "contested last")


def test_empty_inputs(orb):
    f, sc = _frame_tuple(ps.Scene(9))
    assert orb.search_by_projection([f], sc, 1.0, 0.8)[0][0] == 0
    s = ps.scene_radius()
    s.pts, s.pdesc = [], []
    f, sc = _frame_tuple(s)
    n, pm, kpm = orb.search_by_projection([f], sc, 1.0, 0.8)[0]
    assert n == 0 and len(pm) == 0 and np.all(kpm == -1)


def test_last_rejects_angles_outside_0_360(orb):
    """The rotation check bins angle differences mod 360: an angle outside
    [0, 360] would index past the histogram (the reference asserts)."""
    s = ps.scene_last_rotation()
    s.pts[0] = s.pts[0][:3] + (400.0,) + s.pts[0][4:]
    f, sc = _frame_tuple(s)
    with pytest.raises(orb.OrbError):
        orb.search_by_projection_last([f], sc, 15.0, False, True)
    assert orb.search_by_projection_last([f], sc, 15.0, False, False)[0][0] > 0   # no check: accepted


@pytest.mark.parametrize("limit", [0, 1, 3])
def test_in_order_pass_and_round_limits(orb, oracle, limit):
    """With the round limit at 0 the in-order pass runs alone; with a few
    rounds, frames that have not settled fall back to it.  Both equal the
    oracle, for both overloads."""
    ctx = orb.Context(0)
    ctx.set_round_limit(limit)
    scenes = [ps.random_frame(400 + i, n_kp=1500, n_pts=1500) for i in range(3)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got = orb.search_by_projection(fs, ps.SCALE, 3.0, 0.8, context=ctx)
    for i, (f, g) in enumerate(zip(fs, got)):
        _assert_same(g, _oracle_local(oracle, f, ps.SCALE, 3.0), f"local {i}")
    assert ctx.last_rounds()[0] == limit
    if limit == 0:
        assert ctx.last_rounds()[1]
    scenes = [ps.random_frame(500 + i, n_kp=1500, n_pts=1500, last=True) for i in range(3)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got = orb.search_by_projection_last(fs, ps.SCALE, 15.0, False, True, context=ctx)
    for i, (f, g) in enumerate(zip(fs, got)):
        _assert_same(g, _oracle_last(oracle, f, ps.SCALE, 15.0, False), f"last {i}")


def test_timed_forms_return_the_calls_results(orb, oracle):
    scenes = [ps.random_frame(600 + i, n_kp=1200, n_pts=1000) for i in range(2)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got, ms = orb.search_by_projection(fs, ps.SCALE, 3.0, 0.8, timing_steps=3)
    assert ms > 0
    for f, g in zip(fs, got):
        _assert_same(g, _oracle_local(oracle, f, ps.SCALE, 3.0), "timed local")
    scenes = [ps.random_frame(700 + i, n_kp=1200, n_pts=1000, last=True) for i in range(2)]
    fs = [_frame_tuple(s)[0] for s in scenes]
    got, ms = orb.search_by_projection_last(fs, ps.SCALE, 15.0, False, True, timing_steps=3)
    assert ms > 0
    for f, g in zip(fs, got):
        _assert_same(g, _oracle_last(oracle, f, ps.SCALE, 15.0, False), "timed last")
