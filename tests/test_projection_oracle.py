"""CPU tier: pins for the oracle's SearchByProjection (both per-frame
overloads, ORBmatcher.cc:72-183 and :1710-1889).

The reference ships no tests for them.  The oracle is compared with a second
statement written independently (projection_scenes.py) on hand-built scenes
whose outcomes are also stated per point, and on dense random frames; every
planted mistake must change the outcome of at least one scene.
"""
import numpy as np
import pytest

import projection_scenes as ps


def _run_local(oracle, s, th):
    kp, desc, ur, occ, b, sc, pts, pd = s.arrays()
    got = oracle.search_by_projection(kp, desc, ur, occ, b, sc, pts, pd, th, 0.8)
    want = ps.search_local(kp, desc, ur, occ, b, sc, pts, pd, th, 0.8)
    return got, want


def _run_last(oracle, s, th, mono):
    kp, desc, ur, occ, b, sc, pts, pd = s.arrays()
    got = oracle.search_by_projection_last(kp, desc, ur, occ, b, sc, s.cam, mono, pts, pd, th, True)
    want = ps.search_last(kp, desc, ur, occ, b, sc, s.cam, mono, pts, pd, th, True)
    return got, want


def _same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def _all_scene_results(oracle):
    out = {}
    for name, build, th in ps.LOCAL_SCENES:
        out[name] = _run_local(oracle, build(), th)
    for name, build, th, mono in ps.LAST_SCENES:
        out[name] = _run_last(oracle, build(), th, mono)
    return out


def test_scenes_match_independent_statement(oracle):
    for name, (got, want) in _all_scene_results(oracle).items():
        assert _same(got, want), (name, got, want)


def test_scene_known_outcomes(oracle):
    r = {k: v[0] for k, v in _all_scene_results(oracle).items()}
    # radius: 0.999 and 0.998f -> 2.5 (the 3 px key is missed), 0.99 and 0.9 -> 4.0
    _, pm, _ = r["radius"]
    assert list(pm[0::2] >= 0) == [False, True, False, True]
    assert np.all(pm[1::2] >= 0)                               # 2.4 px keys always found
    # levels: [1, 2] window, ratio only within a level, TH_HIGH inclusive
    n, pm, _ = r["levels"]
    assert pm[0] == 1 and pm[1] == -1 and pm[2] == 6 and pm[3] == 8 and pm[4] == -1
    # window: the (3, 3) key is outside the circle; distance == r is outside
    _, pm, _ = r["window"]
    assert pm[0] == 0 and pm[1] == 3
    # claims
    n, pm, kpm = r["claims"]
    assert pm[0] == 1                                          # the occupied key is skipped
    assert pm[1] == 2 and pm[2] == 3                           # second point falls back to the 40-bit key
    assert pm[3] == 4 and pm[4] == 4 and kpm[4] == 4           # no-observation claim does not block
    assert pm[5] == -1 and pm[6] == -1                         # not in view / bad
    # stereo: er 4.5 > 4 rejected, 4.0 kept, mvuRight <= 0 unchecked
    _, pm, _ = r["stereo"]
    assert list(pm >= 0) == [False, True, True, True]
    _, pm, _ = r["th"]
    assert list(pm >= 0) == [True, False]
    # last frame: level windows by motion (groups at octave 0, 2, 4; key level = position in group)
    found = {name: r[name][1].reshape(3, 8) >= 0 for name in ("forward", "backward", "still", "mono")}
    lv = np.arange(8)
    for g, o in enumerate((0, 2, 4)):
        assert list(found["forward"][g]) == list(lv >= o)             # o = 0: no level check at all
        assert list(found["backward"][g]) == list(lv <= o)
        assert list(found["still"][g]) == list(abs(lv - o) <= 1)
        assert list(found["mono"][g]) == list(abs(lv - o) <= 1)       # bMono: neither direction
    n, pm, kpm = r["projection"]
    assert list(pm[:3]) == [-1, 0, -1]                         # behind, found, left of mnMinX
    assert list(pm[3:] >= 0) == [True, True, False]            # er 0, er == radius kept, above rejected
    n, pm, kpm = r["rotation"]
    # 26 claims; bins 30 deg (12 + the first claim of the shared key), 90 deg
    # (6) and 6 deg (3, bin 1) stay; 150 deg (2), 200 deg (1) and the shared
    # key's second claim (bin 0) are cut, and that key ends NULL
    assert n == 26 - 4
    assert np.sum(kpm == -2) == 4 and kpm[-1] == -2


PLANTED = {1: "viewCos compared in float", 2: "square window", 3: "ratio regardless of level",
           4: "claims never block", 5: "stereo er >= radius rejected", 6: "TH_HIGH exclusive",
           7: "backward window [o-1, o]", 8: "rotation bin rounded half to even",
           9: "claims without observations block"}


@pytest.mark.parametrize("variant", sorted(PLANTED))
def test_scenes_reject_planted_mistakes(oracle, variant):
    with oracle.planted(oracle.PLANT_PROJECTION, variant):
        res = _all_scene_results(oracle)
    assert any(not _same(got, want) for got, want in res.values()), PLANTED[variant]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_frames_local(oracle, seed):
    s = ps.random_frame(seed, last=False)
    (n, pm, kpm), want = _run_local(oracle, s, 1.0 if seed != 3 else 2.0)
    assert _same((n, pm, kpm), want)
    assert n > 200


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_random_frames_last(oracle, seed):
    s = ps.random_frame(seed, last=True)
    got, want = _run_last(oracle, s, 7.0 if seed == 4 else 15.0, seed == 6)
    assert _same(got, want)
    assert got[0] > 200
