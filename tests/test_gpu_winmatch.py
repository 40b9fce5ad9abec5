"""GPU parity of SearchForInitialization over the keypoint grid: several
frame pairs per call (extracted keypoints, different motions and sizes),
windows and orientation check on/off, the hand-built steal and tie scenes,
and degenerate pairs -- matches, counts and vbPrevMatched equal the oracle's."""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import synth

pytestmark = pytest.mark.gpu


def _frames(oracle, w, h, seed, angle, shift, nf=2000):
    a, b = synth.frame_pair(w, h, seed, angle_deg=angle, shift=shift)
    k1, d1 = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)(a)
    k2, d2 = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)(b)
    return k1, d1, k2, d2, (0.0, float(w), 0.0, float(h)), np.stack([k1["x"], k1["y"]], 1)


@pytest.mark.parametrize("window,check", [(100, True), (50, False), (200, True)])
def test_batch_matches_oracle(orb, oracle, window, check):
    pairs = [_frames(oracle, 640, 480, 1, 3.0, (6.0, 4.0)), _frames(oracle, 752, 480, 2, 0.5, (2.0, 1.0), 1000),
             _frames(oracle, 1241, 376, 3, 8.0, (15.0, -5.0)), _frames(oracle, 640, 480, 4, 0.0, (0.0, 0.0))]
    got = orb.search_for_initialization(pairs, window, 0.9, check)
    for p, (n, m, pv) in zip(pairs, got):
        on, om, opv = oracle.search_for_initialization(*p, window, 0.9, check)
        assert n == on
        np.testing.assert_array_equal(m, om)
        np.testing.assert_array_equal(pv.view(np.uint32), opv.view(np.uint32))


def test_duplicates_and_ties(orb, oracle):
    """Duplicated F2 descriptors and keypoints make ties in distance: the first
    candidate in grid order must win, as in the reference."""
    k1, d1, k2, d2, b, prev = _frames(oracle, 640, 480, 5, 1.0, (3.0, 2.0))
    rng = np.random.default_rng(0)
    dup = rng.choice(len(k2), size=len(k2) // 3, replace=False)
    k2b = np.concatenate([k2, k2[dup]])
    k2b["x"][len(k2):] += rng.integers(-3, 4, size=len(dup))
    d2b = np.concatenate([d2, d2[dup]])
    p = (k1, d1, k2b, d2b, b, prev)
    (n, m, pv), = orb.search_for_initialization([p])
    on, om, opv = oracle.search_for_initialization(*p)
    assert n == on
    np.testing.assert_array_equal(m, om)


def test_degenerate_pairs(orb, oracle):
    k1, d1, k2, d2, b, prev = _frames(oracle, 640, 480, 6, 2.0, (4.0, 4.0))
    empty_k = np.zeros(0, orb.KEYPOINT_DTYPE)
    empty_d = np.zeros((0, 32), np.uint8)
    high = k1.copy()
    high["octave"] = 1
    pairs = [(k1, d1, empty_k, empty_d, b, prev), (high, d1, k2, d2, b, prev), (k1, d1, k2, d2, b, prev + 1000)]
    got = orb.search_for_initialization(pairs)
    for p, (n, m, pv) in zip(pairs, got):
        on, om, opv = oracle.search_for_initialization(*p)
        assert n == on == 0
        np.testing.assert_array_equal(m, om)
        np.testing.assert_array_equal(pv, opv)


def _check_pairs(orb, oracle, pairs, ctx, window=100, check=True):
    got = orb.search_for_initialization(pairs, window, 0.9, check, context=ctx)
    for i, (p, (n, m, pv)) in enumerate(zip(pairs, got)):
        on, om, opv = oracle.search_for_initialization(*p, window, 0.9, check)
        assert n == on, (i, n, on)
        np.testing.assert_array_equal(m, om, err_msg=f"pair {i}")
        np.testing.assert_array_equal(pv.view(np.uint32), opv.view(np.uint32), err_msg=f"pair {i}")


def test_rounds_and_in_order_pass_agree(orb, oracle):
    """The parallel rounds settle on the in-order pass's result; with the
    round limit at 0 the in-order pass runs alone, and with 2 the pairs that
    have not settled fall back to it.  All three equal the oracle."""
    ctx = orb.Context(0)
    pairs = [_frames(oracle, 640, 480, 11, 3.0, (6.0, 4.0)), _frames(oracle, 1241, 376, 12, 8.0, (15.0, -5.0)),
             _frames(oracle, 752, 480, 13, 1.0, (2.0, 2.0), 1000)]
    _check_pairs(orb, oracle, pairs, ctx)
    rounds, seq = ctx.last_rounds()
    assert not seq and 2 <= rounds <= 48
    for limit in (0, 2):
        ctx.set_round_limit(limit)
        _check_pairs(orb, oracle, pairs, ctx)
        assert ctx.last_rounds() == (limit, True) or (limit == 2 and not ctx.last_rounds()[1])
    ctx.set_round_limit(48)


def test_contested_targets(orb, oracle):
    """Many F1 keypoints over few F2 keypoints: one target is the best of many
    queries (more claims than its slots in the first rounds) and is stolen
    along long chains; the result is still the in-order pass's."""
    ctx = orb.Context(0)
    k1, d1, k2, d2, b, prev = _frames(oracle, 640, 480, 14, 2.0, (3.0, 3.0))
    rng = np.random.default_rng(3)
    keep = rng.choice(np.flatnonzero(k2["octave"] == 0), size=12, replace=False)
    k2s, d2s = k2[keep], d2[keep]
    reps = 60
    k1c = np.repeat(k2s, reps)
    k1c["x"] += rng.uniform(-20, 20, size=len(k1c)).astype(np.float32)
    k1c["y"] += rng.uniform(-20, 20, size=len(k1c)).astype(np.float32)
    d1c = np.repeat(d2s, reps, axis=0)
    flip = rng.integers(0, 256, size=(len(d1c), 32), dtype=np.uint8) & (rng.random((len(d1c), 32)) < 0.08).astype(np.uint8) * 0x11
    d1c = d1c ^ flip
    p = (k1c, d1c, k2s, d2s, b, np.stack([k1c["x"], k1c["y"]], 1))
    _check_pairs(orb, oracle, [p], ctx, window=100, check=True)
    _check_pairs(orb, oracle, [p], ctx, window=100, check=False)


def test_timed_form_returns_the_calls_results(orb, oracle):
    """The measurement aid repeats the device work with vbPrevMatched restored
    each time, so its outputs are exactly one call's."""
    pairs = [_frames(oracle, 640, 480, 15, 3.0, (6.0, 4.0)), _frames(oracle, 640, 480, 16, 1.0, (2.0, 3.0))]
    want = orb.search_for_initialization(pairs)
    got, ms = orb.search_for_initialization(pairs, timing_steps=3)
    assert ms > 0
    for (n, m, pv), (wn, wm, wpv) in zip(got, want):
        assert n == wn
        np.testing.assert_array_equal(m, wm)
        np.testing.assert_array_equal(pv.view(np.uint32), wpv.view(np.uint32))
