"""Serving loop: per-request results are identical alone and in the 16-stream
mix (and equal the oracle); admission by HBM budget; Poisson run completes."""
import numpy as np
import pytest

from orb_slam2_detailed_comments_amd import serving, synth

pytestmark = pytest.mark.gpu

STREAMS = serving.MIXED_16
NF = 3   # frames per stream


@pytest.fixture(scope="module")
def frames():
    return serving.make_stream_frames(STREAMS, NF, seed=200)


def _serve_all(orb, streams, frames, only=None):
    srv = orb.Server(streams, max_inflight=3)
    tickets = {}
    order = [(s, k) for k in range(NF) for s in range(len(streams)) if only is None or s in only]
    for s, k in order:
        tickets[srv.submit(s, frames[s][k])] = (s, k)
    srv.drain(120)
    out = {}
    recs = srv.poll()
    assert len(recs) == len(order)
    for r in recs:
        assert r["status"] == 0
        out[tickets[r["ticket"]]] = srv.fetch(r["ticket"])
    st = srv.stats()
    srv.close()
    return out, st


def test_alone_equals_mix_equals_oracle(orb, oracle, frames):
    mix, st = _serve_all(orb, STREAMS, frames)
    assert st["batches"] < st["requests"]          # requests were batched together
    for s in (0, 5, 9, 14):
        alone, _ = _serve_all(orb, STREAMS, frames, only={s})
        w, h, nf = STREAMS[s]
        ex = oracle.OracleExtractor(nf, 1.2, 8, 20, 7)
        prev = None
        for k in range(NF):
            ka, da, ia, xa = alone[(s, k)]
            km, dm, im, xm = mix[(s, k)]
            np.testing.assert_array_equal(ka.view(np.uint8), km.view(np.uint8))
            np.testing.assert_array_equal(da, dm)
            np.testing.assert_array_equal(ia, im)
            np.testing.assert_array_equal(xa, xm)
            ok, od = ex(frames[s][k])
            np.testing.assert_array_equal(ka.view(np.uint8), ok.view(np.uint8))
            np.testing.assert_array_equal(da, od)
            if prev is None:
                assert np.all(ia == -1)
            else:
                oi, odist = oracle.match_bf(od, prev, 50, 0.6)
                np.testing.assert_array_equal(ia, oi)
                np.testing.assert_array_equal(xa, odist)
            prev = od


def test_admission_by_hbm_budget(orb, frames):
    small = [(640, 480, 1000), (1280, 720, 2000)]
    with pytest.raises(orb.OrbError):
        orb.Server(small, hbm_budget=1 << 10)            # not even the resident descriptors fit
    srv = orb.Server(small, hbm_budget=12 << 20, max_inflight=2)
    t0 = srv.submit(0, frames[0][0])
    t1 = srv.submit(1, frames[12][0])
    srv.drain(60)
    recs = {r["ticket"]: r for r in srv.poll()}
    assert recs[t0]["status"] == 0 and recs[t0]["n_keypoints"] > 0
    assert recs[t1]["status"] == orb.ORB_E_ADMISSION
    assert srv.stats()["hbm_used"] <= 12 << 20
    srv.close()


def test_poisson_run_16_streams(orb, frames):
    srv = orb.Server(STREAMS, max_inflight=3)
    recs, _ = serving.run(srv, frames, fps=30.0, duration=1.0, seed=3)
    s = serving.summarize(recs)
    assert s["errors"] == 0 and s["requests"] == len(recs) > 300
    assert s["p99_ms"] < 100.0
    srv.close()


def test_native_replay_16_streams(orb, frames):
    """orb_server_replay: native submitters at 240 fps per stream for 0.5 s.
    Every scheduled frame is served without error, the count matches the
    schedule's, and the server's own request counter agrees."""
    srv = orb.Server(STREAMS, max_inflight=3)
    r = srv.replay(frames, fps=240.0, duration=0.5, seed=5)
    assert r["errors"] == 0 and r["refused"] == 0
    # 16 Poisson streams of mean 120 arrivals each: 1920 +- a few sqrt(1920)
    assert 1700 < r["requests"] < 2150, r
    assert 0 < r["p50_ms"] <= r["p99_ms"] <= r["max_ms"] < 1000.0
    assert r["frames_per_s"] > 0
    assert srv.stats()["requests"] == r["requests"]
    with pytest.raises(orb.OrbError, match="bad argument"):
        srv.replay(frames, fps=0.0, duration=0.5)
    srv.close()
