"""CPU tier: host logic of the serving driver (frame order and arrivals)."""
import numpy as np

from orb_slam2_detailed_comments_amd import serving


def test_sequence_index_pairs_neighbours():
    for n in (2, 3, 4, 8):
        idx = [serving.sequence_index(k, n) for k in range(5 * n)]
        assert idx[:n] == list(range(n))                 # forwards first
        assert all(abs(a - b) == 1 for a, b in zip(idx, idx[1:]))
        assert set(idx) == set(range(n))
    assert [serving.sequence_index(k, 1) for k in range(4)] == [0, 0, 0, 0]


def test_poisson_schedule_rate_and_order():
    ev = serving.poisson_schedule(16, 30.0, 10.0, seed=3)
    t = np.array([e[0] for e in ev])
    assert np.all(np.diff(t) >= 0) and t.max() < 10.0
    assert abs(len(ev) / (16 * 10.0) - 30.0) < 3.0          # ~30 arrivals/s per stream
    per = {}
    for _, s, k in ev:                                      # per-stream indices count up from 0
        assert k == per.get(s, 0)
        per[s] = k + 1
