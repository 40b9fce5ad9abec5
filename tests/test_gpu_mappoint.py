"""GPU parity of MapPoint::ComputeDistinctiveDescriptors (many points per
launch) against the oracle: the chosen row and descriptor, for point sizes
from 0 to 1024 observations, clustered (tie-prone) and random sets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _points(rng, sizes, ands):
    pts = []
    for n in sizes:
        base = rng.integers(0, 256, size=32, dtype=np.uint8)
        noise = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        for _ in range(ands - 1):
            noise &= rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        pts.append(base ^ noise)
    return pts


@pytest.mark.parametrize("ands", [1, 3, 6])
def test_distinctive_matches_oracle(orb, oracle, ands):
    rng = np.random.default_rng(ands)
    sizes = [0, 1, 2, 3, 4, 5, 7, 30, 63, 64, 65, 128, 200, 1024] + list(rng.integers(1, 60, size=300))
    pts = _points(rng, sizes, ands)
    best, out = orb.distinctive_descriptors(pts)
    want = oracle.distinctive_descriptors(pts)
    np.testing.assert_array_equal(best, want)
    for p, b, o in zip(pts, best, out):
        if b >= 0:
            np.testing.assert_array_equal(o, p[b])
        else:
            assert not o.any()


def test_distinctive_rejects_more_than_1024(orb):
    with pytest.raises(orb.OrbError):
        orb.distinctive_descriptors([np.zeros((1025, 32), np.uint8)])
