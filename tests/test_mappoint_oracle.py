"""CPU pins of the oracle's MapPoint::ComputeDistinctiveDescriptors
(src/MapPoint.cc:365-450): hand-built distance patterns with known answers
for odd and even N (which fix the median index (size_t)(0.5*(N-1)) and the
first-minimum rule), a counting formulation of the order statistic on random
sets, and permutation equivariance."""
import numpy as np
import pytest

oracle = pytest.importorskip("pyoracle")


def _bits(idx):
    d = np.zeros(256, np.uint8)
    d[list(idx)] = 1
    return np.packbits(d, bitorder="little")


def test_known_answers():
    z = _bits([])
    a10 = _bits(range(10))
    b12 = _bits(list(range(9)) + [100, 101, 102])    # d(a10, b12) = 1 + 3 = 4
    # N=3: rows [0,10,12] [10,0,4] [12,4,0] -> medians 10, 4, 4 -> first of the 4s
    assert oracle.distinctive_descriptors([np.stack([z, a10, b12])]).tolist() == [1]
    # N=1, N=2: every median is the self-distance 0 -> row 0
    assert oracle.distinctive_descriptors([z[None], np.stack([a10, z])]).tolist() == [0, 0]
    # N=4: median index 1, i.e. each row's nearest other
    far = _bits(range(200, 256))                      # 56 bits away from z, 66 from a10
    got = oracle.distinctive_descriptors([np.stack([far, z, a10, b12])])
    # nearest others: far->z 56, z->a10 10, a10->b12 4, b12->a10 4 -> first 4 is row 2
    assert got.tolist() == [2]
    # empty point
    assert oracle.distinctive_descriptors([np.zeros((0, 32), np.uint8)]).tolist() == [-1]


def _counting(D):
    N = len(D)
    need = (N - 1) // 2 + 1
    med = [min(v for v in range(257) if (row <= v).sum() >= need) for row in D]
    return int(np.argmin(med))


@pytest.mark.parametrize("seed", range(6))
def test_order_statistic_by_counting(seed):
    rng = np.random.default_rng(seed)
    pts = []
    for n in (1, 2, 3, 4, 5, 8, 17, 40):
        base = rng.integers(0, 256, size=32, dtype=np.uint8)
        noise = rng.integers(0, 256, size=(n, 32), dtype=np.uint8) & rng.integers(0, 256, size=(n, 32),
                                                                                    dtype=np.uint8)
        pts.append(base ^ noise)
    got = oracle.distinctive_descriptors(pts)
    for p, g in zip(pts, got):
        D = np.unpackbits(p[:, None, :] ^ p[None, :, :], axis=-1).sum(-1)
        assert g == _counting(D)


def test_permutation_equivariance():
    rng = np.random.default_rng(9)
    for _ in range(20):
        p = rng.integers(0, 256, size=(11, 32), dtype=np.uint8)
        D = np.unpackbits(p[:, None, :] ^ p[None, :, :], axis=-1).sum(-1)
        med = np.sort(D, axis=1)[:, 5]
        if (med == med.min()).sum() > 1:
            continue                                   # only sets with a unique winner
        perm = rng.permutation(11)
        a = oracle.distinctive_descriptors([p])[0]
        b = oracle.distinctive_descriptors([p[perm]])[0]
        assert perm[b] == a
