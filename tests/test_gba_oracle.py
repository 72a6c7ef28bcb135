"""The global-BA oracle (oracle/gba_oracle.py) against what the mathematics fixes, on the CPU."""
import numpy as np
import pytest

from oracle import gba_oracle as G


def _brute_pairs(olm_sorted, nl):
    offs = [0]
    for l in range(nl):
        offs.append(offs[-1] + sum(1 for v in olm_sorted if v == l))
    pa, pb = [], []
    for l in range(nl):
        for a in range(offs[l], offs[l + 1]):
            for b in range(offs[l], offs[l + 1]):
                pa.append(a)
                pb.append(b)
    return np.array(offs), np.array(pa, dtype=np.int64), np.array(pb, dtype=np.int64)


@pytest.mark.parametrize("seed", range(6))
def test_build_pairs_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(0, 12))
    counts = rng.integers(0, 6, nl) if nl else np.zeros(0, int)
    olm = np.repeat(np.arange(nl), counts)
    offs, pa, pb = G.build_pairs(olm, nl)
    bo, ba, bb = _brute_pairs(list(olm), nl)
    assert offs.dtype == np.int32 and pa.dtype == np.int32 and pb.dtype == np.int32
    assert np.array_equal(offs, bo)
    assert np.array_equal(pa, ba) and np.array_equal(pb, bb)
    assert len(pa) == int(np.sum(counts.astype(np.int64) ** 2))


def test_build_pairs_trailing_and_leading_empty_landmarks():
    olm = np.array([2, 2, 3])
    offs, pa, pb = G.build_pairs(olm, 6)
    assert offs.tolist() == [0, 0, 0, 2, 3, 3, 3]
    assert list(zip(pa.tolist(), pb.tolist())) == [(0, 0), (0, 1), (1, 0), (1, 1), (2, 2)]
