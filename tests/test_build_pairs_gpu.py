"""libxba xba_build_pairs (backend/ba.py:902-927 build_pairs) against the oracle, bit-exact."""
import numpy as np
import pytest

from oracle import gba_oracle as G

pytestmark = pytest.mark.gpu


def _sorted_olm(rng, nl, cmax):
    counts = rng.integers(0, cmax + 1, nl)
    return np.repeat(np.arange(nl, dtype=np.int32), counts)


@pytest.mark.parametrize("nl,cmax,seed", [(1, 1, 0), (7, 4, 1), (1023, 6, 2), (1024, 3, 3), (1025, 9, 4),
                                          (20000, 12, 5), (300000, 8, 6)])
def test_build_pairs_matches_oracle(nl, cmax, seed):
    from xclean_amd.ba import build_pairs
    rng = np.random.default_rng(seed)
    olm = _sorted_olm(rng, nl, cmax)
    print("nl", nl, "n_obs", len(olm), flush=True)
    offs, pa, pb = build_pairs(olm, nl)
    ro, ra, rb = G.build_pairs(olm, nl)
    assert offs.dtype == np.int32 and pa.dtype == np.int32
    assert np.array_equal(offs, ro)
    assert np.array_equal(pa, ra) and np.array_equal(pb, rb)


def test_build_pairs_edge_cases():
    from xclean_amd.ba import build_pairs
    # no observations, no landmarks
    offs, pa, pb = build_pairs(np.zeros(0, np.int32), 0)
    assert offs.tolist() == [0] and len(pa) == 0 and len(pb) == 0
    # landmarks without observations around one landmark with the kernel's largest segment (64 rows)
    olm = np.full(64, 3, np.int32)
    offs, pa, pb = build_pairs(olm, 5)
    ro, ra, rb = G.build_pairs(olm, 5)
    assert np.array_equal(offs, ro) and np.array_equal(pa, ra) and np.array_equal(pb, rb)
    assert len(pa) == 64 * 64


@pytest.mark.parametrize("olm,nl", [([0, 2, 1], 3), ([0, 1, 3], 3), ([-1, 0], 2)])
def test_build_pairs_rejects_unsorted_or_out_of_range(olm, nl):
    from xclean_amd.ba import build_pairs
    with pytest.raises(ValueError):
        build_pairs(np.array(olm, np.int32), nl)


def test_build_pairs_size_query_and_short_capacity():
    import ctypes as C
    import torch
    from xclean_amd import _lib
    L = _lib.load()
    olm = torch.tensor([0, 0, 1, 1, 1], dtype=torch.int32, device="cuda")
    offs = torch.empty(3, dtype=torch.int32, device="cuda")
    pa = torch.empty(12, dtype=torch.int32, device="cuda")
    pb = torch.empty(12, dtype=torch.int32, device="cuda")
    n = C.c_int64(-1)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.xba_build_pairs(C.c_void_p(olm.data_ptr()), 5, 2, C.c_void_p(offs.data_ptr()), None, None, 0, C.byref(n), s)
    assert rc == 0 and n.value == 13 and offs.tolist() == [0, 2, 5]
    rc = L.xba_build_pairs(C.c_void_p(olm.data_ptr()), 5, 2, C.c_void_p(offs.data_ptr()), C.c_void_p(pa.data_ptr()),
                           C.c_void_p(pb.data_ptr()), 12, C.byref(n), s)
    assert rc == _lib.XBA_E_SIZE and n.value == 13
