"""The allocation-free Euclidean local BA (libxba xba_np_solve_noalloc) against the allocating entry point
xba_np_solve it mirrors, through the C ABI: on two calls in a row, the second with a smaller problem (the static
workspace keeps the first call's contents), the poses, points, statistics and per-iteration updates agree within
1e-10 relative (the same kernels on the same landmark grouping; the Schur complement's FP64 atomics vary the last
bits from run to run, for either entry point), with a prior
and rotation factors, under both LM policies; sizes beyond the workspace and out-of-range indices are refused with
the documented codes; and repeated calls take no device memory."""
import ctypes

import numpy as np
import pytest

from xclean_amd import _lib, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


class _Rig:
    def __init__(self, r):
        self.fx_rect, self.cx, self.cy = r.fx_rect, r.cx, r.cy


def _case(nk, nl, seed, obs_per=4, extras=False):
    """make_problem(), and with ``extras`` an inverse-range prior and gyro rotation factors shaped like those of the
    reference-written golden case (tests/golden/make_golden.py)."""
    p = synth.make_problem(nk=nk, nl=nl, obs_per=obs_per, seed=seed)
    prior = rotf = None
    if extras:
        rng = np.random.default_rng(seed)
        n = len(p["pts0"])
        c = p["pts_true"] + rng.normal(0, 0.3, (n, 3))
        rho0 = 1.0 / np.linalg.norm(p["pts_true"] - c, axis=1) * (1 + 0.05 * rng.normal(size=n))
        prior = (c, rho0, 2.0, (rng.uniform(size=n) > 0.2).astype(float))
        Rcw = [np.linalg.inv(T)[:3, :3] for T in p["poses_true"]]
        rotf = [(i, i + 1, Rcw[i] @ Rcw[i + 1].T @ synth.so3_exp(rng.normal(0, 0.002, 3)), 0.01)
                for i in range(0, nk - 1, 3)]
    return p, prior, rotf


@pytest.mark.parametrize("policy", [0, 1])
def test_noalloc_matches_the_allocating_entry_point_on_two_calls(policy):
    from xclean_amd.ba import _solve_np
    for (nk, nl, seed, extras) in [(24, 3000, 3, True), (15, 600, 5, False)]:
        p, prior, rotf = _case(nk, nl, seed, extras=extras)
        print("noalloc case nk=%d nl=%d N=%d policy=%d" % (nk, nl, len(p["okf"]), policy), flush=True)
        args = (p["poses0"], p["pts0"], p["okf"], p["olm"], p["ouv"], _Rig(p["rig"]), 0, 5)
        kw = dict(prior=prior, rotf=rotf, history=True, _policy=policy)
        ref = _solve_np(*args, **kw)
        got = _solve_np(*args, noalloc=True, **kw)
        assert _rel(got[0], ref[0]) <= TOL
        assert _rel(got[1], ref[1]) <= TOL
        for k in ref[2]:
            assert abs(got[2][k] - ref[2][k]) <= TOL * abs(ref[2][k]), k
        assert _rel(got[3], ref[3]) <= TOL
        # the reference's own invariant (tests/test_ba_solver.py:26-33) holds for the no-allocation path too
        e0 = np.sqrt(np.mean(np.sum((p["poses0"][:, :3, 3] - p["poses_true"][:, :3, 3]) ** 2, 1)))
        e1 = np.sqrt(np.mean(np.sum((got[0][:, :3, 3] - p["poses_true"][:, :3, 3]) ** 2, 1)))
        assert e1 < 0.6 * e0


def test_noalloc_refuses_what_it_cannot_hold_and_bad_indices():
    from xclean_amd.ba import _solve_np
    p, _, _ = _case(8, 300, 7)
    args = [p["poses0"], p["pts0"], p["okf"], p["olm"], p["ouv"], _Rig(p["rig"]), 0, 2]
    with pytest.raises(_lib.XbaError, match="noalloc"):
        big = synth.make_problem(nk=40, nl=200, obs_per=3, seed=1)       # nk beyond XBA_NOALLOC_MAX_NK
        _solve_np(big["poses0"], big["pts0"], big["okf"], big["olm"], big["ouv"], _Rig(big["rig"]), 0, 1,
                  noalloc=True)
    bad = np.array(p["olm"]).copy()
    bad[5] = len(p["pts0"]) + 3
    with pytest.raises(_lib.XbaError, match="noalloc"):
        _solve_np(*args[:3], bad, *args[4:], noalloc=True)
    # after a refused call the entry point still solves correctly
    ref = _solve_np(*args)
    got = _solve_np(*args, noalloc=True)
    assert _rel(got[0], ref[0]) <= TOL and _rel(got[1], ref[1]) <= TOL


def test_noalloc_does_not_allocate_device_memory():
    """The caching allocator's and the device's used memory do not grow across no-allocation calls."""
    import torch

    from xclean_amd.ba import _solve_np
    p, _, _ = _case(12, 1000, 9)
    args = (p["poses0"], p["pts0"], p["okf"], p["olm"], p["ouv"], _Rig(p["rig"]), 0, 3)
    _solve_np(*args, noalloc=True)
    torch.cuda.synchronize()
    free0, _total = torch.cuda.mem_get_info()
    for _ in range(3):
        _solve_np(*args, noalloc=True)
    torch.cuda.synchronize()
    free1, _total = torch.cuda.mem_get_info()
    # the inputs' torch tensors come from torch's cache; the library itself takes nothing from the device
    assert free1 >= free0 - (1 << 20)
    L = _lib.load()
    assert L.xba_np_solve_noalloc is not None and ctypes.cast(L.xba_np_solve_noalloc, ctypes.c_void_p).value
