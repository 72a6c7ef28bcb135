"""Parity of the HIP window solve (xclean_amd.window -> libxba.so) with the CPU FP64 oracle.

Bars: the FP64 path matches the oracle's update dx within 1e-9 relative (the oracle restates vio/window.py's solve);
costs within 1e-9 relative; the mixed-precision path within 1e-6 relative of the FP64 oracle's dx.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _gpu_solve(pbs, iters, precision="fp64", **kw):
    from xclean_amd.window import BatchSolver
    bs = BatchSolver(pbs, precision=precision, **kw)
    st = bs.solve(iters=iters)
    return bs, st


def _case(name):
    if name == "vo_small":
        return synth.make_window(P=6, L=300, views=(2, 6), seed=[1, 1])[0]
    if name == "cfg0":
        return synth.make_config(0)[0]
    if name == "vi_small":
        return synth.make_window(P=6, L=400, views=(2, 6), seed=[1, 2], vi=True)[0]
    if name == "ur":
        return synth.make_window(P=7, L=300, views=(1, 7), seed=[1, 3], stereo="ur")[0]
    if name == "posefree":
        return synth.make_window(P=6, L=400, views=(1, 6), seed=[1, 4], posefree=True)[0]
    raise KeyError(name)


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("name", ["vo_small", "vi_small", "ur", "posefree"])
def test_linearize_matches_oracle(name, det):
    from xclean_amd.window import BatchSolver
    pb = _case(name)
    lam = 1e-4
    c, H6, b6, Hi, bl, Hlp, ok, chi = O.window_build(pb.copy(), lam)
    ref = pb.copy()
    H, b, crest = O.window_full_system(ref, H6, b6)
    bs = BatchSolver([pb], precision="fp64", use_graph=False, deterministic=det)
    bs.upload(lam)
    r = bs.linearize(lam)[0]
    n = H.shape[0]
    assert r["H"].shape == (n, n)
    print(name, "H rel", _rel(r["H"], H), "b rel", _rel(r["b"], b))
    assert _rel(r["H"], H) < 1e-12
    assert _rel(r["b"], b) < 1e-12
    cost_ref = c + crest
    if pb.pf is not None:
        pass  # the pose-free cost is part of solve()'s total, not build_system's
    else:
        assert abs(r["cost"] - cost_ref) <= 1e-12 * abs(cost_ref)
        np.testing.assert_array_equal(r["lm_ok"][:pb.L], ok)
        fin = np.isfinite(chi)
        np.testing.assert_array_equal(np.isfinite(r["ob_chi2"]), fin)
        np.testing.assert_allclose(r["ob_chi2"][fin], chi[fin], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(r["Hll_inv"][ok], Hi[ok], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("row", ["host_frame", "other_frame"])
def test_nonfinite_observation_matches_oracle(row, det):
    """A NaN pixel in a host-frame row only drops its landmark (its Hll fails the determinant guard,
    vio/window.py:421): the reduced system stays finite and equal to the oracle's.  In a row of another keyframe it
    reaches the pose blocks, and the reduced system is non-finite in both."""
    from xclean_amd.window import BatchSolver
    pb = _case("vo_small")
    st, host, pose = np.asarray(pb.ob_start), np.asarray(pb.lm_host), np.asarray(pb.ob_pose)
    want_host = row == "host_frame"
    k = [k for l in range(pb.L) for k in range(st[l], st[l + 1])
         if (pose[k] == host[l]) == want_host and st[l + 1] - st[l] > 3][3]
    l = int(np.searchsorted(st, k, side="right") - 1)
    z = np.asarray(pb.ob_z, np.float64).reshape(-1, 2).copy()
    z[k, 0] = np.nan
    pb.ob_z = z
    lam = 1e-4
    c, H6, b6, Hi, bl, Hlp, ok, chi = O.window_build(pb.copy(), lam)
    H, b, _ = O.window_full_system(pb.copy(), H6, b6)
    bs = BatchSolver([pb], precision="fp64", use_graph=False, deterministic=det)
    bs.upload(lam)
    r = bs.linearize(lam)[0]
    assert not ok[l]
    np.testing.assert_array_equal(r["lm_ok"][:pb.L], ok)
    assert np.isfinite(r["H"]).all() == np.isfinite(H).all() == want_host
    if want_host:
        assert _rel(r["H"], H) < 1e-12
        assert _rel(r["b"], b) < 1e-12


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("name", ["vo_small", "cfg0", "vi_small", "ur", "posefree"])
def test_fp64_first_update_matches_oracle(name, det):
    pb = _case(name)
    ref = pb.copy()
    st_ref, dx_ref, acc_ref = O.window_solve(ref, iters=1, history=True)
    bs, st = _gpu_solve([pb], 1, "fp64", use_graph=False, deterministic=det)
    dx = bs.last_dx()[0]
    e = _rel(dx, dx_ref[0])
    print(name, "dx rel err", e, "cost", st[0]["cost"], st_ref["cost"])
    assert e < 1e-9
    assert abs(st[0]["cost0"] - st_ref["cost0"]) <= 1e-9 * abs(st_ref["cost0"])
    assert abs(st[0]["cost"] - st_ref["cost"]) <= 1e-9 * abs(st_ref["cost"])
    assert st[0]["accepted"] == st_ref["accepted"]
    assert _rel(pb.pwb, ref.pwb) < 1e-9
    assert _rel(pb.lmp, ref.lmp) < 1e-9


@pytest.mark.parametrize("name", ["vo_small", "vi_small"])
def test_fp64_five_iterations_match_oracle(name):
    pb = _case(name)
    ref = pb.copy()
    st_ref, dx_ref, acc_ref = O.window_solve(ref, iters=5, history=True)
    bs, st = _gpu_solve([pb], 5, "fp64", use_graph=True)
    assert st[0]["accepted"] == st_ref["accepted"]
    assert abs(st[0]["lam"] - st_ref["lam"]) <= 1e-12 * st_ref["lam"]
    assert _rel(pb.Rwb, ref.Rwb) < 1e-9
    assert _rel(pb.pwb, ref.pwb) < 1e-9
    assert _rel(pb.lmp, ref.lmp) < 1e-9
    assert abs(st[0]["cost"] - st_ref["cost"]) <= 1e-9 * abs(st_ref["cost"])


def test_batch_of_uneven_windows_matches_oracle_per_window():
    # in FP64 n = 180 and n = 150 are factored in global memory (150: rows not a multiple of 4 entries), the others
    # in LDS
    pbs = [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[3, i])[0]
           for i, (P, L) in enumerate([(5, 50), (12, 700), (8, 300), (30, 900), (6, 1), (25, 800)])]
    refs = [pb.copy() for pb in pbs]
    outs = [O.window_solve(r, iters=3, history=True) for r in refs]
    bs, st = _gpu_solve(pbs, 3, "fp64")
    for pb, ref, (sr, dxr, ar), s in zip(pbs, refs, outs, st):
        assert s["accepted"] == sr["accepted"]
        assert _rel(pb.pwb, ref.pwb) < 1e-9
        assert _rel(pb.lmp, ref.lmp) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_deterministic_kernels_graph_replay_is_bitwise_equal_to_eager(precision):
    a = _case("vo_small")
    b = a.copy()
    _gpu_solve([a], 4, precision, use_graph=False, deterministic=True)
    _gpu_solve([b], 4, precision, use_graph=True, deterministic=True)
    np.testing.assert_array_equal(a.Rwb, b.Rwb)
    np.testing.assert_array_equal(a.lmp, b.lmp)


@pytest.mark.parametrize("det", [False, True])
def test_mixed_precision_update_error_vs_fp64_oracle(det):
    pb = _case("cfg0")
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=1, history=True)
    bs, st = _gpu_solve([pb], 1, "mixed", use_graph=False, deterministic=det)
    dx = bs.last_dx()[0]
    e = _rel(dx, dx_ref[0])
    print("mixed dx rel err vs FP64 oracle:", e, "refinement steps", st[0]["refine_steps"])
    assert e < 1e-6


def test_masked_invalid_and_empty_landmarks():
    pb = _case("vo_small")
    pb.lm_mask[::7] = False
    pb.ob_valid[::5] = False
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=2, history=True)
    bs, st = _gpu_solve([pb], 2, "fp64")
    assert _rel(bs.last_dx()[0], dx_ref[1]) < 1e-9
    assert _rel(pb.lmp, ref.lmp) < 1e-9


def test_fast_graph_replay_matches_eager_to_rounding():
    a = _case("vo_small")
    b = a.copy()
    _gpu_solve([a], 4, "fp64", use_graph=False)
    _gpu_solve([b], 4, "fp64", use_graph=True)
    assert _rel(a.lmp, b.lmp) < 1e-12
    assert _rel(a.pwb, b.pwb) < 1e-12


# ---- the tiled multi-workgroup Cholesky (64 x 64 tiles, MFMA trailing updates) of large reduced systems

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["vo_small", "cfg0", "vi_small", "ur", "posefree"])
def test_tiled_cholesky_forced_on_small_windows_matches_oracle(name, precision):
    """big_n = 1 sends every window through the tiled factorisation (one to two tiles per side here)."""
    pb = _case(name)
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=1, history=True)
    bs, st = _gpu_solve([pb], 1, precision, use_graph=False, big_n=1)
    e = _rel(bs.last_dx()[0], dx_ref[0])
    print(name, precision, "tiled dx rel err", e, flush=True)
    assert st[0]["chol_ok"]
    assert e < (1e-9 if precision == "fp64" else 1e-6)
    assert st[0]["accepted"] == st_ref["accepted"]


def _large_windows():
    # n = 300 (20 VI keyframes: 5 tiles of 64 per side) and n = 240 (40 VO keyframes: 4 tiles)
    return [synth.make_window(P=20, L=2000, views=(2, 10), seed=[5, 0], vi=True)[0],
            synth.make_window(P=40, L=3000, views=(2, 10), seed=[5, 1])[0]]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_cholesky_large_windows_match_oracle(precision):
    pbs = _large_windows()
    refs = [pb.copy() for pb in pbs]
    outs = [O.window_solve(r, iters=1, history=True) for r in refs]
    print("large windows", [(pb.P, pb.n_state(), pb.M) for pb in pbs], precision, flush=True)
    bs, st = _gpu_solve(pbs, 1, precision, use_graph=False)
    assert bs.plan_info()["max_n"] == 300
    for dx, (sr, dxr, _a), s in zip(bs.last_dx(), outs, st):
        e = _rel(dx, dxr[0])
        print("n", len(dx), "dx rel err", e, "refinement steps", s["refine_steps"], flush=True)
        assert s["chol_ok"]
        assert e < (1e-9 if precision == "fp64" else 1e-6)
        assert s["accepted"] == sr["accepted"]


def test_tiled_and_one_workgroup_windows_in_one_batch_with_graph_replay():
    """Large and small windows solved together for three LM iterations from a captured graph."""
    pbs = _large_windows() + [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[6, i])[0]
                              for i, (P, L) in enumerate([(5, 80), (12, 600)])]
    refs = [pb.copy() for pb in pbs]
    outs = [O.window_solve(r, iters=3, history=True) for r in refs]
    bs, st = _gpu_solve(pbs, 3, "fp64", use_graph=True)
    for pb, ref, (sr, _d, _a), s in zip(pbs, refs, outs, st):
        assert s["accepted"] == sr["accepted"]
        assert _rel(pb.pwb, ref.pwb) < 1e-9
        assert _rel(pb.Rwb, ref.Rwb) < 1e-9
        assert _rel(pb.lmp, ref.lmp) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_side_stream_solves_graph_replays_match_eager_every_iteration(precision):
    """Windows on both sides of the tiling threshold: the small ones are solved on the second stream beside the tiled
    sequence and joined back.  Four replays of the captured LM iteration, each compared with the eager iteration from
    the same state, with host synchronisation and eager copies between the replays."""
    from xclean_amd.window import BatchSolver
    pbs = _large_windows() + [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[8, i])[0]
                              for i, (P, L) in enumerate([(5, 80), (12, 600), (25, 900)])]
    a = [pb.copy() for pb in pbs]
    b = [pb.copy() for pb in pbs]
    A = BatchSolver(a, precision=precision, use_graph=False)
    B = BatchSolver(b, precision=precision, use_graph=True)
    A.upload()
    B.upload()
    bar = 1e-12 if precision == "fp64" else 1e-10
    for it in range(4):
        A.iterate(1)
        B.iterate(1)
        A.download()
        B.download()
        for pa, pg in zip(a, b):
            assert _rel(pg.pwb, pa.pwb) < bar, it
            assert _rel(pg.Rwb, pa.Rwb) < bar, it
            assert _rel(pg.lmp, pa.lmp) < bar, it
    assert [s["accepted"] for s in A.stats()] == [s["accepted"] for s in B.stats()]


def test_config1_visual_inertial_window_mixed_and_fp64_match_oracle():
    """Config 1: 20 visual-inertial keyframes (n = 300, tiled Cholesky), 20,000 landmarks, ~110,000 observation
    rows, 19 IMU factors.  FP64 within 1e-9 and mixed precision within 1e-6 of the oracle's first update."""
    pb = synth.make_config(1)[0]
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=1, history=True)
    print("config 1", pb.P, pb.n_state(), pb.L, pb.M, flush=True)
    for precision, bar in (("fp64", 1e-9), ("mixed", 1e-6)):
        bs, st = _gpu_solve([pb.copy()], 1, precision, use_graph=False)
        e = _rel(bs.last_dx()[0], dx_ref[0])
        print(precision, "dx rel err", e, "refinement steps", st[0]["refine_steps"], flush=True)
        assert st[0]["chol_ok"]
        assert e < bar
        assert st[0]["accepted"] == st_ref["accepted"]


def _wide_tracks_window():
    """48 keyframes, 800 landmarks; where a landmark is visible from a keyframe 32 or more slots from its host, its last
    observation is moved to such a keyframe (measurement re-projected from the truth plus 1 px noise), so about 200
    landmarks span more slots than one work-item tile holds and get compacted tiles of the slots they touch."""
    pb, tr = synth.make_window(P=48, L=800, views=(2, 8), seed=[7, 0])
    Rwc = np.einsum("pij,jk->pik", tr["Rwb"], pb.Rbc)
    pwc = tr["pwb"] + np.einsum("pij,j->pi", tr["Rwb"], pb.pbc)
    rng = np.random.default_rng(3)
    for l in range(pb.L):
        h = pb.lm_host[l]
        a, b, rho = tr["lmp"][l]
        X = Rwc[h] @ np.array([a, b, 1.0]) / rho + pwc[h]
        q = np.einsum("pji,pj->pi", Rwc, X[None] - pwc)
        with np.errstate(all="ignore"):
            u = pb.fx * q[:, 0] / q[:, 2] + pb.cx
            v = pb.fy * q[:, 1] / q[:, 2] + pb.cy
        vis = (q[:, 2] > 0.3) & (u >= 0) & (u < 752) & (v >= 0) & (v < 480)
        far = np.nonzero(vis & (np.abs(np.arange(pb.P) - h) >= 32))[0]
        k = pb.ob_start[l + 1] - 1
        if len(far) == 0 or pb.ob_pose[k] == h:
            continue
        t = rng.choice(far)
        pb.ob_pose[k] = t
        pb.ob_z[k] = [u[t] + rng.normal(), v[t] + rng.normal()]
    return pb


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_landmarks_spanning_more_than_a_tile_match_oracle(precision):
    from xclean_amd import roofline
    pb = _wide_tracks_window()
    span = roofline.landmark_span(pb.lm_host, pb.ob_start, pb.ob_pose)
    assert (span > 32).sum() > 150
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=3, history=True)
    bs, st = _gpu_solve([pb], 3, precision, use_graph=True)
    print(precision, "wide-span landmarks", int((span > 32).sum()), "plan", bs.plan_info(), flush=True)
    assert st[0]["accepted"] == st_ref["accepted"]
    bar = 1e-9 if precision == "fp64" else 1e-6
    assert _rel(pb.pwb, ref.pwb) < bar
    assert _rel(pb.lmp, ref.lmp) < bar
    # the first update, from a fresh solve
    pb2 = _wide_tracks_window()
    bs2, _ = _gpu_solve([pb2], 1, precision, use_graph=False)
    assert _rel(bs2.last_dx()[0], dx_ref[0]) < bar


def test_config4_shaped_window_fp64_matches_oracle():
    """60 visual-inertial keyframes (n = 900: tiled Cholesky), 20,000 landmarks with 2-13 views (some spanning more
    than 32 keyframes), one LM iteration, FP64 and mixed precision against the oracle."""
    pb = synth.make_window(P=60, L=20000, views=(2, 13), seed=[5, 0], vi=True)[0]
    ref = pb.copy()
    st_ref, dx_ref, _ = O.window_solve(ref, iters=1, history=True)
    for precision, bar in (("fp64", 1e-9), ("mixed", 1e-6)):
        bs, st = _gpu_solve([pb.copy()], 1, precision, use_graph=False)
        e = _rel(bs.last_dx()[0], dx_ref[0])
        print("n 900", precision, "dx rel err", e, "refinement steps", st[0]["refine_steps"], flush=True)
        assert st[0]["chol_ok"]
        assert e < bar
        assert st[0]["accepted"] == st_ref["accepted"]
