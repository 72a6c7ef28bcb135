"""The mixed-precision solve's safety net and the window model's special branches, on the GPU against the oracle.

- Ill-conditioned windows (no gauge prior and lambda 1e-8, so the six gauge directions of the damped reduced
  system are held only by the damping: Jacobi-scaled condition ~1e9, beyond FP32 refinement, the regime of
  backend/kernels.py:12-16): the FP32 refinement cannot converge, the window is flagged and solved in FP64, and its
  update and accept decision match the FP64 oracle.  The same for a window forced onto the tiled factorisation.
- A well-conditioned batch takes no fallback and its refinement converges; with the plan options that factor the
  tiled windows or every window in FP64, those windows take one solve and the update matches the FP64 path.
- A landmark behind one of its cameras at the initial state (vio/window.py:483-485): build_system skips the row but
  the trial cost adds 1e6, so every step is rejected and lambda grows x10 per iteration, as in the oracle.
- The noise-free truth: every whitened chi2 of the GPU linearisation is zero to rounding (measurements made by the
  generator from the true world points).
"""
import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _ill(seed, P=8, L=400):
    pb, _ = synth.make_window(P=P, L=L, views=(2, min(P, 8)), seed=[41, seed], gauge=False)
    assert pb.prior is None
    return pb


def _cond_scaled(pb, lam):
    _c, H6, b6, *_ = O.window_build(pb.copy(), lam)
    H, _b, _ = O.window_full_system(pb.copy(), H6, b6)
    Hd = H + np.diag(lam * np.diag(H) + 1e-9)
    s = 1.0 / np.sqrt(np.diag(Hd))
    return np.linalg.cond(Hd * s[:, None] * s[None, :])


@pytest.mark.parametrize("lam", [1e-8, 1e-6])
@pytest.mark.parametrize("big_n", [0, 32])
def test_ill_conditioned_window_falls_back_to_fp64_and_matches_the_oracle(big_n, lam):
    """lambda 1e-8: condition ~3e8, the FP32 factor fails or its refinement diverges.  lambda 1e-6: condition ~3e6,
    the FP32 factor succeeds and its refinement contracts too slowly to meet the test within the six solves."""
    from xclean_amd.window import BatchSolver
    pbs = [_ill(0), _ill(1, P=6, L=300)]
    conds = [_cond_scaled(pb, lam) for pb in pbs]
    print("scaled condition numbers", conds)
    assert min(conds) > (1e8 if lam < 1e-7 else 1e6)
    refs = [pb.copy() for pb in pbs]
    outs = [O.window_solve(r, iters=1, lam0=lam, history=True) for r in refs]
    bs = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=False, big_n=big_n)
    bs.upload(lam)
    bs.iterate(1)
    st = bs.stats()
    dx = bs.last_dx()
    for w, ((sr, dxr, ar), s) in enumerate(zip(outs, st)):
        print(w, "fallback", s["fallback"], "refine", s["refine_steps"], "dx err", _rel(dx[w], dxr[0]))
        assert s["fallback"] == 1 and s["chol_ok"] and not s["refine_converged"]
        if lam > 1e-7:
            assert s["refine_steps"] == 6   # factored in FP32, refinement stalled
        assert s["accepted"] == int(ar[0])
        assert _rel(dx[w], dxr[0]) < 1e-6
        assert abs(s["cost"] - sr["cost"]) <= 1e-6 * abs(sr["cost"])
    # the FP64 GPU path agrees and never flags a window
    bs64 = BatchSolver([pb.copy() for pb in pbs], precision="fp64", use_graph=False, big_n=big_n)
    bs64.upload(lam)
    bs64.iterate(1)
    for w, d in enumerate(bs64.last_dx()):
        assert _rel(dx[w], d) < 1e-6
        assert bs64.stats()[w]["fallback"] == 0


def test_ill_conditioned_window_several_iterations_match_the_oracle():
    from xclean_amd.window import BatchSolver
    pb = _ill(2)
    ref = pb.copy()
    sr, dxr, ar = O.window_solve(ref, iters=4, lam0=1e-8, history=True)
    bs = BatchSolver([pb], precision="mixed", use_graph=True)
    bs.upload(1e-8)
    bs.iterate(4)
    s = bs.stats()[0]
    assert s["fallback"] >= 1
    assert s["accepted"] == sr["accepted"]
    assert abs(s["lam"] - sr["lam"]) <= 1e-12 * sr["lam"]
    bs.download()
    assert _rel(pb.pwb, ref.pwb) < 1e-6
    assert _rel(pb.lmp, ref.lmp) < 1e-6


def test_well_conditioned_batch_takes_no_fallback():
    from xclean_amd.window import BatchSolver
    pbs = [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[42, i])[0]
           for i, (P, L) in enumerate([(5, 200), (12, 700), (30, 900), (20, 600)])]
    for big_n in (0, 64):
        bs = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=False, big_n=big_n)
        bs.upload(1e-4)
        bs.iterate(3)
        for s in bs.stats():
            assert s["fallback"] == 0 and s["refine_converged"] and s["chol_ok"]
            assert 1 <= s["refine_steps"] <= 6


@pytest.mark.parametrize("big_fp64", [1, 2])
def test_fp64_factor_plan_options_match_the_fp64_path(big_fp64):
    """Mixed precision with the tiled windows (big_fp64 = 1) or every window (2) factored in FP64: one solve per
    such window, no fallback, and the first update within 1e-6 of the FP64 path (whose only difference is the FP64
    observation record)."""
    from xclean_amd.window import BatchSolver
    pbs = [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[42, i])[0]
           for i, (P, L) in enumerate([(5, 200), (12, 700), (30, 900), (20, 600)])]
    ref = BatchSolver([pb.copy() for pb in pbs], precision="fp64", use_graph=False, big_n=64)
    ref.upload(1e-4)
    ref.iterate(1)
    dx_ref = ref.last_dx()
    bs = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=False, big_n=64, big_fp64=big_fp64)
    bs.upload(1e-4)
    bs.iterate(1)
    dx = bs.last_dx()
    for pb, a, b, st in zip(pbs, dx, dx_ref, bs.stats()):
        assert _rel(a, b) < 1e-6
        assert st["fallback"] == 0 and st["chol_ok"] and st["refine_converged"]
        if pb.n_state() > 64 or big_fp64 == 2:
            assert st["refine_steps"] == 1  # factored in FP64: one solve
        else:
            assert 1 <= st["refine_steps"] <= 6


def _behind_window():
    """A landmark seen by its host and one other keyframe, moved to a point 0.2 m behind that keyframe's camera
    (still in front of the host), with its host pixel and inverse-depth measurement made consistent with the new
    point: its gradient vanishes, so the step leaves it behind the camera and every trial costs the 1e6."""
    pb, _ = synth.make_window(P=8, L=300, views=(2, 8), seed=[33, 1])
    Rwc = np.einsum("pij,jk->pik", pb.Rwb, pb.Rbc)
    pwc = pb.pwb + np.einsum("pij,j->pi", pb.Rwb, pb.pbc)
    for l in range(pb.L):
        k0, k1 = pb.ob_start[l], pb.ob_start[l + 1]
        h = pb.lm_host[l]
        if k1 - k0 != 2:
            continue
        kt = [k for k in range(k0, k1) if pb.ob_pose[k] != h]
        kh = [k for k in range(k0, k1) if pb.ob_pose[k] == h]
        if len(kt) != 1 or len(kh) != 1:
            continue
        t = pb.ob_pose[kt[0]]
        X = pwc[t] + Rwc[t] @ np.array([0.05, -0.03, -0.2])
        Xh = Rwc[h].T @ (X - pwc[h])
        if Xh[2] > 0.3:
            a, b, rho = Xh[0] / Xh[2], Xh[1] / Xh[2], 1.0 / Xh[2]
            pb.lmp[l] = [a, b, rho]
            pb.ob_z[kh[0]] = [pb.fx * a + pb.cx, pb.fy * b + pb.cy]
            pb.rho_meas[l] = rho
            return pb
    raise AssertionError("no landmark to move behind a camera")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_behind_camera_row_rejects_every_step_like_the_oracle(precision):
    from xclean_amd.window import BatchSolver
    pb = _behind_window()
    ref = pb.copy()
    sr, dxr, ar = O.window_solve(ref, iters=3, history=True)
    assert not ar.any() and abs(sr["lam"] - 1e-4 * 1e3) <= 1e-12 * sr["lam"]
    bs = BatchSolver([pb], precision=precision, use_graph=False)
    bs.upload(1e-4)
    bs.iterate(3)
    s = bs.stats()[0]
    assert s["accepted"] == 0
    assert abs(s["lam"] - sr["lam"]) <= 1e-12 * sr["lam"]
    assert abs(s["cost0"] - sr["cost0"]) <= 1e-9 * abs(sr["cost0"])
    assert abs(s["cost"] - sr["cost"]) <= 1e-9 * abs(sr["cost"])
    tol = 1e-9 if precision == "fp64" else 1e-6
    assert _rel(bs.last_dx()[0], dxr[2]) < tol


@pytest.mark.parametrize("vi", [False, True])
def test_gpu_chi2_vanishes_at_the_noise_free_truth(vi):
    from xclean_amd.window import BatchSolver
    pb, tr = synth.make_window(P=6, L=300, views=(2, 6), seed=[31, 7], vi=vi, stereo="ur", outlier_frac=0.0,
                               noise_scale=0.0)
    pb.Rwb[:], pb.pwb[:], pb.lmp[:] = tr["Rwb"], tr["pwb"], tr["lmp"]
    pb.prior = None
    pb.imu = [] if vi else pb.imu
    bs = BatchSolver([pb], precision="fp64", use_graph=False)
    bs.upload(0.0)
    r = bs.linearize(0.0)[0]
    chi = r["ob_chi2"]
    assert np.isfinite(chi).all()
    assert chi.max() < 1e-18            # whitened squared residual: |e| < 1e-9 px at sigma 1 px
    assert np.abs(r["b"]).max() < 1e-6 * max(1.0, np.abs(r["H"]).max()) ** 0.5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_failed_cholesky_raises_lambda_and_keeps_the_state_like_the_oracle(precision):
    """A non-finite prior makes the damped reduced system non-finite: the Cholesky fails on every iteration
    (vio/window.py:936-941), lambda grows x10 per iteration, nothing is accepted and the state is not touched."""
    from xclean_amd.window import BatchSolver
    pb, _ = synth.make_window(P=6, L=200, views=(2, 6), seed=[3, 3])
    pb.prior.S[0, 0] = np.nan
    ref = pb.copy()
    sr, dxr, ar = O.window_solve(ref, iters=3, lam0=1e-4, history=True)
    assert not ar.any() and abs(sr["lam"] - 0.1) <= 1e-12
    before = (pb.Rwb.copy(), pb.pwb.copy(), pb.lmp.copy())
    bs = BatchSolver([pb], precision=precision, use_graph=True)
    bs.upload(1e-4)
    bs.iterate(3)
    s = bs.stats()[0]
    assert not s["chol_ok"] and s["accepted"] == 0
    assert abs(s["lam"] - sr["lam"]) <= 1e-12 * sr["lam"]
    bs.download()
    for a, b in zip((pb.Rwb, pb.pwb, pb.lmp), before):
        assert np.array_equal(a, b)


def test_landmark_with_more_rows_than_a_wavefront_is_refused():
    """A landmark with more than XBA_MAX_OBS_PER_LM (64) observation rows: the planner refuses the window with
    XBA_E_TOO_MANY_OBS (one wavefront lane per row of a landmark) instead of solving something else."""
    from xclean_amd import _lib
    from xclean_amd.window import BatchSolver
    pb, _ = synth.make_window(P=8, L=50, views=(2, 8), seed=[3, 4])
    k0, k1 = pb.ob_start[0], pb.ob_start[1]
    rep = np.concatenate([np.arange(k0, k1)] * (70 // (k1 - k0) + 1))[:70]
    rest = np.arange(k1, pb.M)
    idx = np.concatenate([rep, rest])
    for k in ("ob_pose", "ob_cam", "ob_z", "ob_isig", "ob_valid"):
        setattr(pb, k, np.ascontiguousarray(getattr(pb, k)[idx]))
    pb.ob_start = np.concatenate([[0], pb.ob_start[1:] - (k1 - k0) + 70])
    assert pb.ob_start[1] == 70 and pb.ob_start[-1] == pb.M
    with pytest.raises(_lib.XbaError, match="XBA_MAX_OBS_PER_LM"):
        BatchSolver([pb], precision="fp64", use_graph=False)
