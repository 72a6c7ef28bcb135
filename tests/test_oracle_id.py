"""The oracle's restatement of the inverse-depth local BA (backend/ba.py:1212-1342 _solve_np_id), pinned by:
- golden vectors written by the reference's own _solve_np_id (tests/golden/make_golden.py) on three problems:
  out-of-window hosts, a held-keyframe mask with the bearing / inverse-range prior, rotation factors and an active
  inverse-depth floor, and a hard problem with rejected steps -- poses, inverse depths, bearings and stats at 1e-10;
- central finite differences of the residual for the three analytic Jacobian blocks (observer pose, host pose,
  landmark (x_n, y_n, rho)) of _analytic_blocks_id (ba.py:1138-1209), including the observer == host case where
  the two pose blocks cancel.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


class _Rig:
    def __init__(self, fx, cx, cy):
        self.fx_rect, self.cx, self.cy = fx, cx, cy


def load_id_case(name):
    z = np.load(os.path.join(GOLD, "solve_np_id_%s.npz" % name))
    d = {k: z[k] for k in z.files}
    prior = None
    if "pr_m0" in d:
        prior = (d["pr_m0"], d["pr_rho0"], d["pr_wbear"], d["pr_invsig"], d["pr_valid"])
    rotf = None
    if "rot_ab" in d:
        rotf = [(int(a), int(b), M, float(s)) for (a, b), M, s in zip(d["rot_ab"], d["rot_M"], d["rot_sig"])]
    free = d["free"]
    free = int(free) if free.ndim == 0 else free.astype(bool)
    return d, _Rig(float(d["fx"]), float(d["cx"]), float(d["cy"])), free, prior, rotf


@pytest.mark.parametrize("case", ["default", "mask_prior_rotf", "rejects"])
def test_oracle_solve_np_id_matches_reference_golden(case):
    d, rig, free, prior, rotf = load_id_case(case)
    p, rho, mb, st = O.solve_np_id(d["poses0"], d["rho0"], d["mbear0"], d["hidx"], d["hpose_ext"], d["okf"], d["olm"],
                                   d["ouv"], rig, free, int(d["iters"]), float(d["lam0"]), prior=prior, rotf=rotf,
                                   rho_floor=float(d["rho_floor"]))
    np.testing.assert_allclose(p, d["out_poses"], rtol=0, atol=1e-10)
    # inverse depths within 1e-9 relative: on the hard problem (551 px initial rms) the far points' depths are the
    # worst-conditioned unknowns and the summation order of the Schur terms shows at ~6e-10
    np.testing.assert_allclose(rho, d["out_rho"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(mb, d["out_mbear"], rtol=0, atol=1e-10)
    np.testing.assert_allclose([st["chi2_0"], st["chi2_n"], st["rms_0"], st["rms_n"]], d["stats"], rtol=1e-10)


def _residual(Tcw_o, Twc_h, mb, rho, fx, cx, cy, uv):
    Ph = mb / max(rho, 1e-12)
    Pw = Twc_h[:3, :3] @ Ph + Twc_h[:3, 3]
    Pc = Tcw_o[:3, :3] @ Pw + Tcw_o[:3, 3]
    z = max(Pc[2], 1e-6)
    return np.array([fx * Pc[0] / z + cx - uv[0], fx * Pc[1] / z + cy - uv[1]]), Pc, Ph


@pytest.mark.parametrize("same", [False, True])
def test_id_jacobian_blocks_by_central_finite_differences(same):
    rng = np.random.default_rng(5)
    fx, cx, cy = 458.654, 367.215, 248.375
    Twc_h = synth.se3_exp(rng.normal(0, 0.2, 6))
    Twc_o = Twc_h if same else Twc_h @ synth.se3_exp(np.r_[rng.normal(0, 0.3, 3), rng.normal(0, 0.05, 3)])
    Tcw_o = synth.se3_inv(Twc_o)
    mb = np.array([0.12, -0.07, 1.0])
    rho = 0.25
    uv = np.array([300.0, 200.0])
    r, Pc, Ph = _residual(Tcw_o, Twc_h, mb, rho, fx, cx, cy, uv)
    Jo, Jh, Jl, rw = O.id_blocks(Tcw_o, Twc_h[:3, :3], rho, Pc, Ph, True, fx, r, 1.0)
    np.testing.assert_allclose(rw, r, rtol=1e-14)
    h = 1e-6
    Fo, Fh, Fl = np.zeros((2, 6)), np.zeros((2, 6)), np.zeros((2, 3))
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        # observer: T_cw <- exp(xi) T_cw; host: T_cw_h <- exp(xi) T_cw_h (T_wc_h <- T_wc_h exp(-xi))
        if same:
            rp = _residual(synth.se3_exp(e) @ Tcw_o, synth.se3_inv(synth.se3_exp(e) @ Tcw_o), mb, rho, fx, cx, cy, uv)[0]
            rm = _residual(synth.se3_exp(-e) @ Tcw_o, synth.se3_inv(synth.se3_exp(-e) @ Tcw_o), mb, rho, fx, cx, cy,
                           uv)[0]
            Fo[:, i] = (rp - rm) / (2 * h)
            continue
        Fo[:, i] = (_residual(synth.se3_exp(e) @ Tcw_o, Twc_h, mb, rho, fx, cx, cy, uv)[0] -
                    _residual(synth.se3_exp(-e) @ Tcw_o, Twc_h, mb, rho, fx, cx, cy, uv)[0]) / (2 * h)
        Fh[:, i] = (_residual(Tcw_o, Twc_h @ synth.se3_exp(-e), mb, rho, fx, cx, cy, uv)[0] -
                    _residual(Tcw_o, Twc_h @ synth.se3_exp(e), mb, rho, fx, cx, cy, uv)[0]) / (2 * h)
    for c in range(3):
        mp, mm = mb.copy(), mb.copy()
        rp_, rm_ = rho, rho
        if c < 2:
            mp[c] += h
            mm[c] -= h
        else:
            rp_, rm_ = rho + h, rho - h
        Fl[:, c] = (_residual(Tcw_o, Twc_h, mp, rp_, fx, cx, cy, uv)[0] -
                    _residual(Tcw_o, Twc_h, mm, rm_, fx, cx, cy, uv)[0]) / (2 * h)
    # d(Ph)/d(x_n) = e1 / rho scaled the bearing: the block is w.r.t. (x_n, y_n, rho) as the reference states
    if same:
        np.testing.assert_allclose(Jo + Jh, Fo, atol=1e-6 * np.abs(Jo).max())   # the pose blocks cancel
        np.testing.assert_allclose(Fo, 0.0, atol=1e-5)
    else:
        np.testing.assert_allclose(Jo, Fo, rtol=1e-6, atol=1e-6 * np.abs(Fo).max())
        np.testing.assert_allclose(Jh, Fh, rtol=1e-6, atol=1e-6 * np.abs(Fh).max())
    np.testing.assert_allclose(Jl, Fl, rtol=1e-6, atol=1e-6 * np.abs(Fl).max())


def test_oracle_solve_np_id_moves_poses_toward_the_truth():
    pb = synth.make_problem(nk=10, nl=300, obs_per=4, seed=21)
    q = synth.to_inverse_depth(pb, ext_every=6)
    p, rho, mb, st = O.solve_np_id(q["poses0"], q["rho0"], q["mbear0"], q["hidx"], q["hpose_ext"], q["okf"],
                                   q["olm"], q["ouv"], q["rig"], 0, 6)
    e0 = np.linalg.norm(q["poses0"][:, :3, 3] - pb["poses_true"][:, :3, 3])
    e1 = np.linalg.norm(p[:, :3, 3] - pb["poses_true"][:, :3, 3])
    assert e1 < 0.6 * e0 and st["rms_n"] < st["rms_0"]
