"""Pins the oracle's window model (oracle/xba_oracle.c restating vio/window.py and vio/marg.py) against facts fixed
outside the oracle, so that a sign slip or a transposed rotation in the shared reading of the reference would fail:

- residuals: exact projections of the true world points (made by the generator from world coordinates, never from
  the oracle's inverse-depth formula) give zero residuals at the true state, left and right-camera rows, VO and VI
  with a non-identity body-camera transform; at arbitrary states the oracle's residual equals a world-point
  pinhole projection written here in numpy (vio/window.py:125-146 in another parametrisation);
- the trial cost: a row whose point lies behind its camera costs exactly 1e6 (vio/window.py:483-485), the other
  rows their Huber cost computed here;
- the prior terms (vio/marg.py:102-127): against central differences of the prior cost 1/2 ||S d + r||^2 with d
  formed here with scipy's rotation log, and against S^T (S d + r) away from x0;
- the LM driver: it moves a clean window's poses toward the truth (the window analogue of
  tests/test_ba_solver.py:26-33), and one Gauss-Newton step from the truth of a noise-free window is zero.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from xclean_amd import synth
from xclean_amd.problem import Prior


def _cam(pb, Rwb, pwb):
    """T_wc = T_wb T_bc for every slot."""
    Rwc = np.einsum("pij,jk->pik", Rwb, pb.Rbc)
    pwc = pwb + np.einsum("pij,j->pi", Rwb, pb.pbc)
    return Rwc, pwc


def _world_point_residual(pb, Rwb, pwb, lmp, l, k):
    """The residual of observation k of landmark l by the pinhole projection of the landmark's world point."""
    Rwc, pwc = _cam(pb, Rwb, pwb)
    h, t = pb.lm_host[l], pb.ob_pose[k]
    a, b, rho = lmp[l]
    X = Rwc[h] @ (np.array([a, b, 1.0]) / rho) + pwc[h]
    Xc = Rwc[t].T @ (X - pwc[t])
    if pb.ob_cam[k] == 0:
        return np.array([pb.fx * Xc[0] / Xc[2] + pb.cx - pb.ob_z[k, 0], pb.fy * Xc[1] / Xc[2] + pb.cy - pb.ob_z[k, 1]])
    return np.array([pb.fx * (Xc[0] - pb.B) / Xc[2] + pb.cx - pb.ob_z[k, 0]])


def _at_truth(pb, tr):
    q = pb.copy()
    q.Rwb[:] = tr["Rwb"]
    q.pwb[:] = tr["pwb"]
    q.lmp[:] = tr["lmp"]
    return q


CASES = [dict(vi=False, stereo="prior"), dict(vi=False, stereo="ur"), dict(vi=True, stereo="prior"),
         dict(vi=True, stereo="ur")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % ("vi" if c["vi"] else "vo", c["stereo"]))
def test_residuals_vanish_at_the_noise_free_truth(case):
    pb, tr = synth.make_window(P=6, L=120, views=(2, 6), seed=[31, int(case["vi"]), len(case["stereo"])],
                               outlier_frac=0.0, noise_scale=0.0, **case)
    if case["vi"]:
        assert np.abs(pb.Rbc - np.eye(3)).max() > 0.1 and np.abs(pb.pbc).max() > 0.01
    q = _at_truth(pb, tr)
    worst = {0: 0.0, 1: 0.0}
    for l in range(q.L):
        for k in range(q.ob_start[l], q.ob_start[l + 1]):
            e, *_ = O.obs_residual_jac(q, l, k)
            worst[int(q.ob_cam[k])] = max(worst[int(q.ob_cam[k])], float(np.abs(e).max()))
    assert worst[0] < 1e-9
    if case["stereo"] == "ur":
        assert (q.ob_cam == 1).any() and worst[1] < 1e-9
    # the robust cost of the visual rows (and the exact inverse-depth priors) is zero there too
    assert O.visual_cost(q, q.Rwb, q.pwb, q.lmp) < 1e-15
    # while the generator's perturbed initial state is not at the truth
    assert O.visual_cost(pb, pb.Rwb, pb.pwb, pb.lmp) > 1.0


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % ("vi" if c["vi"] else "vo", c["stereo"]))
def test_residuals_equal_the_world_point_projection_at_random_states(case):
    pb, _ = synth.make_window(P=5, L=60, views=(2, 5), seed=[32, int(case["vi"]), len(case["stereo"])], **case)
    rng = np.random.default_rng(5)
    Rwb = np.einsum("pij,pjk->pik", pb.Rwb, Rotation.from_rotvec(rng.normal(0, 0.02, (pb.P, 3))).as_matrix())
    pwb = pb.pwb + rng.normal(0, 0.05, pb.pwb.shape)
    lmp = pb.lmp * (1 + rng.normal(0, 0.02, pb.lmp.shape))
    q = pb.copy()
    q.Rwb[:], q.pwb[:], q.lmp[:] = Rwb, pwb, lmp
    n = 0
    for l in range(q.L):
        for k in range(q.ob_start[l], q.ob_start[l + 1]):
            e, *_ = O.obs_residual_jac(q, l, k)
            want = _world_point_residual(q, Rwb, pwb, lmp, l, k)
            assert e.shape == want.shape
            assert np.abs(e - want).max() <= 1e-9 * max(1.0, np.abs(want).max()), (l, k, e, want)
            n += 1
    assert n == q.M


def _huber_cost(en, k):
    return 0.5 * en * en if en <= k else k * (en - 0.5 * k)


def test_a_row_behind_its_camera_costs_1e6():
    """vio/window.py:477-489: a row with q2 <= 1e-9 adds 1e6 to the trial cost and nothing else; the landmark's
    other rows and its inverse-depth prior add their Huber costs."""
    pb, _ = synth.make_window(P=8, L=200, views=(3, 8), seed=[33, 0])
    Rwc, pwc = _cam(pb, pb.Rwb, pb.pwb)
    # a landmark and one of its observers such that a point just behind that camera is in front of the host
    found = None
    for l in range(pb.L):
        h = pb.lm_host[l]
        for k in range(pb.ob_start[l], pb.ob_start[l + 1]):
            t = pb.ob_pose[k]
            if t == h:
                continue
            X = pwc[t] + Rwc[t] @ np.array([0.05, -0.03, -0.2])
            Xh = Rwc[h].T @ (X - pwc[h])
            if Xh[2] > 0.3:
                found = (l, t, Xh)
                break
        if found:
            break
    assert found is not None
    l, t, Xh = found
    q = pb.copy()
    q.lmp[l] = [Xh[0] / Xh[2], Xh[1] / Xh[2], 1.0 / Xh[2]]
    masked = q.copy()
    masked.lm_mask[l] = False
    extra = 0.0
    behind = 0
    for k in range(q.ob_start[l], q.ob_start[l + 1]):
        Xc = Rwc[q.ob_pose[k]].T @ (Rwc[q.lm_host[l]] @ (np.array([q.lmp[l, 0], q.lmp[l, 1], 1.0]) / q.lmp[l, 2])
                                    + pwc[q.lm_host[l]] - pwc[q.ob_pose[k]])
        if q.lmp[l, 2] * Xc[2] <= 1e-9:
            extra += 1e6
            behind += 1
        else:
            e = _world_point_residual(q, q.Rwb, q.pwb, q.lmp, l, k) * q.ob_isig[k]
            extra += _huber_cost(float(np.linalg.norm(e)), q.huber_k)
    if q.rho_isig[l] > 0:
        extra += _huber_cost(abs((q.lmp[l, 2] - q.rho_meas[l]) * q.rho_isig[l]), q.huber_k)
    assert behind >= 1
    c_full = O.visual_cost(q, q.Rwb, q.pwb, q.lmp)
    c_masked = O.visual_cost(masked, masked.Rwb, masked.pwb, masked.lmp)
    assert abs((c_full - c_masked) - extra) <= 1e-9 * extra


def _boxminus(R, p, v, bg, ba, x0, dim):
    R0, p0, v0, bg0, ba0 = x0
    d = [Rotation.from_matrix(R0.T @ R).as_rotvec(), p - p0]
    if dim == 15:
        d += [v - v0, bg - bg0, ba - ba0]
    return np.concatenate(d)


def _prior_cost(pr, Rwb, pwb, vel, bgs, bas):
    d = np.concatenate([_boxminus(Rwb[i], pwb[i], None if vel is None else vel[i], None if bgs is None else bgs[i],
                                  None if bas is None else bas[i], pr.x0[k], pr.dims[k])
                        for k, i in enumerate(pr._idx)])
    res = pr.S @ d + pr.r
    return 0.5 * float(res @ res), d


def _random_prior(pb, slots, dims, rng, shift=0.0):
    x0 = []
    for s, dm in zip(slots, dims):
        R0 = pb.Rwb[s] @ Rotation.from_rotvec(rng.normal(0, shift, 3)).as_matrix()
        p0 = pb.pwb[s] + rng.normal(0, shift, 3)
        if dm == 15:
            x0.append((R0, p0, pb.vel[s] + rng.normal(0, shift, 3), pb.bgs[s] + rng.normal(0, shift, 3),
                       pb.bas[s] + rng.normal(0, shift, 3)))
        else:
            x0.append((R0, p0, None, None, None))
    m = int(sum(dims))
    S = rng.normal(0, 1.0, (m + 2, m)) * 3.0
    r = rng.normal(0, 0.5, m + 2)
    pr = Prior(slots, dims, x0, S, r)
    pr.bind({s: s for s in slots})
    return pr


def _perturb(state, s, c, h):
    """State with coordinate c of slot s's block moved by h (rotation: right perturbation R Exp(h e_c))."""
    Rwb, pwb, vel, bgs, bas = (None if a is None else a.copy() for a in state)
    if c < 3:
        Rwb[s] = Rwb[s] @ Rotation.from_rotvec(h * np.eye(3)[c]).as_matrix()
    elif c < 6:
        pwb[s, c - 3] += h
    elif c < 9:
        vel[s, c - 6] += h
    elif c < 12:
        bgs[s, c - 9] += h
    else:
        bas[s, c - 12] += h
    return Rwb, pwb, vel, bgs, bas


@pytest.mark.parametrize("vi", [False, True])
def test_prior_terms_against_central_differences_of_the_prior_cost(vi):
    """At x0 (d = 0) the prior's gradient is the derivative of 1/2 ||S d + r||^2 under the solver's perturbation
    and its Hessian the (exact) second derivative; away from x0, b = S^T (S d + r) and the cost follow d."""
    pb, _ = synth.make_window(P=5, L=30, views=(2, 5), seed=[34, int(vi)], vi=vi)
    rng = np.random.default_rng(7)
    slots, dims = ([1, 3], [15, 6]) if vi else ([1, 3], [6, 6])
    off = np.concatenate([[0], np.cumsum(pb.dims if vi else np.full(pb.P, 6))])
    cols = np.concatenate([np.arange(off[s], off[s] + d) for s, d in zip(slots, dims)])
    n = pb.n_state()
    for shift in (0.0, 0.05):
        q = pb.copy()
        q.imu = []                       # the prior alone in the full system
        q.prior = _random_prior(q, slots, dims, rng, shift)
        state = (q.Rwb, q.pwb, q.vel, q.bgs, q.bas)
        H, b, c = O.window_full_system(q, np.zeros((6 * q.P, 6 * q.P)), np.zeros(6 * q.P))
        c_np, d = _prior_cost(q.prior, *state)
        assert abs(c - c_np) <= 1e-12 * c_np
        other = np.setdiff1d(np.arange(n), cols)
        assert np.abs(H[other]).max() == 0.0 and np.abs(b[other]).max() == 0.0
        S, r = q.prior.S, q.prior.r
        assert np.abs(H[np.ix_(cols, cols)] - S.T @ S).max() <= 1e-12 * np.abs(S.T @ S).max()
        assert np.abs(b[cols] - S.T @ (S @ d + r)).max() <= 1e-12 * np.abs(S.T @ (S @ d + r)).max()
        if shift == 0.0:
            assert np.abs(d).max() < 1e-12
            h = 1e-4
            k = 0
            grad = np.zeros(len(cols))
            hess = np.zeros((len(cols), len(cols)))
            coords = [(s, cc) for s, dm in zip(slots, dims) for cc in range(dm)]
            for a, (s, cc) in enumerate(coords):
                cp = _prior_cost(q.prior, *_perturb(state, s, cc, h))[0]
                cm = _prior_cost(q.prior, *_perturb(state, s, cc, -h))[0]
                grad[a] = (cp - cm) / (2 * h)
                hess[a, a] = (cp - 2 * c_np + cm) / (h * h)
                k += 1
            assert np.abs(grad - b[cols]).max() <= 1e-6 * np.abs(b[cols]).max()
            assert np.abs(np.diag(hess) - np.diag(H)[cols]).max() <= 1e-5 * np.abs(np.diag(H)[cols]).max()
            # an off-diagonal entry by the mixed second difference
            (s1, c1), (s2, c2) = coords[0], coords[-1]
            pp = _prior_cost(q.prior, *_perturb(_perturb(state, s1, c1, h), s2, c2, h))[0]
            pm = _prior_cost(q.prior, *_perturb(_perturb(state, s1, c1, h), s2, c2, -h))[0]
            mp = _prior_cost(q.prior, *_perturb(_perturb(state, s1, c1, -h), s2, c2, h))[0]
            mm = _prior_cost(q.prior, *_perturb(_perturb(state, s1, c1, -h), s2, c2, -h))[0]
            mixed = (pp - pm - mp + mm) / (4 * h * h)
            assert abs(mixed - H[cols[0], cols[-1]]) <= 1e-5 * np.abs(H[np.ix_(cols, cols)]).max()


def _pose_errors(pb, tr):
    dp = np.linalg.norm(pb.pwb - tr["pwb"], axis=1)
    dr = np.array([np.linalg.norm(Rotation.from_matrix(tr["Rwb"][i].T @ pb.Rwb[i]).as_rotvec()) for i in range(pb.P)])
    return float(np.sqrt(np.mean(dp[1:] ** 2))), float(np.sqrt(np.mean(dr[1:] ** 2)))


@pytest.mark.parametrize("vi", [False, True])
def test_window_solve_moves_poses_toward_the_truth(vi):
    """The window analogue of tests/test_ba_solver.py:26-33: on a clean window (1 px noise, no outliers) five LM
    iterations cut the position error of the free keyframes below 0.6 of the initial one and lower the cost."""
    pb, tr = synth.make_window(P=8, L=600, views=(2, 8), seed=[35, int(vi)], vi=vi, outlier_frac=0.0,
                               pose_sigma_t=0.05, pose_sigma_r_deg=1.0)
    e0, r0 = _pose_errors(pb, tr)
    q = pb.copy()
    st = O.window_solve(q, iters=5)
    e1, r1 = _pose_errors(q, tr)
    assert st["cost"] < st["cost0"]
    assert e1 < 0.6 * e0, (e0, e1)
    assert r1 < 0.6 * r0, (r0, r1)


def test_gauss_newton_step_from_the_noise_free_truth_is_zero():
    """Measurements made by the generator from the true world points (no oracle residual involved): the window is
    at its minimum at the truth, so the step there is zero."""
    pb, tr = synth.make_window(P=5, L=80, views=(2, 5), seed=[9, 4], outlier_frac=0.0, noise_scale=0.0,
                               stereo="ur")
    q = _at_truth(pb, tr)
    q.prior = None
    st, dxh, _ = O.window_solve(q.copy(), iters=1, history=True)
    assert st["cost0"] < 1e-15
    assert np.abs(dxh[0]).max() < 1e-9
