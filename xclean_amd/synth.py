"""Seeded synthetic BA problems.

- ``make_problem``: the Euclidean local-BA window of backend/ba_synth.py:14-64 (same parameters, same
  ``np.random.RandomState`` call sequence, so the same seed gives the same problem).
- ``make_window``: an inverse-depth stereo window in the form of vio/window.py's ``Problem`` (VO or VI), with the
  camera, noise, outlier and trajectory model of the benchmark configurations:

  ============ ======================================================================================================
  ``cfg0``     10 keyframes (6-DoF), 2,000 landmarks, depth U(1, 20) m, 2-10 views, 1 LM iteration, FP64
  ``cfg1``     20 keyframes (15-DoF) with 200 Hz IMU factors, 20,000 landmarks (~120k observations)
  ``cfg2``     256 windows, 5-30 keyframes, landmarks log-uniform in [500, 30000], 2..min(10, P) views
  ``cfg3``     64 windows x 15 keyframes x 5,000 landmarks
  ``cfg4``     60 VI keyframes, 200,000 landmarks, 2-13 views (mean 7.5 -> ~1.5M observations)
  ============ ======================================================================================================

Camera for every config: fx = fy = 458, cx = 367, cy = 248, 752 x 480, rectified baseline 0.11 m.  Pixel noise
N(0, 1 px) per coordinate; 5 % of the non-host observations are replaced by a uniform pixel over the image.  The
stereo measurement enters as the host inverse-depth prior (the reference's default stereo form, rho_meas =
d / (fx B), sigma_rho = 0.5 / (fx B)); ``stereo="ur"`` adds a right-image u_R observation per view instead.
Poses start perturbed by N(0, 2 cm) per axis and N(0, 0.5 deg) per rotation axis; keyframe 0 is held by the
1e10 gauge prior (vio/marg.py:205-209).  Observation rows counted by the benchmarks are the entries of ``ob_pose``.
"""
from __future__ import annotations

import math

import numpy as np

from .problem import PoseFree, Preint, Problem, gauge_prior

CAM = dict(fx=458.0, fy=458.0, cx=367.0, cy=248.0, W=752, H=480, B=0.11)
HUBER_K = 1.345                      # core/derive.py:47
FEATURE_SIGMA_PX = 1.0               # core/derive.py:61
DISPARITY_NOISE_PX = 0.5             # core/derive.py:58
# EuRoC-like IMU (gyro / accel densities as configured; the random walks are EuRoC's)
IMU_NOISE = dict(sigma_g=1.7e-4, sigma_a=2.0e-3, rw_g=1.9393e-5, rw_a=3.0e-3, rate_hz=200.0)
G_W = np.array([0.0, 0.0, -9.81])


# ------------------------------------------------------------------------------------------------ small SO(3) --
def hat(v):
    v = np.asarray(v, float)
    return np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])


def so3_exp(w):
    """Rodrigues with the series near 0 (closed form of exp on so(3))."""
    w = np.asarray(w, float)
    t2 = float(w @ w)
    t = math.sqrt(t2)
    if t < 0.126:
        a = 1.0 - t2 / 6.0 + t2 * t2 / 120.0 - t2 ** 3 / 5040.0
        b = 0.5 - t2 / 24.0 + t2 * t2 / 720.0 - t2 ** 3 / 40320.0
    else:
        a = math.sin(t) / t
        b = (1.0 - math.cos(t)) / t2
    K = hat(w)
    return np.eye(3) + a * K + b * (K @ K)


def so3_jr(w):
    """Right Jacobian of SO(3)."""
    w = np.asarray(w, float)
    t2 = float(w @ w)
    t = math.sqrt(t2)
    if t < 0.126:
        b = 0.5 - t2 / 24.0 + t2 * t2 / 720.0
        c = 1.0 / 6.0 - t2 / 120.0 + t2 * t2 / 5040.0
    else:
        b = (1.0 - math.cos(t)) / t2
        c = (t - math.sin(t)) / (t2 * t)
    K = hat(w)
    return np.eye(3) - b * K + c * (K @ K)


def se3_exp(xi):
    """[rho, phi] twist -> 4x4 (the convention of core/lie.py:301-314)."""
    xi = np.asarray(xi, float)
    rho, phi = xi[:3], xi[3:]
    t2 = float(phi @ phi)
    t = math.sqrt(t2)
    if t < 0.126:
        b = 0.5 - t2 / 24.0 + t2 * t2 / 720.0 - t2 ** 3 / 40320.0
        c = 1.0 / 6.0 - t2 / 120.0 + t2 * t2 / 5040.0 - t2 ** 3 / 362880.0
    else:
        b = (1.0 - math.cos(t)) / t2
        c = (t - math.sin(t)) / (t2 * t)
    K = hat(phi)
    V = np.eye(3) + b * K + c * (K @ K)
    T = np.eye(4)
    T[:3, :3] = so3_exp(phi)
    T[:3, 3] = V @ rho
    return T


def se3_inv(T):
    out = np.eye(4)
    out[:3, :3] = T[:3, :3].T
    out[:3, 3] = -T[:3, :3].T @ T[:3, 3]
    return out


# ---------------------------------------------------------------------------- the Euclidean drop-in problem --
class Rig:
    """The rig of backend/ba_synth.py:14-17."""
    fx_rect = 458.654
    cx, cy = 367.215, 248.375
    width, height = 752, 480


def make_problem(nk=15, nl=600, obs_per=4, seed=5, pose_noise=0.02, pt_noise=0.05, px_noise=0.5):
    """backend/ba_synth.py:20-64: nk keyframes on an arc, nl points in front, obs_per views per point."""
    rng = np.random.RandomState(seed)
    rig = Rig()
    poses_true = np.stack([se3_exp([0.09 * i, 0.01 * i, 0.0, 0.0, 0.02 * i, 0.0]) for i in range(nk)])
    pts_true = np.stack([rng.uniform(-2.5, 2.5, nl), rng.uniform(-1.5, 1.5, nl), rng.uniform(1.5, 8.0, nl)], axis=1)
    pts_true[:, 0] += 0.6
    T_cw = np.stack([se3_inv(T) for T in poses_true])
    okf, olm, ouv = [], [], []
    for j in range(nl):
        ks = rng.choice(nk, size=min(obs_per, nk), replace=False)
        for k in ks:
            Pc = T_cw[k, :3, :3] @ pts_true[j] + T_cw[k, :3, 3]
            if Pc[2] < 0.5:
                continue
            u = rig.fx_rect * Pc[0] / Pc[2] + rig.cx
            v = rig.fx_rect * Pc[1] / Pc[2] + rig.cy
            if not (0 <= u < rig.width and 0 <= v < rig.height):
                continue
            okf.append(k)
            olm.append(j)
            ouv.append((u + rng.randn() * px_noise, v + rng.randn() * px_noise))
    okf = np.array(okf, np.int32)
    olm = np.array(olm, np.int32)
    ouv = np.array(ouv, np.float64)
    uniq, olm = np.unique(olm, return_inverse=True)
    pts_true = pts_true[uniq]
    poses0 = poses_true.copy()
    for i in range(1, nk):
        poses0[i] = poses_true[i] @ se3_exp(rng.randn(6) * pose_noise)
    pts0 = pts_true + rng.randn(*pts_true.shape) * pt_noise
    return dict(rig=rig, poses_true=poses_true, pts_true=pts_true, poses0=poses0, pts0=pts0, okf=okf,
                olm=olm.astype(np.int32), ouv=ouv)


def to_inverse_depth(pb, ext_every=0, ext_shift=None):
    """The inverse-depth form of a ``make_problem`` problem for ``_solve_np_id`` (backend/ba.py:1212): every
    landmark is hosted by the keyframe of its first observation, with its bearing (x/z, y/z, 1) and inverse depth
    1/z in that keyframe's camera frame at the initial poses and points.  With ``ext_every`` = m > 0, every m-th
    landmark's host is a constant pose outside the window (hidx = -1; the first observer's initial pose, moved by
    the twist ``ext_shift`` when given), as the reference's cached out-of-window hosts.  Returns dict(poses0, rho0,
    mbear0, hidx, hpose_ext (nl, 12), okf, olm, ouv, rig)."""
    poses0, pts0 = pb["poses0"], pb["pts0"]
    okf, olm = np.asarray(pb["okf"], np.int64), np.asarray(pb["olm"], np.int64)
    nl = len(pts0)
    first = np.full(nl, -1, np.int64)
    for k in range(len(okf) - 1, -1, -1):
        first[olm[k]] = okf[k]
    host_T = poses0[first].copy()
    hidx = first.astype(np.int32)
    if ext_every > 0:
        ext = np.arange(nl) % ext_every == ext_every - 1
        hidx[ext] = -1
        if ext_shift is not None:
            host_T[ext] = np.einsum("nij,jk->nik", host_T[ext], se3_exp(ext_shift))
    Tcw_h = np.stack([se3_inv(T) for T in host_T])
    Xh = np.einsum("nij,nj->ni", Tcw_h[:, :3, :3], pts0) + Tcw_h[:, :3, 3]
    mbear = np.stack([Xh[:, 0] / Xh[:, 2], Xh[:, 1] / Xh[:, 2], np.ones(nl)], 1)
    rho = 1.0 / Xh[:, 2]
    hext = np.zeros((nl, 12))
    ext = hidx < 0
    hext[ext] = host_T[ext, :3, :4].reshape(-1, 12)
    return dict(poses0=poses0.copy(), rho0=rho, mbear0=mbear, hidx=hidx, hpose_ext=hext, okf=pb["okf"],
                olm=pb["olm"], ouv=pb["ouv"], rig=pb["rig"], poses_true=pb["poses_true"])


# ------------------------------------------------------------------------------------ inverse-depth windows --
R_CAM_IN_BODY = np.array([[0.0, 0.0, 1.0], [-1.0, 0.0, 0.0], [0.0, -1.0, 0.0]])  # camera z forward = body x


def _rot_zyx(yaw, pitch, roll):
    cz, sz, cy_, sy, cx_, sx = (math.cos(yaw), math.sin(yaw), math.cos(pitch), math.sin(pitch), math.cos(roll),
                                math.sin(roll))
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1.0]])
    Ry = np.array([[cy_, 0, sy], [0, 1.0, 0], [-sy, 0, cy_]])
    Rx = np.array([[1.0, 0, 0], [0, cx_, -sx], [0, sx, cx_]])
    return Rz @ Ry @ Rx


def _vo_trajectory(rng, P, spacing=0.5, yaw_deg=10.0):
    """Body (= camera, T_bc = I) poses on a random smooth path: 0.5 m spacing, heading within +/- yaw_deg."""
    amp = math.radians(yaw_deg)
    om = rng.uniform(0.2, 0.6)
    ph = rng.uniform(0, 2 * math.pi)
    Rwb = np.zeros((P, 3, 3))
    pwb = np.zeros((P, 3))
    pos = np.zeros(3)
    for i in range(P):
        yaw = amp * math.sin(om * i + ph)
        pitch = math.radians(1.0) * math.sin(0.7 * om * i + 1.3 * ph)
        Rb = _rot_zyx(yaw, pitch, 0.0)
        Rwb[i] = Rb @ R_CAM_IN_BODY
        pwb[i] = pos
        pos = pos + spacing * np.array([math.cos(yaw), math.sin(yaw), 0.02 * math.sin(om * i)])
    return Rwb, pwb


def _imu_signals(rng, T, rate):
    """A smooth analytic body motion sampled at ``rate``: body angular rate w_b(t), world acceleration a_w(t)."""
    fr = rng.uniform(0.6, 1.4, 6)
    ph = rng.uniform(0, 2 * math.pi, 6)

    def w_b(t):
        return np.stack([0.10 * np.sin(0.7 * fr[0] * t + ph[0]), 0.08 * np.sin(0.9 * fr[1] * t + ph[1]),
                         0.30 * np.sin(0.5 * fr[2] * t + ph[2])], -1)

    def a_w(t):
        return np.stack([0.30 * np.sin(0.4 * fr[3] * t + ph[3]), 0.20 * np.cos(0.6 * fr[4] * t + ph[4]),
                         0.10 * np.sin(0.8 * fr[5] * t + ph[5])], -1)
    return w_b, a_w


def imu_streams(n, duration, rate=None, seed=0):
    """``n`` synthetic IMU streams of ``duration`` seconds at ``rate`` Hz (the configured IMU by default): the
    analytic body rates and accelerations of ``_imu_signals`` plus gravity along the body z axis, constant biases and
    white noise at the configured densities.  Returns [(t, gyro, acc)]; used as preintegration workloads."""
    rate = IMU_NOISE["rate_hz"] if rate is None else float(rate)
    rng = np.random.default_rng(seed)
    ns = int(round(duration * rate)) + 1
    t = np.arange(ns) / rate
    out = []
    for _ in range(n):
        w_b, a_w = _imu_signals(rng, duration, rate)
        bg = rng.normal(0, 2e-3, 3)
        ba = rng.normal(0, 2e-2, 3)
        gyro = w_b(t) + bg + rng.normal(0, IMU_NOISE["sigma_g"] * math.sqrt(rate), (ns, 3))
        acc = a_w(t) + np.array([0.0, 0.0, 9.81]) + ba + rng.normal(0, IMU_NOISE["sigma_a"] * math.sqrt(rate),
                                                                        (ns, 3))
        out.append((t.copy(), gyro, acc))
    return out


def _integrate_truth(w_b, a_w, T, rate, v0, sub=8):
    """Truth (R_wb, p, v) at the sample times by fine midpoint integration of the analytic signals."""
    n = int(round(T * rate)) + 1
    ts = np.arange(n) / rate
    R = np.eye(3)
    p = np.zeros(3)
    v = np.array(v0, float)
    Rs, ps, vs = [R.copy()], [p.copy()], [v.copy()]
    h = 1.0 / (rate * sub)
    t = 0.0
    for k in range(n - 1):
        for _ in range(sub):
            tm = t + 0.5 * h
            am = a_w(np.array(tm))
            p = p + v * h + 0.5 * am * h * h
            v = v + am * h
            R = R @ so3_exp(w_b(np.array(tm)) * h)
            t += h
        u, _s, vt = np.linalg.svd(R)
        R = u @ vt
        Rs.append(R.copy())
        ps.append(p.copy())
        vs.append(v.copy())
    return ts, np.stack(Rs), np.stack(ps), np.stack(vs)


def preintegrate(ts, gyro, acc, k0, k1, bg_lin, ba_lin, noise):
    """Midpoint preintegration over samples k0..k1 (Forster et al., T-RO 2017, discrete form) with first-order bias
    Jacobians and the 15x15 covariance in residual order [rR, rv, rp, rbg, rba] (bias blocks: random walk only)."""
    dR = np.eye(3)
    dv = np.zeros(3)
    dp = np.zeros(3)
    JRg = np.zeros((3, 3))
    Jvg = np.zeros((3, 3))
    Jva = np.zeros((3, 3))
    Jpg = np.zeros((3, 3))
    Jpa = np.zeros((3, 3))
    C = np.zeros((9, 9))
    sg2, sa2 = noise["sigma_g"] ** 2, noise["sigma_a"] ** 2
    for k in range(k0, k1):
        h = ts[k + 1] - ts[k]
        w = 0.5 * (gyro[k] + gyro[k + 1]) - bg_lin
        a = 0.5 * (acc[k] + acc[k + 1]) - ba_lin
        dRk = so3_exp(w * h)
        Jr = so3_jr(w * h)
        Rm = dR @ so3_exp(0.5 * w * h)
        Ax = Rm @ hat(a)
        A = np.eye(9)
        A[0:3, 0:3] = dRk.T
        A[3:6, 0:3] = -Ax * h
        A[6:9, 0:3] = -0.5 * Ax * h * h
        A[6:9, 3:6] = np.eye(3) * h
        Bg = np.zeros((9, 3))
        Bg[0:3] = Jr * h
        Ba = np.zeros((9, 3))
        Ba[3:6] = Rm * h
        Ba[6:9] = 0.5 * Rm * h * h
        C = A @ C @ A.T + Bg @ Bg.T * (sg2 / h) + Ba @ Ba.T * (sa2 / h)
        Jpg = Jpg + Jvg * h - 0.5 * Ax @ JRg * h * h
        Jpa = Jpa + Jva * h - 0.5 * Rm * h * h
        Jvg = Jvg - Ax @ JRg * h
        Jva = Jva - Rm * h
        JRg = dRk.T @ JRg - Jr * h
        dp = dp + dv * h + 0.5 * (Rm @ a) * h * h
        dv = dv + (Rm @ a) * h
        dR = dR @ dRk
    Tt = ts[k1] - ts[k0]
    cov = np.zeros((15, 15))
    cov[:9, :9] = 0.5 * (C + C.T)
    cov[9:12, 9:12] = np.eye(3) * noise["rw_g"] ** 2 * Tt
    cov[12:15, 12:15] = np.eye(3) * noise["rw_a"] ** 2 * Tt
    J = np.zeros((9, 6))
    J[0:3, 0:3] = JRg
    J[3:6, 0:3] = Jvg
    J[3:6, 3:6] = Jva
    J[6:9, 0:3] = Jpg
    J[6:9, 3:6] = Jpa
    u, _s, vt = np.linalg.svd(dR)
    return Preint(t0=ts[k0], t1=ts[k1], dt=Tt, dR=u @ vt, dv=dv, dp=dp, J=J, bg_lin=np.array(bg_lin, float),
                  ba_lin=np.array(ba_lin, float), cov=cov)


def sqrt_information(cov, kappa=1.0):
    """W = L^-1 with L L^T = kappa C (vio/imu_factor.py:707-746 with no BIASCOV term)."""
    L = np.linalg.cholesky(kappa * cov)
    return np.linalg.solve(L, np.eye(15))


def huber_k_dof(dof, k=HUBER_K):
    """sqrt(chi2_dof^-1(P(|z| <= k))) (vio/imu_factor.py:1096-1101)."""
    from scipy.special import erf, gammaincinv
    q = float(erf(k / math.sqrt(2.0)))
    return float(math.sqrt(2.0 * gammaincinv(0.5 * float(dof), q)))


def _project(Rwc, pwc, X):
    """Camera-frame coordinates of world points X (L,3) in every slot: (P, L, 3)."""
    return np.einsum("pji,plj->pli", Rwc, X[None, :, :] - pwc[:, None, :])


def make_window(P=10, L=2000, views=(2, 10), seed=0, vi=False, stereo="prior", outlier_frac=0.05,
                px_sigma=FEATURE_SIGMA_PX, pose_sigma_t=0.02, pose_sigma_r_deg=0.5, depth=(1.0, 20.0),
                spacing=0.5, posefree=False, gauge=True, noise_scale=1.0):
    """A seeded inverse-depth stereo window (module docstring).  ``views`` = (min, max) views per landmark including
    the host's.  ``posefree`` moves landmarks whose only observation is the host-left one into ``Problem.pf``.
    ``noise_scale`` multiplies the pixel and stereo-depth noise drawn (0: every measurement is the exact projection
    of the true world point, whitened as if its sigma were ``px_sigma``); the random stream is the same."""
    rng = np.random.default_rng(seed)
    fx, fy, cx, cy, Wd, Hd, B = (CAM[k] for k in ("fx", "fy", "cx", "cy", "W", "H", "B"))
    if vi:
        dtk = spacing / 1.0
        rate = IMU_NOISE["rate_hz"]
        Ttot = dtk * (P - 1)
        w_b, a_w = _imu_signals(rng, Ttot, rate)
        ts, Rs, ps, vs = _integrate_truth(w_b, a_w, Ttot, rate, v0=[1.0, 0.0, 0.0])
        step = int(round(dtk * rate))
        kidx = np.arange(P) * step
        Rwb_t, pwb_t, vel_t = Rs[kidx], ps[kidx], vs[kidx]
        Rbc = R_CAM_IN_BODY @ so3_exp(np.radians([0.3, -0.2, 0.4]))
        pbc = np.array([0.06, -0.02, 0.01])
    else:
        Rwb_t, pwb_t = _vo_trajectory(rng, P, spacing)
        Rbc = np.eye(3)
        pbc = np.zeros(3)
    Rwc = np.einsum("pij,jk->pik", Rwb_t, Rbc)
    pwc = pwb_t + np.einsum("pij,j->pi", Rwb_t, pbc)

    # landmarks: host, host pixel, depth
    host = rng.integers(0, P, L)
    u0 = rng.uniform(0, Wd, L)
    v0 = rng.uniform(0, Hd, L)
    d = rng.uniform(depth[0], depth[1], L)
    m = np.stack([(u0 - cx) / fx, (v0 - cy) / fy, np.ones(L)], 1)
    Xw = np.einsum("lij,lj->li", Rwc[host], m * d[:, None]) + pwc[host]
    q = _project(Rwc, pwc, Xw)                                  # (P, L, 3)
    z = q[:, :, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        uu = fx * q[:, :, 0] / z + cx
        vv = fy * q[:, :, 1] / z + cy
    vis = (z > 0.3) & (uu >= 0) & (uu < Wd) & (vv >= 0) & (vv < Hd)
    if stereo == "ur":
        ur = fx * (q[:, :, 0] - B) / z + cx
        vis &= (ur >= 0) & (ur < Wd)
    vis[host, np.arange(L)] = True
    nv = rng.integers(views[0], min(views[1], P) + 1, L)
    slots = np.arange(P)[:, None]
    score = np.abs(slots - host[None, :]) + rng.uniform(0, 0.5, (P, L))
    score = np.where(vis, score, np.inf)
    score[host, np.arange(L)] = -1.0                            # the host view first
    order = np.argsort(score, axis=0)                           # (P, L)
    rank = np.empty_like(order)
    np.put_along_axis(rank, order, np.arange(P)[:, None].repeat(L, 1), axis=0)
    take = (rank < nv[None, :]) & np.isfinite(score) | (score < 0)
    # observations, CSR by landmark, slots ascending inside a landmark
    lm_idx, sl_idx = np.nonzero(take.T)
    cnt = np.bincount(lm_idx, minlength=L)
    obs_l, obs_t = lm_idx, sl_idx
    n = len(obs_l)
    zl = np.stack([uu[obs_t, obs_l], vv[obs_t, obs_l]], 1) + noise_scale * rng.normal(0, px_sigma, (n, 2))
    is_host = obs_t == host[obs_l]
    out = (rng.uniform(0, 1, n) < outlier_frac) & ~is_host
    zl[out] = np.stack([rng.uniform(0, Wd, out.sum()), rng.uniform(0, Hd, out.sum())], 1)
    rho_t = 1.0 / d
    if stereo == "ur":
        zr = (fx * (q[obs_t, obs_l, 0] - B) / z[obs_t, obs_l] + cx) + noise_scale * rng.normal(0, px_sigma, n)
        zr[out] = rng.uniform(0, Wd, out.sum())
        # interleave: left then right per view
        ob_pose = np.repeat(obs_t, 2)
        ob_cam = np.tile([0, 1], n).astype(np.uint8)
        ob_z = np.zeros((2 * n, 2))
        ob_z[0::2] = zl
        ob_z[1::2, 0] = zr
        cnt = cnt * 2
        rho_isig = np.zeros(L)
        rho_meas = rho_t.copy()
        rho0 = np.clip(rho_t * np.exp(rng.normal(0, 0.05, L)), 1e-3, None)
    else:
        ob_pose = obs_t
        ob_cam = np.zeros(n, np.uint8)
        ob_z = zl
        sig_rho = DISPARITY_NOISE_PX / (fx * B)
        rho_meas = np.clip(rho_t + noise_scale * rng.normal(0, sig_rho, L), 1e-3, None)
        rho_isig = np.full(L, 1.0 / sig_rho)
        rho0 = rho_meas.copy()
    ob_start = np.zeros(L + 1, np.int64)
    ob_start[1:] = np.cumsum(cnt)
    M = len(ob_pose)
    # initial landmark state from the (noisy) host observation and the stereo measurement
    hz = zl[is_host]
    hl = obs_l[is_host]
    a0 = np.empty(L)
    b0 = np.empty(L)
    a0[hl] = (hz[:, 0] - cx) / fx
    b0[hl] = (hz[:, 1] - cy) / fy
    lmp = np.stack([a0, b0, rho0], 1)
    lmp_true = np.stack([m[:, 0], m[:, 1], rho_t], 1)

    # initial poses: perturbed, slot 0 exact
    Rwb0 = Rwb_t.copy()
    pwb0 = pwb_t.copy()
    for s in range(1, P):
        Rwb0[s] = Rwb_t[s] @ so3_exp(rng.normal(0, math.radians(pose_sigma_r_deg), 3))
        pwb0[s] = pwb_t[s] + rng.normal(0, pose_sigma_t, 3)
    pb = Problem(Rwb=Rwb0, pwb=pwb0, Rbc=Rbc.copy(), pbc=pbc.copy(), lmp=lmp, lm_host=host.astype(np.int64),
                 lm_mask=np.ones(L, np.bool_), rho_meas=rho_meas, rho_isig=rho_isig, ob_start=ob_start,
                 ob_pose=ob_pose.astype(np.int64), ob_cam=ob_cam.astype(np.int64), ob_z=np.ascontiguousarray(ob_z),
                 ob_isig=np.full(M, 1.0 / px_sigma), ob_valid=np.ones(M, np.bool_), fx=fx, fy=fy, cx=cx, cy=cy,
                 B=B, huber_k=HUBER_K)
    if gauge:
        pb.prior = gauge_prior(0, Rwb0[0], pwb0[0])
        pb.prior.bind({0: 0})
    truth = dict(Rwb=Rwb_t, pwb=pwb_t, lmp=lmp_true)
    if vi:
        noise = IMU_NOISE
        bg_t = rng.normal(0, 2e-3, 3)
        ba_t = rng.normal(0, 2e-2, 3)
        n_s = len(ts)
        gyro = np.stack([w_b(t) for t in ts]) + bg_t + rng.normal(0, noise["sigma_g"] * math.sqrt(rate), (n_s, 3))
        acc_true = np.einsum("nji,nj->ni", Rs, np.stack([a_w(t) for t in ts]) - G_W)
        acc = acc_true + ba_t + rng.normal(0, noise["sigma_a"] * math.sqrt(rate), (n_s, 3))
        imu = []
        for i in range(P - 1):
            pre = preintegrate(ts, gyro, acc, kidx[i], kidx[i + 1], np.zeros(3), np.zeros(3), noise)
            imu.append((i, i + 1, pre, sqrt_information(pre.cov)))
        pb.dims = np.full(P, 15, np.int64)
        pb.vel = vel_t + rng.normal(0, 0.05, (P, 3))
        pb.bgs = np.zeros((P, 3))
        pb.bas = np.zeros((P, 3))
        pb.imu = imu
        pb.g_w = G_W.copy()
        truth.update(vel=vel_t, bg=bg_t, ba=ba_t)
    if posefree:
        single = (np.diff(ob_start) == 1) & (ob_cam[ob_start[:-1].clip(max=M - 1)] == 0)
        single &= ob_pose[ob_start[:-1].clip(max=M - 1)] == host
        idx = np.nonzero(single)[0]
        if len(idx):
            pb.pf = PoseFree(host[idx], ob_z[ob_start[idx]], 1.0 / px_sigma, rho_meas[idx], rho_isig[idx], lmp[idx])
            keep = np.ones(L, bool)
            keep[idx] = False
            _drop_landmarks(pb, keep)
    return pb, truth


def _drop_landmarks(pb, keep):
    """Keep only landmarks ``keep`` (bool mask) and their observations."""
    cnt = np.diff(pb.ob_start)
    okeep = np.repeat(keep, cnt)
    for k in ("lmp", "lm_host", "lm_mask", "rho_meas", "rho_isig"):
        setattr(pb, k, np.ascontiguousarray(getattr(pb, k)[keep]))
    for k in ("ob_pose", "ob_cam", "ob_z", "ob_isig", "ob_valid"):
        setattr(pb, k, np.ascontiguousarray(getattr(pb, k)[okeep]))
    st = np.zeros(int(keep.sum()) + 1, np.int64)
    st[1:] = np.cumsum(cnt[keep])
    pb.ob_start = st


CONFIGS = {
    0: dict(P=10, L=2000, views=(2, 10), vi=False),
    1: dict(P=20, L=20000, views=(2, 10), vi=True),
    3: dict(P=15, L=5000, views=(2, 10), vi=False),
    4: dict(P=60, L=200000, views=(2, 13), vi=True),
}


def make_config(cfg, seed=0, **kw):
    """The single-window configurations 0, 1, 3 (one window) and 4; seeds follow [20261203, cfg, window]."""
    args = dict(CONFIGS[cfg])
    args.update(kw)
    return make_window(seed=[20261203, cfg, seed], **args)


def make_batch(nwin=256, seed=0, P_range=(5, 30), L_range=(500, 30000), views_max=10, scale=1.0):
    """The batched configuration: ``nwin`` independent VO windows, P ~ U{5..30}, L log-uniform in L_range
    (times ``scale``), views 2..min(10, P)."""
    rng = np.random.default_rng([20261203, 2, seed])
    out = []
    for w in range(nwin):
        P = int(rng.integers(P_range[0], P_range[1] + 1))
        L = int(round(math.exp(rng.uniform(math.log(L_range[0]), math.log(L_range[1]))) * scale))
        pb, _t = make_window(P=P, L=max(L, 4), views=(2, min(views_max, P)), seed=[20261203, 2, seed, w])
        out.append(pb)
    return out


def make_gba(nk=60, nl=3000, track=(2, 10), seed=0, reach=15, px_sigma=0.5, disp_sigma=0.5, outlier_frac=0.02,
             stereo_frac=0.8, prior_valid_frac=0.9, pose_sigma_t=0.05, pose_sigma_r_deg=0.5, depth=(1.0, 20.0),
             spacing=0.5, noise_scale=1.0):
    """A seeded whole-trajectory problem in the argument form of backend/global_ba.py:694 ``global_ba_id``.

    ``nk`` keyframes on the VO trajectory model (camera = body), ``nl`` landmarks, each hosted at a random keyframe
    at a uniform pixel and depth U(depth), observed by its host and by up to track[1] - 1 keyframes within
    ``reach`` of it that see it (track length U{track}).  Pixels carry N(0, px_sigma); ``outlier_frac`` of the
    non-host rows are shifted by U(10, 40) px.  ``stereo_frac`` of the rows carry a disparity fB / z + N(0,
    disp_sigma) (the others 0: none measured).  The landmark prior holds the host's measured bearing and inverse
    depth (whitening fx / px_sigma and fB / disp_sigma; ``prior_valid_frac`` of them valid).  The state starts at
    the measured bearings and inverse depths and at poses perturbed by a random walk (pose_sigma_t m, pose_sigma_r
    deg per step); keyframe 0 is at the truth.  ``noise_scale`` scales every measurement noise (0: exact).

    -> dict(kf_ids, poses (nk,4,4) T_wc, rho, mbear, host_idx, okf, olm, ouv, odisp, rig, prior, truth)."""
    import types
    rng = np.random.default_rng(seed)
    fx, cx, cy, Wd, Hd, B = (CAM[k] for k in ("fx", "cx", "cy", "W", "H", "B"))
    fB = fx * B
    Rwc, pwc = _vo_trajectory(rng, nk, spacing)
    host = rng.integers(0, nk, nl)
    u0, v0 = rng.uniform(0, Wd, nl), rng.uniform(0, Hd, nl)
    dep = rng.uniform(depth[0], depth[1], nl)
    m_true = np.stack([(u0 - cx) / fx, (v0 - cy) / fx, np.ones(nl)], 1)
    Xw = np.einsum("lij,lj->li", Rwc[host], m_true * dep[:, None]) + pwc[host]
    offs = np.arange(-reach, reach + 1)
    cand = host[:, None] + offs[None, :]                                   # (nl, 2 reach + 1)
    inside = (cand >= 0) & (cand < nk)
    c = np.clip(cand, 0, nk - 1)
    q = np.einsum("lkji,lkj->lki", Rwc[c], Xw[:, None, :] - pwc[c])          # camera-frame points
    z = q[..., 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        uu, vv = fx * q[..., 0] / z + cx, fx * q[..., 1] / z + cy
    vis = inside & (z > 0.3) & (uu >= 0) & (uu < Wd) & (vv >= 0) & (vv < Hd)
    vis[:, reach] = True                                                    # the host row
    ntr = rng.integers(track[0], track[1] + 1, nl)
    score = np.where(vis, np.abs(offs)[None, :] + rng.uniform(0, 0.5, vis.shape), np.inf)
    score[:, reach] = -1.0
    rank = np.argsort(np.argsort(score, axis=1), axis=1)
    take = (rank < ntr[:, None]) & np.isfinite(score)
    lm_i, k_i = np.nonzero(take)
    okf = cand[lm_i, k_i]
    n = len(okf)
    ouv = np.stack([uu[lm_i, k_i], vv[lm_i, k_i]], 1) + noise_scale * rng.normal(0, px_sigma, (n, 2))
    is_host = okf == host[lm_i]
    bad = (rng.uniform(0, 1, n) < outlier_frac) & ~is_host
    ang = rng.uniform(0, 2 * math.pi, n)
    ouv[bad] += rng.uniform(10, 40, n)[bad, None] * np.stack([np.cos(ang), np.sin(ang)], 1)[bad]
    zt = z[lm_i, k_i]
    odisp = fB / zt + noise_scale * rng.normal(0, disp_sigma, n)
    odisp[rng.uniform(0, 1, n) >= stereo_frac] = 0.0
    odisp = np.maximum(odisp, 0.0)
    # the host row's measurement gives the prior and the starting landmark state
    hrow = np.flatnonzero(is_host)
    m0 = np.zeros((nl, 2))
    m0[lm_i[hrow]] = np.stack([(ouv[hrow, 0] - cx) / fx, (ouv[hrow, 1] - cy) / fx], 1)
    rho0 = 1.0 / dep + noise_scale * rng.normal(0, disp_sigma / fB, nl)
    rho0 = np.maximum(rho0, 1e-3)
    valid = (rng.uniform(0, 1, nl) < prior_valid_frac).astype(np.float64)
    prior = (m0, rho0, np.full(nl, fx / px_sigma), np.full(nl, fB / disp_sigma), valid)
    mbear = np.concatenate([m0, np.ones((nl, 1))], 1)
    # poses: keyframe 0 exact, then a random walk of errors
    dth = np.cumsum(rng.normal(0, math.radians(pose_sigma_r_deg), (nk, 3)), 0)
    dpt = np.cumsum(rng.normal(0, pose_sigma_t, (nk, 3)), 0)
    dth[0] = 0.0
    dpt[0] = 0.0
    poses = np.zeros((nk, 4, 4))
    for k in range(nk):
        poses[k, :3, :3] = so3_exp(dth[k]) @ Rwc[k]
        poses[k, :3, 3] = pwc[k] + dpt[k]
        poses[k, 3, 3] = 1.0
    truth = np.zeros((nk, 4, 4))
    truth[:, :3, :3], truth[:, :3, 3], truth[:, 3, 3] = Rwc, pwc, 1.0
    rig = types.SimpleNamespace(fx_rect=fx, cx=cx, cy=cy, fB=fB, rho_sigma=0.5 / fB)
    return {"kf_ids": np.arange(nk), "poses": poses, "rho": rho0.copy(), "mbear": mbear, "host_idx": host,
            "okf": okf.astype(np.int64), "olm": lm_i.astype(np.int64), "ouv": ouv, "odisp": odisp, "rig": rig,
            "prior": prior, "truth": {"poses": truth, "rho": 1.0 / dep, "mbear": m_true, "Xw": Xw,
                                      "outlier": bad}}
