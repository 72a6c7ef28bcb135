"""Pins the oracle's IMU preintegration (oracle/xba_oracle.c o_preint, restating vio/imu_factor.py:392-704) and its
whitening (o_sqrt_info, :707-746).  No reference outputs exist (the reference needs Numba), so every check here is
one the mathematics fixes:

- the mean against an independent numpy/scipy evaluation of the documented midpoint rule (imu_factor.py:46-62) on
  irregular streams, and closed forms (constant rates: dR = Exp(w T); w = 0: dv = a T, dp = a T^2 / 2);
- the bias Jacobians against central differences of the oracle's own mean;
- the covariance in closed form where the scheme is the exact discretisation of the continuous error dynamics
  (w = 0: accelerometer white noise and walk, gyro white noise through a constant [a]x, gyro walk with a = 0), the
  single-sub-interval covariance against the Van Loan integral of the frozen linear SDE at any rates, and the
  composition identity at a sample stamp (mean, bias Jacobians and the full 15x15 covariance of [t0, t1] from those
  of [t0, tm] and [tm, t1]);
- the interval rules: sample-support scale on regular streams, gaps and holds, the wholly held interval, causality
  (values past the horizon never read), the refusal codes;
- the whitening: W lower triangular, W^T W = (kappa C + BIASCOV)^-1, refusal when not positive definite.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O

TN = 0.005  # 200 Hz


def hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], float)


def exp3(v):
    return expm(hat(v))


def log3(R):
    w = np.zeros(3)
    O.lib().o_log_quat(O._p(np.ascontiguousarray(R, np.float64), O.f64p), O._p(w, O.f64p))
    return w


def stream(seed, n=140, jitter=0.0, drop=()):
    rng = np.random.default_rng(seed)
    t = np.arange(n) * TN
    if jitter:
        t = t + rng.uniform(-jitter, jitter, n) * TN
        t[0] = 0.0
    keep = np.ones(n, bool)
    keep[list(drop)] = False
    t = t[keep]
    f = rng.uniform(0.5, 3.0, 6)
    ph = rng.uniform(0, 6.28, 6)
    gyr = np.stack([0.6 * np.sin(f[i] * 6.28 * t + ph[i]) for i in range(3)], 1) + rng.normal(0, 0.02, (len(t), 3))
    acc = np.stack([1.5 * np.sin(f[3 + i] * 6.28 * t + ph[3 + i]) for i in range(3)], 1) + [0.3, -0.2, 9.81]
    return t, gyr, acc + rng.normal(0, 0.05, (len(t), 3))


def pre(t, g, a, t0, t1, horizon=None, bg=(0, 0, 0), ba=(0, 0, 0), q=(0.0, 0.0, 0.0, 0.0), tnom=TN):
    return O.preintegrate(t, g, a, t0, t1, horizon, np.asarray(bg, float), np.asarray(ba, float), *q, tnom)


def numpy_mean(t, g, a, t0, t1, horizon=None, bg=np.zeros(3), ba=np.zeros(3)):
    """The midpoint rule as documented (imu_factor.py:46-62): edges t0, released samples inside, min(t1, t_end),
    t1; rates = np.interp at each sub-interval's midpoint (the last released sample held past t_end); rotation
    advanced by Exp(w h), acceleration rotated by the midpoint attitude."""
    rel = t <= (np.inf if horizon is None else horizon)
    tr, gr, ar = t[rel], g[rel], a[rel]
    te = tr[-1]
    edges = [t0] + [s for s in tr if t0 < s < min(t1, te)] + [min(t1, te)] + ([t1] if t1 > te else [])
    dR, dv, dp = np.eye(3), np.zeros(3), np.zeros(3)
    for e0, e1 in zip(edges[:-1], edges[1:]):
        h = e1 - e0
        if h <= 0:
            continue
        m = 0.5 * (e0 + e1)
        w = np.array([np.interp(m, tr, gr[:, c]) for c in range(3)]) - bg
        acc = np.array([np.interp(m, tr, ar[:, c]) for c in range(3)]) - ba
        am = dR @ exp3(0.5 * w * h) @ acc
        dp = dp + dv * h + 0.5 * am * h * h
        dv = dv + am * h
        dR = dR @ exp3(w * h)
    return dR, dv, dp


def test_mean_matches_numpy_midpoint_rule():
    for seed, jit, drop, (t0, t1) in [(1, 0.0, (), (0.0123, 0.4567)), (2, 0.3, (), (0.003, 0.61)),
                                      (3, 0.2, (40, 41, 42), (0.1, 0.5)), (4, 0.0, (), (0.1, 0.8))]:
        t, g, a = stream(seed, jitter=jit, drop=drop)
        r = pre(t, g, a, t0, t1, bg=(0.01, -0.02, 0.005), ba=(0.1, 0.0, -0.05))
        assert r["info"][O.PI_STATUS] == 0
        dR, dv, dp = numpy_mean(t, g, a, t0, t1, bg=np.array([0.01, -0.02, 0.005]), ba=np.array([0.1, 0.0, -0.05]))
        assert np.abs(r["dR"] - dR).max() < 1e-12
        assert np.abs(r["dv"] - dv).max() < 1e-11 * max(1, np.abs(dv).max())
        assert np.abs(r["dp"] - dp).max() < 1e-11 * max(1, np.abs(dp).max())
        assert np.abs(r["dR"].T @ r["dR"] - np.eye(3)).max() < 1e-15


def test_mean_closed_forms():
    t = np.arange(120) * TN
    w = np.array([0.4, -0.3, 0.2])
    g = np.tile(w, (120, 1))
    a0 = np.array([0.5, -1.0, 9.8])
    r = pre(t, g, np.tile(a0, (120, 1)), 0.0021, 0.5)
    assert np.abs(r["dR"] - exp3(w * (0.5 - 0.0021))).max() < 1e-14
    r = pre(t, np.zeros((120, 3)), np.tile(a0, (120, 1)), 0.0021, 0.5)
    T = 0.5 - 0.0021
    assert np.abs(r["dv"] - a0 * T).max() < 1e-13 and np.abs(r["dp"] - a0 * T * T / 2).max() < 1e-13


def test_bias_jacobians_match_central_differences():
    t, g, a = stream(5, jitter=0.25)
    bg0, ba0 = np.array([0.02, -0.01, 0.03]), np.array([0.05, 0.1, -0.02])
    r = pre(t, g, a, 0.0071, 0.52, bg=bg0, ba=ba0)
    eps = 1e-6
    for c in range(6):
        d = np.zeros(6)
        d[c] = eps
        rp = pre(t, g, a, 0.0071, 0.52, bg=bg0 + d[:3], ba=ba0 + d[3:])
        rm = pre(t, g, a, 0.0071, 0.52, bg=bg0 - d[:3], ba=ba0 - d[3:])
        dphi = (log3(r["dR"].T @ rp["dR"]) - log3(r["dR"].T @ rm["dR"])) / (2 * eps)
        dv = (rp["dv"] - rm["dv"]) / (2 * eps)
        dp = (rp["dp"] - rm["dp"]) / (2 * eps)
        fd = np.concatenate([dphi, dv, dp])
        assert np.abs(r["J"][:, c] - fd).max() < 2e-7 * max(1.0, np.abs(fd).max()), c


def _internal(Cm):
    """The internal covariance (errors computed minus true): the residual order flips the preint / bias block."""
    P = Cm.copy()
    P[:9, 9:] *= -1
    P[9:, :9] *= -1
    return P


def test_covariance_closed_forms_where_the_scheme_is_exact():
    t = np.arange(120) * TN
    T = 100 * TN
    a0 = np.array([0.7, -0.4, 9.8])
    Ax = hat(a0)
    I = np.eye(3)
    # accelerometer white noise and walk, w = 0 (any a)
    sa2, swa2 = 2e-3, 3e-4
    P = _internal(pre(t, np.zeros((120, 3)), np.tile(a0, (120, 1)), 0.0, T, q=(0, sa2, 0, swa2))["C"])
    assert np.allclose(P[3:6, 3:6], (sa2 * T + swa2 * T ** 3 / 3) * I, rtol=1e-12, atol=0)
    assert np.allclose(P[6:9, 3:6], (sa2 * T ** 2 / 2 + swa2 * T ** 4 / 8) * I, rtol=1e-12, atol=0)
    assert np.allclose(P[6:9, 6:9], (sa2 * T ** 3 / 3 + swa2 * T ** 5 / 20) * I, rtol=1e-12, atol=0)
    assert np.allclose(P[3:6, 12:15], swa2 * T ** 2 / 2 * I, rtol=1e-12, atol=0)
    assert np.allclose(P[6:9, 12:15], swa2 * T ** 3 / 6 * I, rtol=1e-12, atol=0)
    assert np.allclose(P[12:15, 12:15], swa2 * T * I, rtol=1e-12, atol=0)
    assert np.abs(P[0:3]).max() == 0 and np.abs(P[9:12]).max() == 0
    # gyro white noise through a constant [a]x, w = 0: phi = int n, v = -[a]x int phi, p = int v
    sg2 = 1e-3
    P = _internal(pre(t, np.zeros((120, 3)), np.tile(a0, (120, 1)), 0.0, T, q=(sg2, 0, 0, 0))["C"])
    K = Ax @ Ax.T
    assert np.allclose(P[0:3, 0:3], sg2 * T * I, rtol=1e-12, atol=0)
    assert np.allclose(P[3:6, 0:3], -Ax * sg2 * T ** 2 / 2, rtol=1e-11, atol=1e-20)
    assert np.allclose(P[6:9, 0:3], -Ax * sg2 * T ** 3 / 6, rtol=1e-11, atol=1e-20)
    assert np.allclose(P[3:6, 3:6], K * sg2 * T ** 3 / 3, rtol=1e-11, atol=1e-20)
    assert np.allclose(P[6:9, 3:6], K * sg2 * T ** 4 / 8, rtol=1e-11, atol=1e-20)
    assert np.allclose(P[6:9, 6:9], K * sg2 * T ** 5 / 20, rtol=1e-11, atol=1e-20)
    # gyro walk with a = 0: phi = int eps_g
    swg2 = 5e-4
    P = _internal(pre(t, np.zeros((120, 3)), np.zeros((120, 3)), 0.0, T, q=(0, 0, swg2, 0))["C"])
    assert np.allclose(P[0:3, 0:3], swg2 * T ** 3 / 3 * I, rtol=1e-12, atol=0)
    assert np.allclose(P[0:3, 9:12], swg2 * T ** 2 / 2 * I, rtol=1e-12, atol=0)
    assert np.allclose(P[9:12, 9:12], swg2 * T * I, rtol=1e-12, atol=0)


def test_single_subinterval_covariance_is_the_van_loan_integral():
    """One sub-interval from t0: P = Q, the exact covariance of the frozen error SDE phi' = eps_g + n_g,
    v' = -M phi + Rm (eps_a + n_a), p' = v, eps' = n_w (M = Rm [a]x, Rm the midpoint attitude)."""
    t = np.array([0.0, 0.02])
    g = np.array([[0.8, -0.5, 0.3], [1.2, 0.4, -0.6]])
    a = np.array([[0.5, 0.2, 9.7], [-0.3, 0.9, 9.9]])
    sg2, sa2, swg2, swa2 = 1e-3, 2e-3, 3e-2, 4e-2
    t0, t1 = 0.004, 0.013
    r = pre(t, g, a, t0, t1, q=(sg2, sa2, swg2, swa2), tnom=1.0)  # tnom large: scale 1
    h, m = t1 - t0, 0.5 * (t0 + t1)
    w = g[0] + (g[1] - g[0]) * m / 0.02
    acc = a[0] + (a[1] - a[0]) * m / 0.02
    Rm = exp3(0.5 * w * h)
    Mx = Rm @ hat(acc)
    A = np.zeros((15, 15))
    A[0:3, 9:12] = np.eye(3)
    A[3:6, 0:3] = -Mx
    A[3:6, 12:15] = Rm
    A[6:9, 3:6] = np.eye(3)
    G = np.zeros((15, 12))
    G[0:3, 0:3] = np.eye(3)
    G[3:6, 3:6] = Rm
    G[9:12, 6:9] = np.eye(3)
    G[12:15, 9:12] = np.eye(3)
    Qc = np.diag([sg2] * 3 + [sa2] * 3 + [swg2] * 3 + [swa2] * 3)
    VL = np.zeros((30, 30))
    VL[:15, :15] = -A
    VL[:15, 15:] = G @ Qc @ G.T
    VL[15:, 15:] = A.T
    E = expm(VL * h)
    Q = E[15:, 15:].T @ E[:15, 15:]
    P = _internal(r["C"])
    assert np.abs(P - Q).max() < 1e-12 * np.abs(Q).max()


def test_composition_at_a_sample_stamp():
    t, g, a = stream(7, jitter=0.2)
    q = (1e-4, 2e-3, 1e-5, 3e-4)
    bg, ba = (0.01, 0.02, -0.01), (0.05, -0.03, 0.02)
    t0, tm, t1 = 0.0123, t[45], 0.4417
    r = pre(t, g, a, t0, t1, bg=bg, ba=ba, q=q)
    r1 = pre(t, g, a, t0, tm, bg=bg, ba=ba, q=q)
    r2 = pre(t, g, a, tm, t1, bg=bg, ba=ba, q=q)
    R1, R2, v2, p2, T2 = r1["dR"], r2["dR"], r2["dv"], r2["dp"], t1 - tm
    assert np.abs(r["dR"] - R1 @ R2).max() < 1e-14
    assert np.abs(r["dv"] - (r1["dv"] + R1 @ v2)).max() < 1e-13
    assert np.abs(r["dp"] - (r1["dp"] + r1["dv"] * T2 + R1 @ p2)).max() < 1e-13
    J1, J2 = r1["J"], r2["J"]
    Jc = np.zeros((9, 6))
    Jc[0:3, 0:3] = R2.T @ J1[0:3, 0:3] + J2[0:3, 0:3]
    Jc[3:6, 0:3] = J1[3:6, 0:3] - R1 @ hat(v2) @ J1[0:3, 0:3] + R1 @ J2[3:6, 0:3]
    Jc[3:6, 3:6] = J1[3:6, 3:6] + R1 @ J2[3:6, 3:6]
    Jc[6:9, 0:3] = J1[6:9, 0:3] + T2 * J1[3:6, 0:3] - R1 @ hat(p2) @ J1[0:3, 0:3] + R1 @ J2[6:9, 0:3]
    Jc[6:9, 3:6] = J1[6:9, 3:6] + T2 * J1[3:6, 3:6] + R1 @ J2[6:9, 3:6]
    assert np.abs(r["J"] - Jc).max() < 1e-12 * np.abs(Jc).max()
    Phi = np.eye(15)
    Phi[0:3, 0:3] = R2.T
    Phi[0:3, 9:12] = -J2[0:3, 0:3]
    Phi[3:6, 0:3] = -R1 @ hat(v2)
    Phi[3:6, 9:12] = -R1 @ J2[3:6, 0:3]
    Phi[3:6, 12:15] = -R1 @ J2[3:6, 3:6]
    Phi[6:9, 0:3] = -R1 @ hat(p2)
    Phi[6:9, 3:6] = T2 * np.eye(3)
    Phi[6:9, 9:12] = -R1 @ J2[6:9, 0:3]
    Phi[6:9, 12:15] = -R1 @ J2[6:9, 3:6]
    D = np.eye(15)
    D[3:6, 3:6] = R1
    D[6:9, 6:9] = R1
    Pc = Phi @ _internal(r1["C"]) @ Phi.T + D @ _internal(r2["C"]) @ D.T
    P = _internal(r["C"])
    assert np.abs(P - Pc).max() < 1e-11 * np.abs(P).max()


def test_sample_support_gaps_and_holds():
    t, g, a = stream(8)
    q = (1e-4, 2e-3, 1e-5, 3e-4)
    r = pre(t, g, a, 0.1, 0.4, q=q)  # regular stream, inside
    assert abs(r["info"][O.PI_SMAX] - 1.0) < 1e-12 and r["info"][O.PI_EXTRAP] == 0.0
    assert r["info"][O.PI_NSUB] == 60 and r["info"][O.PI_NINNER] == 59
    tg, gg, ag = stream(8, drop=(50,))  # a gap of two periods: a full bracket scaled by L / T_nom = 2
    r = pre(tg, gg, ag, 0.1, 0.4, q=q)
    assert abs(r["info"][O.PI_SMAX] - 2.0) < 1e-9 and abs(r["info"][O.PI_GAPMAX] - 2 * TN) < 1e-12
    te = t[-1]
    r = pre(t, g, a, te - 0.05, te + 0.02, q=q)  # runs past the last sample: held, nu = 2
    assert abs(r["info"][O.PI_EXTRAP] - 0.02) < 1e-12 and abs(r["info"][O.PI_SMAX] - 2 * 0.02 / TN) < 1e-9
    r = pre(t, g, a, te + 0.01, te + 0.04, q=q)  # wholly held: dR = Exp(g_last T), nu = 1
    assert np.abs(r["dR"] - exp3(g[-1] * 0.03)).max() < 1e-14
    assert abs(r["info"][O.PI_SMAX] - 0.03 / TN) < 1e-9 and r["info"][O.PI_NINNER] == 0


def test_causality_and_refusals():
    t, g, a = stream(9)
    hz = t[80] + 1e-4
    q = (1e-4, 2e-3, 1e-5, 3e-4)
    r = pre(t, g, a, 0.1, 0.45, horizon=hz, q=q)
    g2, a2 = g.copy(), a.copy()
    g2[81:] = np.nan
    a2[81:] = 1e30
    r2 = pre(t, g2, a2, 0.1, 0.45, horizon=hz, q=q)
    for k in ("dR", "dv", "dp", "J", "C", "info"):
        np.testing.assert_array_equal(r[k], r2[k])
    assert r["info"][O.PI_EXTRAP] > 0
    assert pre(t, g, a, 0.2, 0.2)["info"][O.PI_STATUS] == 1
    assert pre(t, g, a, 0.1, 0.2, horizon=-1.0)["info"][O.PI_STATUS] == 2
    assert pre(t + 0.01, g, a, 0.005, 0.2)["info"][O.PI_STATUS] == 3
    g3 = g.copy()
    g3[30, 1] = np.nan
    assert pre(t, g3, a, 0.1, 0.2)["info"][O.PI_STATUS] == 4
    assert pre(t, g, a, 0.1, 0.2, q=q, tnom=0.0)["info"][O.PI_STATUS] == 5
    assert pre(t, g, a, 0.1, 0.2, tnom=0.0)["info"][O.PI_STATUS] == 0  # mean only needs no rate


def test_sqrt_info_whitens():
    t, g, a = stream(10)
    r = pre(t, g, a, 0.05, 0.5, q=(1e-4, 2e-3, 1e-5, 3e-4))
    for kappa, sbg2, sba2 in [(1.0, 0.0, 0.0), (2.5, 1e-4, 3e-3)]:
        W = O.sqrt_info(r["C"], kappa, sbg2, sba2, r["J"])
        A = kappa * r["C"].copy()
        Jg, Ja = r["J"][:, :3], r["J"][:, 3:]
        A[:9, :9] += sbg2 * Jg @ Jg.T + sba2 * Ja @ Ja.T
        assert np.abs(np.triu(W, 1)).max() == 0
        assert np.abs(W.T @ W @ A - np.eye(15)).max() < 1e-9
    assert O.sqrt_info(-r["C"], 1.0, 0, 0, r["J"]) is None
