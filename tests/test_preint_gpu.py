"""Parity of the batched GPU IMU preintegration (xclean_amd.imu -> xba_preintegrate) with the CPU oracle
(oracle/xba_oracle.c o_preint / o_sqrt_info, restating vio/imu_factor.py:392-746), FP64: dR within 1e-13, the other
outputs within 1e-12 relative (max-norm), the info vector exactly, on one batch that mixes regular and jittered
streams, gaps, holds past the last sample, wholly held intervals, causal horizons with poisoned later samples,
non-zero linearisation biases, several intervals per stream, and every refusal code."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TN = 0.005


def _stream(seed, n=160, jitter=0.0, drop=()):
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
    g = np.stack([0.6 * np.sin(f[i] * 6.28 * t + ph[i]) for i in range(3)], 1) + rng.normal(0, 0.02, (len(t), 3))
    a = np.stack([1.5 * np.sin(f[3 + i] * 6.28 * t + ph[3 + i]) for i in range(3)], 1) + [0.3, -0.2, 9.81]
    return t, g, a + rng.normal(0, 0.05, (len(t), 3))


def _case():
    streams = [_stream(1), _stream(2, jitter=0.3), _stream(3, drop=(40, 41, 90)), _stream(4, jitter=0.1)]
    t_poison, g_poison, a_poison = _stream(5)
    hz = t_poison[100] + 1e-4
    g_poison = g_poison.copy()
    g_poison[101:] = np.nan
    streams.append((t_poison, g_poison, a_poison))
    t_bad, g_bad, a_bad = _stream(6)
    g_bad = g_bad.copy()
    g_bad[30, 2] = np.nan
    streams.append((t_bad, g_bad, a_bad))
    rng = np.random.default_rng(0)
    iv = []
    for si in range(4):
        te = streams[si][0][-1]
        for t0, t1 in [(0.0123, 0.1555), (0.05, 0.45), (0.2, 0.2 + 0.0031), (te - 0.1, te + 0.03),
                       (te + 0.01, te + 0.05), (0.1, 0.1 + 2 * TN)]:
            iv.append((si, t0, t1, None, rng.normal(0, 0.02, 3), rng.normal(0, 0.05, 3)))
    iv.append((4, 0.2, 0.7, hz, np.zeros(3), np.zeros(3)))       # causal: held past the horizon
    iv.append((4, 0.6, 0.9, hz, np.zeros(3), np.zeros(3)))       # wholly past the horizon
    iv.append((5, 0.1, 0.3, None, np.zeros(3), np.zeros(3)))     # non-finite sample used: refused (4)
    iv.append((0, 0.3, 0.3, None, np.zeros(3), np.zeros(3)))     # t1 <= t0 (1)
    iv.append((0, 0.1, 0.2, -1.0, np.zeros(3), np.zeros(3)))     # nothing released (2)
    iv.append((0, -0.01, 0.2, None, np.zeros(3), np.zeros(3)))   # before the first sample (3)
    return streams, iv


def _oracle(streams, iv, noise, kappa, sbg, sba):
    from xclean_amd.imu import _as_noise, t_nom
    nz = _as_noise(noise)
    q = (0.0, 0.0, 0.0, 0.0) if nz is None else nz.squares()
    out = []
    for si, t0, t1, hz, bg, ba in iv:
        t, g, a = streams[si]
        tn = t_nom(nz, t, np.inf if hz is None else hz)
        r = O.preintegrate(t, g, a, t0, t1, hz, bg, ba, *q, tn)
        r["W"] = O.sqrt_info(r["C"], kappa, sbg ** 2, sba ** 2, r["J"]) if r["info"][0] == 0 else None
        out.append(r)
    return out


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("noise", [None, (1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3), (1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3, 190.0)])
def test_batch_matches_oracle(noise):
    from xclean_amd.imu import preintegrate_many
    streams, iv = _case()
    kappa, sbg, sba = 1.5, 2e-3, 1e-2
    gpu = preintegrate_many(streams, iv, noise, kappa, sbg, sba, whiten=True)
    ref = _oracle(streams, iv, noise, kappa, sbg, sba)
    statuses = set()
    for f, (p, r) in enumerate(zip(gpu, ref)):
        np.testing.assert_array_equal(np.array([p.status, p.n_sub, p.n_inner, p.n_released]),
                                      r["info"][[0, 1, 2, 6]], err_msg="interval %d" % f)
        statuses.add(p.status)
        if not p.ok:
            continue
        for x, y in [(p.extrap_span, r["info"][3]), (p.t_data_end, r["info"][4]), (p.gap_max, r["info"][5]),
                     (p.t_nom, r["info"][7]), (p.white_scale_max, r["info"][8])]:
            assert abs(x - y) <= 1e-15 * max(1.0, abs(y)), f
        assert np.abs(p.dR - r["dR"]).max() < 1e-13, f
        assert _rel(p.dv, r["dv"]) < 1e-12 and _rel(p.dp, r["dp"]) < 1e-12, f
        assert _rel(p.J, r["J"]) < 1e-12, f
        if noise is not None:
            assert _rel(p.cov, r["C"]) < 1e-12, f
            assert (p.W is None) == (r["W"] is None), f
            if p.W is not None:
                assert _rel(p.W, r["W"]) < 1e-10, f
        else:
            assert np.abs(p.cov).max() == 0.0
    assert statuses >= ({0, 1, 2, 3, 4} if noise is not None else {0, 1, 2, 3, 4})


def test_refused_without_rate_and_window_records():
    import torch

    from xclean_amd import imu
    streams, iv = _case()
    iv = iv[:6]
    noise = imu.ImuNoise(1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3)
    t, g, a, s0, ns, par = imu.pack(streams, iv, noise, 1.0, 0.0, 0.0)
    par[2, 13] = 0.0  # no sample period with densities: refused (5)
    T = lambda x, dt: torch.as_tensor(x, dtype=dt, device="cuda").contiguous()
    out = imu.launch(T(t, torch.float64), T(g, torch.float64), T(a, torch.float64), T(s0, torch.int64),
                     T(ns, torch.int32), T(par, torch.float64), want_rec=True)
    info = out["info"].cpu().numpy()
    assert info[2, 0] == 5 and (info[[0, 1, 3, 4, 5], 0] == 0).all()
    rec = out["rec"].cpu().numpy()
    for f in (0, 1, 3):
        r = O.preintegrate(t[s0[f]:s0[f] + ns[f]], g[s0[f]:s0[f] + ns[f]], a[s0[f]:s0[f] + ns[f]], par[f, 0],
                           par[f, 1], None, par[f, 3:6], par[f, 6:9], *par[f, 9:14])
        W = O.sqrt_info(r["C"], 1.0, 0.0, 0.0, r["J"])
        want = np.concatenate([r["dR"].reshape(9), r["dv"], r["dp"], [par[f, 1] - par[f, 0]], r["J"].reshape(54),
                               par[f, 3:6], par[f, 6:9], W.reshape(225)])
        assert np.abs(rec[f] - want).max() <= 1e-10 * np.abs(want).max(), f
    assert np.abs(rec[2]).max() == 0.0


def test_single_interval_api_matches_reference_signature():
    from xclean_amd import imu
    t, g, a = _stream(11)
    pre = imu.preintegrate(t, g, a, 0.0123, 0.4567, None, np.zeros(3), np.zeros(3),
                           {"gyroscope_noise_density": 1.7e-4, "accelerometer_noise_density": 2e-3,
                            "gyroscope_random_walk": 1.9e-5, "accelerometer_random_walk": 3e-3})
    r = O.preintegrate(t, g, a, 0.0123, 0.4567, None, np.zeros(3), np.zeros(3), 1.7e-4 ** 2, 2e-3 ** 2,
                       1.9e-5 ** 2, 3e-3 ** 2, float(np.median(np.diff(t))))
    assert pre.ok and not pre.extrapolated and pre.n_inner == r["info"][2]
    assert _rel(pre.cov, r["C"]) < 1e-12 and np.abs(pre.dR - r["dR"]).max() < 1e-13
    assert pre.JRg.shape == (3, 3) and pre.cov9.shape == (9, 9)


def test_sqrt_information_honours_kappa_and_bias_sigmas():
    """vio/imu_factor.py:1051-1063: W for the caller's kappa and bias sigmas, against the oracle's sqrt_info_core
    restatement and against the defining identity W^T W = (kappa C + BIASCOV)^-1."""
    from xclean_amd import imu
    t, g, a = _stream(12)
    noise = imu.ImuNoise(1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3, rate_hz=200.0)
    pres = [imu.preintegrate(t, g, a, t0, t1, None, np.full(3, 0.01), np.full(3, -0.02), noise)
            for t0, t1 in ((0.01, 0.31), (0.2, 0.7), (0.05, 0.06))]
    for kappa, sg, sa in ((1.0, 0.0, 0.0), (4.0, 0.0, 0.0), (2.5, 1e-3, 2e-2), (0.3, 5e-4, 0.0)):
        Ws = imu.sqrt_information_many(pres, kappa, sg, sa)
        for p, W in zip(pres, Ws):
            want = O.sqrt_info(p.cov, kappa, sg ** 2, sa ** 2, p.J)
            assert W is not None and want is not None
            assert _rel(W, want) < 1e-12
            A = kappa * p.cov.copy()
            A[:9, :9] += sg ** 2 * p.J[:, :3] @ p.J[:, :3].T + sa ** 2 * p.J[:, 3:] @ p.J[:, 3:].T
            assert _rel(W.T @ W, np.linalg.inv(A)) < 1e-8
            assert np.abs(np.triu(W, 1)).max() == 0.0
        W1 = imu.sqrt_information(pres[1], kappa, sg, sa)
        assert _rel(W1, Ws[1]) == 0.0
    # kappa scales W by 1 / sqrt(kappa) when there is no bias term
    W1 = imu.sqrt_information(pres[0], 1.0)
    W4 = imu.sqrt_information(pres[0], 4.0)
    assert _rel(W4, 0.5 * W1) < 1e-13


def test_sqrt_information_refusals():
    from xclean_amd import imu
    t, g, a = _stream(13)
    noise = imu.ImuNoise(1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3, rate_hz=200.0)
    good = imu.preintegrate(t, g, a, 0.02, 0.3, None, np.zeros(3), np.zeros(3), noise)
    refused = imu.preintegrate(t, g, a, 0.3, 0.2, None, np.zeros(3), np.zeros(3), noise)
    assert not refused.ok
    with pytest.raises(ValueError):
        imu.sqrt_information(refused, 1.0)
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            imu.sqrt_information(good, bad)
    # a mean-only interval has a zero covariance: not positive definite
    mean_only = imu.preintegrate(t, g, a, 0.02, 0.3, None, np.zeros(3), np.zeros(3), None)
    assert imu.sqrt_information_many([mean_only])[0] is None
    with pytest.raises(ValueError):
        imu.sqrt_information(mean_only)


def test_engine_call_pair_preintegrate_then_sqrt_information():
    """The engine's call pair (vio/engine_live.py:898, 1134): preintegrate(...) then sqrt_information(pre, kappa)
    equals the whitened batch launch at that kappa."""
    from xclean_amd import imu
    t, g, a = _stream(14)
    noise = imu.ImuNoise(1.7e-4, 2.0e-3, 1.9e-5, 3.0e-3)
    pre = imu.preintegrate(t, g, a, 0.0123, 0.4567, None, np.zeros(3), np.zeros(3), noise)
    W = imu.sqrt_information(pre, 3.0)
    batch = imu.preintegrate_many([(t, g, a)], [(0, 0.0123, 0.4567, None, np.zeros(3), np.zeros(3))], noise,
                                  kappa=3.0, whiten=True)[0]
    assert _rel(W, batch.W) < 1e-13
    assert _rel(W, O.sqrt_info(pre.cov, 3.0, 0.0, 0.0, pre.J)) < 1e-12
