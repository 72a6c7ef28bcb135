"""Parity of the GPU marginalisation algebra (xclean_amd.marg -> xba_marg_schur / xba_prior_from_info, FP64 Jacobi)
with the numpy oracle restating vio/marg.py, and of marg_system with the oracle's final linearisation on a masked
window.  Bars: Schur complements and the information form of every square root (S^T S, S^T r) within 1e-10 relative
(max-norm), ranks equal, eigenvalues within 1e-12 of the largest; the engine-style marginalisation of a window's
oldest keyframe within 1e-9."""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pd(n, rng, cond=1e4):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.geomspace(1.0, cond, n)) @ Q.T


def _hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _psd(n, rank, rng):
    A = rng.standard_normal((rank, n))
    return A.T @ A


def test_schur_batch_matches_oracle():
    from xclean_amd import marg
    rng = np.random.default_rng(0)
    cases = []
    for n, m in [(12, 6), (60, 15), (200, 6), (90, 30), (7, 1)]:
        H, b = _pd(n, rng), rng.standard_normal(n)
        mg = np.sort(rng.choice(n, m, replace=False))
        cases.append((H, b, np.setdiff1d(np.arange(n), mg), mg))
    # a rank-deficient marginalised block (its null direction uncoupled)
    n, m = 40, 6
    H = _psd(n, n - 2, rng)
    cases.append((H, H @ rng.standard_normal(n), np.arange(n - m), np.arange(n - m, n)))
    got = marg.schur_batch([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], [c[3] for c in cases])
    for (H, b, k, mg), (H2, b2) in zip(cases, got):
        W2, w2 = O.marg_schur(H, b, k, mg)
        assert _rel(H2, W2) < 1e-10 and _rel(b2, w2) < 1e-10


def test_from_info_batch_matches_oracle():
    from xclean_amd import marg
    rng = np.random.default_rng(1)
    items = [(_pd(24, rng, 1e8), rng.standard_normal(24)), (_pd(150, rng, 1e6), rng.standard_normal(150))]
    H = _psd(60, 54, rng)
    items.append((H, H @ rng.standard_normal(60)))
    items.append((np.diag([4.0, 1.0, 9.0]), np.array([1.0, 2.0, 3.0])))
    got = marg.from_info_batch(items)
    for (H, b), (S, r, lam) in zip(items, got):
        Sr, rr = O.prior_from_info(H, b)
        assert S.shape == Sr.shape
        assert _rel(S.T @ S, Sr.T @ Sr) < 1e-10 and _rel(S.T @ r, Sr.T @ rr) < 1e-10
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))
        assert np.abs(lam - ev).max() < 1e-12 * ev.max()
        assert np.allclose(np.sum(S * S, 1), ev[len(ev) - len(S):], rtol=1e-10, atol=1e-12 * ev.max())


def test_from_info_every_matrix_layout_matches_oracle():
    """One batch (leading dimension 120) through each in-kernel layout: orders up to 112 keep the packed upper
    triangle and the eigenvectors in LDS (112 the largest, 101 and 7 odd, padded by a decoupled zero row), 120 keeps
    the whole matrix in LDS with the eigenvectors in the global workspace."""
    from xclean_amd import marg
    rng = np.random.default_rng(3)
    items = [(_pd(n, rng, 1e6), rng.standard_normal(n)) for n in (120, 112, 101, 7)]
    H = _psd(101, 95, rng)
    items.append((H, H @ rng.standard_normal(101)))
    got = marg.from_info_batch(items)
    for (H, b), (S, r, lam) in zip(items, got):
        Sr, rr = O.prior_from_info(H, b)
        assert S.shape == Sr.shape
        assert _rel(S.T @ S, Sr.T @ Sr) < 1e-10 and _rel(S.T @ r, Sr.T @ rr) < 1e-10
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))
        assert np.abs(lam - ev).max() < 1e-12 * ev.max()


def _state(rng, R=None):
    return (expm(_hat(rng.standard_normal(3))) if R is None else R, rng.standard_normal(3), rng.standard_normal(3),
            rng.standard_normal(3), rng.standard_normal(3))


def test_combine_drop_shrink_add_match_oracle():
    from xclean_amd import marg
    from xclean_amd.problem import Prior
    rng = np.random.default_rng(2)
    z3 = np.zeros(3)
    x0a, x0b = _state(rng), (expm(_hat(rng.standard_normal(3))), rng.standard_normal(3), z3, z3, z3)
    old = Prior([10, 11], [15, 6], [x0a, x0b], np.eye(21), np.zeros(21))
    states = [_state(rng, x0a[0] @ expm(_hat(0.2 * rng.standard_normal(3)))),
              _state(rng, x0b[0] @ expm(_hat(0.1 * rng.standard_normal(3)))), _state(rng)]
    dims = [15, 15, 6]
    H, b = _pd(36, rng), rng.standard_normal(36)
    p = marg.combine(old, [10, 11, 12], dims, states, H, b)
    x0s, Sr, rr = O.combine(([10, 11], [15, 6], old.x0), [10, 11, 12], dims, states, H, b)
    assert _rel(p.S.T @ p.S, Sr.T @ Sr) < 1e-10 and _rel(p.S.T @ p.r, Sr.T @ rr) < 1e-10
    for a, c in zip(p.x0, x0s):
        for u, v in zip(a, c):
            np.testing.assert_array_equal(u, v)
    # drop / shrink: the oracle's Schur of the information form, then its square root
    Hp, bp = Sr.T @ Sr, Sr.T @ rr
    q = marg.drop_slot(p, 11)
    keep = np.r_[0:15, 30:36]
    W2, w2 = O.marg_schur(Hp, bp, keep, np.r_[15:30])
    S2, r2 = O.prior_from_info(W2, w2)
    assert q.sids == [10, 12] and _rel(q.S.T @ q.S, S2.T @ S2) < 1e-9 and _rel(q.S.T @ q.r, S2.T @ r2) < 1e-9
    q = marg.shrink_slot(p, 10)
    keep = np.r_[0:6, 15:36]
    W2, w2 = O.marg_schur(Hp, bp, keep, np.r_[6:15])
    S2, r2 = O.prior_from_info(W2, w2)
    assert q.dims == [6, 15, 6] and _rel(q.S.T @ q.S, S2.T @ S2) < 1e-9
    # add_block_info on an existing 6-dim block that grows to 15, at a state away from its x0
    st = _state(rng, x0b[0] @ expm(_hat(0.1 * rng.standard_normal(3))))
    Hb, bb = _pd(15, rng), rng.standard_normal(15)
    q = marg.add_block_info(old, 11, 15, st, Hb, bb)
    H0 = np.zeros((30, 30))
    H0[:15, :15] = np.eye(15)
    H0[15:21, 15:21] = np.eye(6)
    H0[15:30, 15:30] += Hb
    xk = (x0b[0], x0b[1], st[2], st[3], st[4])
    b0 = np.zeros(30)
    b0[15:30] = bb - Hb @ O.boxminus(st, xk, 15)
    S2, r2 = O.prior_from_info(H0, b0)
    assert q.dims == [15, 15] and _rel(q.S.T @ q.S, S2.T @ S2) < 1e-10 and _rel(q.S.T @ q.r, S2.T @ r2) < 1e-9


def test_marginalise_oldest_keyframe_like_the_engine():
    """vio/engine_live.py:690-739: the oldest keyframe's hosted landmarks' visual rows and the prior through
    marg_system, the keyframe's 6 columns Schur'd out, the prior rebuilt over the other blocks (FEJ)."""
    from xclean_amd import marg, synth
    pb = synth.make_window(P=8, L=500, views=(2, 6), seed=[9, 1])[0]
    hosted = np.asarray(pb.lm_host) == 0
    H, b, off = marg.marg_system(pb, hosted)
    q = pb.copy()
    q.lm_mask = hosted.copy()
    Hr, br = O.window_final(q)[:2]
    assert _rel(H, Hr) < 1e-12 and _rel(b, br) < 1e-12
    P = pb.P
    mset = set(range(int(off[0]), int(off[1])))
    blocks = [(i, [c for c in range(int(off[i]), int(off[i + 1])) if c not in mset]) for i in range(P)]
    blocks = [(i, c) for i, c in blocks if c and np.abs(np.diag(H)[c]).sum() > 0]
    keep = np.concatenate([np.array(c) for _i, c in blocks])
    mg = np.array(sorted(mset))
    H2, b2 = marg.schur(H, b, keep, mg)
    W2, w2 = O.marg_schur(Hr, br, keep, mg)
    assert _rel(H2, W2) < 1e-9 and _rel(b2, w2) < 1e-9
    sids = [i for i, _c in blocks]
    dims = [len(c) for _i, c in blocks]
    states = [(pb.Rwb[i], pb.pwb[i], None, None, None) for i in sids]
    p = marg.combine(pb.prior, sids, dims, states, H2, b2)
    ost = [(pb.Rwb[i], pb.pwb[i], np.zeros(3), np.zeros(3), np.zeros(3)) for i in sids]
    x0s, Sr, rr = O.combine((pb.prior.sids, pb.prior.dims, pb.prior.x0), sids, dims, ost, W2, w2)
    assert p.S.shape == Sr.shape
    assert _rel(p.S.T @ p.S, Sr.T @ Sr) < 1e-9 and _rel(p.S.T @ p.r, Sr.T @ rr) < 1e-9
