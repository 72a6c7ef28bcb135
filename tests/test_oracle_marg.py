"""Pins the oracle's restatement of the marginalisation algebra (oracle/oracle.py marg_schur, prior_from_info,
combine, boxminus; vio/marg.py:47-56, 73-84, 165-177, 212-242) with checks the mathematics fixes: the block
inverse identity and the reduced solve for the Schur complement, the observable-subspace reduction for a
rank-deficient marginalised block, S^T S = H and S^T r = b (projected on the range), the ascending eigenvalue rows,
the FEJ combine's gradient and Hessian at the current state, and Log(R0^T R0 Exp(w)) = w."""
import numpy as np
from scipy.linalg import expm

from oracle import oracle as O


def _pd(n, rng, cond=1e4):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.geomspace(1.0, cond, n)) @ Q.T


def _hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def test_schur_is_the_block_inverse():
    rng = np.random.default_rng(0)
    for n, m in [(12, 6), (30, 15), (45, 6)]:
        H, b = _pd(n, rng), rng.standard_normal(n)
        marg = np.sort(rng.choice(n, m, replace=False))
        keep = np.setdiff1d(np.arange(n), marg)
        H2, b2 = O.marg_schur(H, b, keep, marg)
        Hi = np.linalg.inv(H)
        assert np.abs(np.linalg.inv(H2) - Hi[np.ix_(keep, keep)]).max() < 1e-9 * np.abs(Hi).max()
        x = np.linalg.solve(H, b)
        assert np.abs(np.linalg.solve(H2, b2) - x[keep]).max() < 1e-9 * np.abs(x).max()


def test_schur_of_a_rank_deficient_block_keeps_its_observable_part():
    rng = np.random.default_rng(1)
    n, m = 20, 6
    # marginalised coordinates rotated by Q; the last rotated direction is unobserved (no information, no coupling)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    Hr = _pd(n - 1, rng)                     # over (kept, observable marginalised)
    br = rng.standard_normal(n - 1)
    nk = n - m
    T = np.eye(n)
    T[nk:, nk:] = Q
    Hfull = np.zeros((n, n))
    Hfull[:n - 1, :n - 1] = Hr
    bfull = np.concatenate([br, [0.0]])
    H = T @ Hfull @ T.T                      # in the original marginalised coordinates
    b = T @ bfull
    keep, marg = np.arange(nk), np.arange(nk, n)
    H2, b2 = O.marg_schur(H, b, keep, marg)
    # reference: ordinary Schur over the observable coordinates
    ko, mo = np.arange(nk), np.arange(nk, n - 1)
    Hmm = Hr[np.ix_(mo, mo)]
    Hkm = Hr[np.ix_(ko, mo)]
    want = Hr[np.ix_(ko, ko)] - Hkm @ np.linalg.solve(Hmm, Hkm.T)
    wb = br[ko] - Hkm @ np.linalg.solve(Hmm, br[mo])
    assert np.abs(H2 - want).max() < 1e-9 * np.abs(want).max()
    assert np.abs(b2 - wb).max() < 1e-9 * np.abs(wb).max()


def test_square_root_of_the_information():
    rng = np.random.default_rng(2)
    H, b = _pd(24, rng, 1e6), rng.standard_normal(24)
    S, r = O.prior_from_info(H, b)
    assert S.shape == (24, 24)
    assert np.abs(S.T @ S - H).max() < 1e-9 * np.abs(H).max()
    assert np.abs(S.T @ r - b).max() < 1e-9 * np.abs(b).max()
    rows = np.sum(S * S, axis=1)
    assert np.all(np.diff(rows) >= -1e-9 * rows.max())          # ascending eigenvalues
    # a gauge-like null space: rank n - 3, b in the range
    A = rng.standard_normal((21, 24))
    H = A.T @ A
    b = H @ rng.standard_normal(24)
    S, r = O.prior_from_info(H, b)
    assert S.shape == (21, 24)
    assert np.abs(S.T @ S - H).max() < 1e-9 * np.abs(H).max()
    assert np.abs(S.T @ r - b).max() < 1e-8 * np.abs(b).max()


def test_boxminus_and_combine_fej():
    rng = np.random.default_rng(3)
    R0 = expm(_hat(rng.standard_normal(3)))
    w = 0.3 * rng.standard_normal(3)
    cur = (R0 @ expm(_hat(w)), rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3),
           rng.standard_normal(3))
    x0 = (R0, rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3))
    d = O.boxminus(cur, x0, 15)
    assert np.abs(d[:3] - w).max() < 1e-13
    assert np.abs(d[3:6] - (cur[1] - x0[1])).max() == 0 and np.abs(d[12:] - (cur[4] - x0[4])).max() == 0
    # combine: blocks 10 (old, 15 dims), 11 (old pose block grown to 15), 12 (new); the new prior's quadratic model
    # has the current information as its Hessian and b_cur as its gradient at the current state
    z3 = np.zeros(3)
    old = ([10, 11], [15, 6], [x0, (expm(_hat(rng.standard_normal(3))), rng.standard_normal(3), z3, z3, z3)])
    states = [cur, (old[2][1][0] @ expm(_hat(0.1 * rng.standard_normal(3))), rng.standard_normal(3),
                    rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)),
              (np.eye(3), rng.standard_normal(3), z3, z3, z3)]
    dims = [15, 15, 6]
    H, b = _pd(36, rng), rng.standard_normal(36)
    x0s, S, r = O.combine(old, [10, 11, 12], dims, states, H, b)
    assert x0s[0] is x0
    assert np.array_equal(x0s[1][0], old[2][1][0]) and np.array_equal(x0s[1][2], states[1][2])
    assert x0s[2] is states[2]
    dcur = np.concatenate([O.boxminus(st, x, dm) for st, x, dm in zip(states, x0s, dims)])
    assert np.abs(S.T @ S - H).max() < 1e-9 * np.abs(H).max()
    assert np.abs(S.T @ (S @ dcur + r) - b).max() < 1e-8 * max(np.abs(b).max(), 1.0)
