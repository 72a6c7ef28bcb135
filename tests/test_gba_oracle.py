"""The global-BA oracle (oracle/gba_oracle.py) against what the mathematics fixes, on the CPU."""
import numpy as np
import pytest

from oracle import gba_oracle as G


def _brute_pairs(olm_sorted, nl):
    offs = [0]
    for l in range(nl):
        offs.append(offs[-1] + sum(1 for v in olm_sorted if v == l))
    pa, pb = [], []
    for l in range(nl):
        for a in range(offs[l], offs[l + 1]):
            for b in range(offs[l], offs[l + 1]):
                pa.append(a)
                pb.append(b)
    return np.array(offs), np.array(pa, dtype=np.int64), np.array(pb, dtype=np.int64)


@pytest.mark.parametrize("seed", range(6))
def test_build_pairs_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(0, 12))
    counts = rng.integers(0, 6, nl) if nl else np.zeros(0, int)
    olm = np.repeat(np.arange(nl), counts)
    offs, pa, pb = G.build_pairs(olm, nl)
    bo, ba, bb = _brute_pairs(list(olm), nl)
    assert offs.dtype == np.int32 and pa.dtype == np.int32 and pb.dtype == np.int32
    assert np.array_equal(offs, bo)
    assert np.array_equal(pa, ba) and np.array_equal(pb, bb)
    assert len(pa) == int(np.sum(counts.astype(np.int64) ** 2))


def test_build_pairs_trailing_and_leading_empty_landmarks():
    olm = np.array([2, 2, 3])
    offs, pa, pb = G.build_pairs(olm, 6)
    assert offs.tolist() == [0, 0, 0, 2, 3, 3, 3]
    assert list(zip(pa.tolist(), pb.tolist())) == [(0, 0), (0, 1), (1, 0), (1, 1), (2, 2)]


# ---------------------------------------------------------------------------------------------------------------
# global BA
# ---------------------------------------------------------------------------------------------------------------
from xclean_amd import synth  # noqa: E402

ARGS = ("kf_ids", "poses", "rho", "mbear", "host_idx", "okf", "olm", "ouv", "rig", "odisp", "prior")


def _args(g, **over):
    a = {k: g[k] for k in ARGS}
    a.update(over)
    return a


def test_gba_residuals_vanish_at_the_noise_free_truth():
    g = synth.make_gba(nk=30, nl=800, seed=3, noise_scale=0.0, outlier_frac=0.0)
    tr = g["truth"]
    T_cw = G.se3_inv(tr["poses"])
    rig = g["rig"]
    r, Pc, _, _ = G.gba_geom(T_cw, tr["rho"], tr["mbear"], g["host_idx"], g["okf"], g["olm"], g["ouv"],
                             rig.fx_rect, rig.cx, rig.cy)
    assert np.abs(r).max() < 1e-9
    d = g["odisp"]
    assert G.stereo_chi(Pc[:, 2], d, rig.fB, 1.0 / rig.rho_sigma) < 1e-18 * len(d)
    # and the camera-frame points are the generator's world points seen from each camera
    Rwc, pwc = tr["poses"][:, :3, :3], tr["poses"][:, :3, 3]
    q = np.einsum("nji,nj->ni", Rwc[g["okf"]], tr["Xw"][g["olm"]] - pwc[g["okf"]])
    assert np.allclose(Pc, q, rtol=0, atol=1e-9)


def _pose_factor_rows(T_cw, rotf, welds):
    """Whitened rotation-factor rows w Log(M^T R_a R_b^T) and weld rows W^(1/2) Log(Z^-1 T_cw_a T_wc_b)."""
    from scipy.spatial.transform import Rotation
    out = []
    for a, b, M, sig in rotf:
        A = T_cw[a][:3, :3] @ T_cw[b][:3, :3].T
        out.append(Rotation.from_matrix(M.T @ A).as_rotvec() / sig)
    for a, b, Z, w in welds:
        E = G.se3_inv(Z) @ T_cw[a] @ G.se3_inv(T_cw[b])
        out.append(np.sqrt(np.asarray(w, np.float64)) * G.se3_log(E))
    return np.concatenate(out) if out else np.zeros(0)


def _whitened_rows(T_cw, rho, mb, g, keep, sw, sws, okd, prior, swp):
    """Every whitened residual of the system with frozen IRLS weights: vision (2 per row), disparity (1 per row
    with one), prior (3 per valid landmark)."""
    rig = g["rig"]
    okf, olm, ouv = g["okf"][keep], g["olm"][keep], g["ouv"][keep]
    r, Pc, _, _ = G.gba_geom(T_cw, rho, mb, g["host_idx"], okf, olm, ouv, rig.fx_rect, rig.cx, rig.cy)
    d = g["odisp"][keep]
    rs = (1.0 / np.maximum(Pc[okd, 2], 1e-6) - d[okd] / rig.fB) / rig.rho_sigma
    m0, rho0, wb, isg, valid = prior
    rp = np.stack([(mb[:, 0] - m0[:, 0]) * wb, (mb[:, 1] - m0[:, 1]) * wb, (rho - rho0) * isg], 1)
    ok = valid > 0.5
    return np.concatenate([(r * sw[:, None]).ravel(), rs * sws, (rp * swp)[ok].ravel()])


@pytest.mark.parametrize("lam,factors", [(0.0, False), (1e-3, False), (1e-3, True)])
def test_gba_reduced_system_matches_a_dense_fd_schur_complement(lam, factors):
    """The reduced system of gba_system against a finite-difference Jacobian of every whitened residual (poses by a
    left-multiplied twist, landmarks by (a, b, rho)), the dense normal equations, the landmark damping
    lam max(tr/3, 1e-9), a dense Schur complement, the gauge rows and the pose damping; with gyro rotation factors
    (a non-zero residual: their Jacobian is exact) and loop welds (at zero residual, where the reference's weld
    Jacobian, which leaves out Jr^-1(e), is exact)."""
    import scipy.linalg as sla
    from scipy.spatial.transform import Rotation
    g = synth.make_gba(nk=6, nl=40, seed=11, reach=6, pose_sigma_t=0.01)
    T0 = G.se3_inv(g["poses"])
    rotf, welds = [], []
    if factors:
        for a, b, sig in ((1, 2, 0.01), (3, 5, 0.02)):
            A = T0[a][:3, :3] @ T0[b][:3, :3].T
            rotf.append((a, b, A @ Rotation.from_rotvec([0.01, -0.02, 0.015]).as_matrix(), sig))
        welds.append((2, 4, T0[2] @ G.se3_inv(T0[4]), np.array([4e2, 5e2, 6e2, 1e3, 2e3, 3e3])))
    rig = g["rig"]
    nk, nl = len(g["poses"]), len(g["rho"])
    T_cw = G.se3_inv(g["poses"])
    keep = g["host_idx"][g["olm"]] != g["okf"]
    okf, olm, ouv = g["okf"][keep], g["olm"][keep], g["ouv"][keep]
    hk = g["host_idx"][olm]
    d = g["odisp"][keep]
    free = np.ones(nk, bool)
    free[0] = False
    cam = (rig.fx_rect, rig.cx, rig.cy, rig.fB, 1.0 / rig.rho_sigma)
    out = G.gba_system(T_cw, g["rho"], g["mbear"], g["host_idx"], okf, olm, ouv, hk, cam, d, g["prior"], lam, free,
                       rotf=rotf, welds=welds)
    # frozen weights at the linearisation point
    r, Pc, _, _ = G.gba_geom(T_cw, g["rho"], g["mbear"], g["host_idx"], okf, olm, ouv, *cam[:3])
    sw = np.sqrt(G._huber_w(np.linalg.norm(r, axis=1)))
    okd = d > 1e-6
    rs = (1.0 / np.maximum(Pc[okd, 2], 1e-6) - d[okd] / rig.fB) / rig.rho_sigma
    sws = np.sqrt(G._huber_w(np.abs(rs)))
    _rp, _J0, swp, _ok = G.prior_rows(g["prior"], g["rho"], g["mbear"])
    args = (g, keep, sw, sws, okd, g["prior"], swp)

    def f(x):
        Tn = T_cw.copy()
        for k in range(nk):
            Tn[k] = G.se3_exp_many(x[6 * k:6 * k + 6])[0] @ T_cw[k]
        lm = x[6 * nk:].reshape(nl, 3)
        mb = g["mbear"].copy()
        mb[:, :2] += lm[:, :2]
        return np.concatenate([_whitened_rows(Tn, g["rho"] + lm[:, 2], mb, *args), _pose_factor_rows(Tn, rotf, welds)])

    nx = 6 * nk + 3 * nl
    h = 1e-6
    J = np.stack([(f(h * e) - f(-h * e)) / (2 * h) for e in np.eye(nx)], 1)
    res = f(np.zeros(nx))
    Hf = J.T @ J
    gf = -J.T @ res
    n = 6 * nk
    App, Apl, All = Hf[:n, :n], Hf[:n, n:], Hf[n:, n:].copy()
    for l in range(nl):
        s = slice(3 * l, 3 * l + 3)
        blk = All[s, s]
        All[s, s] = blk + lam * max(np.trace(blk) / 3.0, 1e-9) * np.eye(3)
    # the landmark block is block diagonal: its exact inverse
    Alli = np.zeros_like(All)
    for l in range(nl):
        s = slice(3 * l, 3 * l + 3)
        Alli[s, s] = np.linalg.inv(All[s, s])
    Hr = App - Apl @ Alli @ Apl.T
    br = gf[:n] - Apl @ Alli @ gf[n:]
    Hr[:6, :] = 0.0
    Hr[:, :6] = 0.0
    Hr[:6, :6] = np.eye(6)
    br[:6] = 0.0
    Hr = Hr + np.diag(lam * np.maximum(np.diag(Hr), 1e-9))
    Hg = out["H"].toarray()
    assert np.abs(Hg - Hr).max() <= 2e-6 * np.abs(Hr).max()
    assert np.abs(out["bp"] - br).max() <= 2e-6 * max(np.abs(br).max(), 1.0)
    # the landmark back-substitution reproduces the full-system step
    if lam > 0:
        Hfull = Hf.copy()
        Hfull[n:, n:] = All
        Hfull[:6, :] = 0.0
        Hfull[:, :6] = 0.0
        Hfull[:6, :6] = np.eye(6)
        gfull = gf.copy()
        gfull[:6] = 0.0
        dx = np.linalg.solve(out["H"].toarray(), out["bp"])
        dxl = np.einsum("lab,lb->la", out["Hlli"], out["bl"] - np.asarray(out["U"].T @ dx).reshape(nl, 3))
        # the pose damping enters the reduced system only: compare against the full solve with it added
        Hfull[:n, :n] += np.diag(lam * np.maximum(np.diag(Hr) / (1 + lam), 1e-9))
        full = sla.solve(Hfull, gfull)
        assert np.allclose(dx, full[:n], rtol=0, atol=1e-6 * np.abs(full).max())
        assert np.allclose(dxl.ravel(), full[n:], rtol=0, atol=1e-6 * np.abs(full).max())


def test_gba_rigid_and_scale_modes_are_null_vectors_of_the_visual_system():
    """With no prior, no disparity rows, (almost) no damping and no gauge the reduced visual system does not see a rigid
    motion or a scaling of the whole map (inverse-depth landmarks are host-relative): the deflation modes of every
    keyframe, stacked, are null vectors of it."""
    g = synth.make_gba(nk=8, nl=300, seed=12, reach=8, outlier_frac=0.0)
    rig = g["rig"]
    nk = len(g["poses"])
    T_cw = G.se3_inv(g["poses"])
    keep = g["host_idx"][g["olm"]] != g["okf"]
    okf, olm, ouv = g["okf"][keep], g["olm"][keep], g["ouv"][keep]
    cam = (rig.fx_rect, rig.cx, rig.cy, rig.fB, 1.0 / rig.rho_sigma)
    # lambda 1e-12: the landmark blocks of one-row landmarks become invertible; their Schur term tends to the
    # projection onto the rows' range, which the limit keeps
    out = G.gba_system(T_cw, g["rho"], g["mbear"], g["host_idx"], okf, olm, ouv, g["host_idx"][olm], cam, None,
                       None, 1e-12, np.ones(nk, bool))
    H = out["H"].toarray()
    M = G.mode_vectors(T_cw, np.ones(nk, bool)).reshape(6 * nk, 7)
    scale = np.abs(H).max() * np.abs(M).max()
    assert np.abs(H @ M).max() <= 1e-7 * scale
    # a mode of the wrong sign on the rotation block is not a null vector
    Mw = G.mode_vectors(T_cw, np.ones(nk, bool))
    Mw[:, :3, 3:6] *= -1.0
    assert np.abs(H @ Mw.reshape(6 * nk, 7)).max() > 1e-3 * scale


def test_dpcg_reaches_the_direct_solution():
    import scipy.sparse.linalg as spl
    g = synth.make_gba(nk=80, nl=4000, seed=5)
    rig = g["rig"]
    nk = len(g["poses"])
    T_cw = G.se3_inv(g["poses"])
    keep = g["host_idx"][g["olm"]] != g["okf"]
    okf, olm, ouv = g["okf"][keep], g["olm"][keep], g["ouv"][keep]
    free = np.ones(nk, bool)
    free[0] = False
    cam = (rig.fx_rect, rig.cx, rig.cy, rig.fB, 1.0 / rig.rho_sigma)
    out = G.gba_system(T_cw, g["rho"], g["mbear"], g["host_idx"], okf, olm, ouv, g["host_idx"][olm], cam,
                       g["odisp"][keep], g["prior"], 1e-4, free)
    H, b = out["H"], out["bp"]
    ref = spl.spsolve(H.tocsc(), b)
    for m in (None, 4, 9):
        x, info = G.dpcg(H, b, T_cw, free, seg_len=m)
        assert info["converged"] and info["reason"] == "tol", info
        assert np.linalg.norm(b - H @ x) <= 1e-8 * np.linalg.norm(b)
        assert np.linalg.norm(x - ref) <= 1e-5 * np.linalg.norm(ref)
        assert info["iters"] < 300


def test_gba_lm_recovers_the_truth_without_noise():
    g = synth.make_gba(nk=40, nl=2500, seed=7, noise_scale=0.0, outlier_frac=0.0, pose_sigma_t=0.01,
                       pose_sigma_r_deg=0.1)
    poses, rho, mb, st = G.global_ba_id(**_args(g), iters=12, passes=1)
    tr = g["truth"]
    assert st["rms_px_after"] < 1e-6 < st["rms_px_before"]
    assert np.abs(poses[:, :3, 3] - tr["poses"][:, :3, 3]).max() < 1e-6
    assert np.abs(rho - tr["rho"]).max() < 1e-6


def test_gba_second_pass_rejects_the_planted_outliers():
    g = synth.make_gba(nk=50, nl=3000, seed=8, outlier_frac=0.05)
    h = []
    poses, rho, mb, st = G.global_ba_id(**_args(g), history=h)
    keep = g["host_idx"][g["olm"]] != g["okf"]
    planted = int(g["truth"]["outlier"][keep].sum())
    assert planted > 50
    assert planted <= st["rejected"] <= planted + 0.05 * st["observations"]
    assert st["rms_px_after"] < 0.8 < st["rms_px_before"]
    assert st["lm_pass1"]["accepted"] >= 1 and st["lm"]["accepted"] >= 1
