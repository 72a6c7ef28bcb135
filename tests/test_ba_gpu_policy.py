"""The GPU-driver twins of the local-BA drop-ins (xclean_amd.ba._solve_gpu / _solve_gpu_id -> libxba with
lm_policy = 1), which mirror backend/ba.py:930-1089 `_solve_gpu` and :1345-1497 `_solve_gpu_id`: the same solves
under that driver's LM rule (the step applied at once, rejected by the next iteration when the Huber-weighted chi2
grew, lambda x4 / x10 / halved, the rotation factors' Hessian in the reduced system, SO(3) projected at the end).

Bars: poses, points (inverse depths, bearings) and stats within 1e-9 relative of the oracle's restatement of that
rule (FP64); the reference's own agreement bar between its GPU and numpy paths (src/xclean/tests/test_ba.py:56-70:
max pose delta < 5e-3 m, final rms within 0.05 px, poses moved toward the truth); build_pairs' pair count.
The reference's GPU drivers need CuPy and NVRTC, which this image lacks, so no reference outputs of them exist:
parity with them is pinned through the shared math (golden vectors of _solve_np / _solve_np_id) and the restated
LM rule only.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import synth

pytestmark = pytest.mark.gpu

KEYS = ("chi2_0", "chi2_n", "rms_0", "rms_n")


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _problems():
    out = [("default", synth.make_problem(), 0, 5, None, None)]
    pb = synth.make_problem(nk=8, nl=200, obs_per=4, seed=3)
    free = np.ones(8, bool)
    free[[0, 5]] = False
    out.append(("mask", pb, free, 4, None, None))
    out.append(("rejects", synth.make_problem(nk=6, nl=80, obs_per=3, seed=11, pose_noise=0.3, pt_noise=0.8,
                                              px_noise=8.0), 0, 6, None, None))
    pb = synth.make_problem(nk=7, nl=120, obs_per=3, seed=7)
    rng = np.random.RandomState(2)
    nl = len(pb["pts0"])
    c = pb["pts_true"] + rng.randn(nl, 3) * 0.3
    rho0 = 1.0 / np.linalg.norm(pb["pts_true"] - c, axis=1) * (1 + 0.05 * rng.randn(nl))
    valid = (rng.rand(nl) > 0.2).astype(float)
    rotf = []
    for (i, j) in ((0, 1), (1, 2), (3, 5)):
        Ri = np.linalg.inv(pb["poses_true"][i])[:3, :3]
        Rj = np.linalg.inv(pb["poses_true"][j])[:3, :3]
        rotf.append((i, j, Ri @ Rj.T @ synth.so3_exp(rng.randn(3) * 0.002), 0.01))
    out.append(("prior_rotf", pb, 0, 4, (c, rho0, 2.0, valid), rotf))
    return out


@pytest.mark.parametrize("idx", range(4))
def test_solve_gpu_matches_the_oracle_under_the_gpu_rule(idx):
    from xclean_amd.ba import _solve_gpu
    name, pb, free, iters, prior, rotf = _problems()[idx]
    args = (pb["poses0"], pb["pts0"], pb["okf"], pb["olm"], pb["ouv"], pb["rig"], free, iters)
    po, qo, so = O.solve_np(*args, prior=prior, rotf=rotf, policy="gpu")
    p, q, st = _solve_gpu(*args, prior=prior, rotf=rotf)
    print(name, "poses %.2e points %.2e" % (_rel(p, po), _rel(q, qo)), st, flush=True)
    assert _rel(p, po) < 1e-9 and _rel(q, qo) < 1e-9
    np.testing.assert_allclose([st[k] for k in KEYS], [so[k] for k in KEYS], rtol=1e-9)
    counts = np.bincount(pb["olm"], minlength=len(pb["pts0"]))
    assert st["n_obs"] == len(pb["okf"]) and st["npair"] == int((counts ** 2).sum())


def test_solve_gpu_meets_the_reference_agreement_bar_against_solve_np():
    from xclean_amd.ba import _solve_gpu, _solve_np
    pb = synth.make_problem()
    args = (pb["poses0"].copy(), pb["pts0"].copy(), pb["okf"], pb["olm"], pb["ouv"], pb["rig"], 0, 5)
    p_ref, q_ref, st_ref = _solve_np(*args)
    p, q, st = _solve_gpu(*args)
    dt = float(np.abs(p[:, :3, 3] - p_ref[:, :3, 3]).max())
    e0 = np.sqrt(np.mean(np.sum((pb["poses0"][:, :3, 3] - pb["poses_true"][:, :3, 3]) ** 2, 1)))
    e1 = np.sqrt(np.mean(np.sum((p[:, :3, 3] - pb["poses_true"][:, :3, 3]) ** 2, 1)))
    print("max pose delta %.2e m, rms %.4f vs %.4f, pose rmse %.4f -> %.4f" % (dt, st["rms_n"], st_ref["rms_n"], e0,
                                                                             e1), flush=True)
    assert dt < 5e-3 and abs(st["rms_n"] - st_ref["rms_n"]) < 0.05
    assert e1 < 0.6 * e0


@pytest.mark.parametrize("seed,nk,nl,ext", [(21, 10, 300, 6), (23, 6, 60, 2)])
def test_solve_gpu_id_matches_the_oracle_under_the_gpu_rule(seed, nk, nl, ext):
    from xclean_amd.ba import _solve_gpu_id
    q = synth.to_inverse_depth(synth.make_problem(nk=nk, nl=nl, obs_per=4, seed=seed), ext_every=ext,
                               ext_shift=[0.02, 0.0, 0.01, 0.0, 0.003, 0.0])
    args = (q["poses0"], q["rho0"], q["mbear0"], q["hidx"], q["hpose_ext"], q["okf"], q["olm"], q["ouv"], q["rig"],
            0, 5)
    po, ro, mo, so = O.solve_np_id(*args, policy="gpu")
    p, rho, mb, st = _solve_gpu_id(*args)
    print("seed", seed, "poses %.2e rho %.2e" % (_rel(p, po), _rel(rho, ro)), flush=True)
    assert _rel(p, po) < 1e-9 and _rel(rho, ro) < 1e-9 and _rel(mb, mo) < 1e-9
    np.testing.assert_allclose([st[k] for k in KEYS], [so[k] for k in KEYS], rtol=1e-9)
