"""Writes the golden fixtures of the drop-in BA entry point from the reference implementation.

Run in a container that has the reference checkout (it is never needed at test time):

    python tests/golden/make_golden.py /root/reference

It imports only the reference's pure-numpy functions whose exact outputs no mathematical check pins:
``xclean.backend.ba_synth.make_problem`` (its exact RNG stream), ``xclean.backend.ba._solve_np`` and
``xclean.backend.ba._solve_np_id`` (the whole fixed-iteration LM loops: Huber weights, Schur, damping, LU solve,
SE(3) update + SVD projection, accept/reject; the inverse-depth one also its host frames, bearing / inverse-range
prior, rotation-factor Hessian and inverse-depth floor).  The inverse-depth inputs are this repository's
``xclean_amd.synth.to_inverse_depth`` of a ``make_problem`` problem.
The window solve (vio/window.py) imports numba, which this image lacks, so it is not run here; its parts are pinned
by the checks in tests/test_oracle.py instead.  Outputs: tests/golden/solve_np_*.npz and solve_np_id_*.npz (inputs and outputs only).
"""
from __future__ import annotations

import os
import sys

import numpy as np


def main(ref_root):
    os.environ["XCLEAN_KEEP_CUDA_PATH"] = "1"      # skip the package's CUDA-toolkit probe at import
    sys.path.insert(0, os.path.join(ref_root, "src"))
    from xclean.backend.ba import _solve_np
    from xclean.backend.ba_synth import make_problem

    out_dir = os.path.dirname(os.path.abspath(__file__))

    # 1. make_problem determinism / RNG stream (tests/test_ba_solver.py:18-23 sizes)
    a = make_problem(nk=6, nl=60, seed=1)
    np.savez_compressed(os.path.join(out_dir, "make_problem_nk6_nl60_s1.npz"),
                        **{k: a[k] for k in ("poses0", "pts0", "okf", "olm", "ouv", "poses_true", "pts_true")})

    def case(name, pb, free, iters, lam0=1e-4, prior=None, rotf=None):
        p, q, st = _solve_np(pb["poses0"].copy(), pb["pts0"].copy(), pb["okf"], pb["olm"], pb["ouv"], pb["rig"],
                             free, iters, lam0, prior=prior, rotf=rotf)
        d = dict(poses0=pb["poses0"], pts0=pb["pts0"], okf=pb["okf"], olm=pb["olm"], ouv=pb["ouv"],
                 poses_true=pb["poses_true"], fx=pb["rig"].fx_rect, cx=pb["rig"].cx, cy=pb["rig"].cy,
                 free=np.asarray(free), iters=iters, lam0=lam0, out_poses=p, out_pts=q,
                 stats=np.array([st["chi2_0"], st["chi2_n"], st["rms_0"], st["rms_n"]]))
        if prior is not None:
            d.update(pr_c=prior[0], pr_rho0=prior[1], pr_invsig=prior[2], pr_valid=prior[3])
        if rotf:
            d.update(rot_ab=np.array([[f[0], f[1]] for f in rotf]), rot_M=np.stack([f[2] for f in rotf]),
                     rot_sig=np.array([f[3] for f in rotf]))
        np.savez_compressed(os.path.join(out_dir, "solve_np_%s.npz" % name), **d)
        print(name, st)

    # 2. the call tests/test_ba_solver.py:26-33 makes
    case("default", make_problem(), 0, 5)
    # 3. a bool free mask with two held keyframes, other damping
    pb = make_problem(nk=8, nl=200, obs_per=4, seed=3)
    free = np.ones(8, bool)
    free[[0, 5]] = False
    case("mask", pb, free, 3, lam0=1e-3)
    # 4. a hard problem: large noise so that steps are rejected and lambda moves both ways
    case("rejects", make_problem(nk=6, nl=80, obs_per=3, seed=11, pose_noise=0.3, pt_noise=0.8, px_noise=8.0), 0, 6)
    # 5. the optional inverse-range prior and gyro rotation factors
    pb = make_problem(nk=7, nl=120, obs_per=3, seed=7)
    rng = np.random.RandomState(2)
    nl = len(pb["pts0"])
    c = pb["pts_true"] + rng.randn(nl, 3) * 0.3
    rho0 = 1.0 / np.linalg.norm(pb["pts_true"] - c, axis=1) * (1 + 0.05 * rng.randn(nl))
    valid = (rng.rand(nl) > 0.2).astype(float)
    rotf = []
    T_true = pb["poses_true"]
    for (i, j) in ((0, 1), (1, 2), (3, 5), (5, 6)):
        Ri = np.linalg.inv(T_true[i])[:3, :3]
        Rj = np.linalg.inv(T_true[j])[:3, :3]
        M = Ri @ Rj.T
        from xclean.core.lie import so3_exp
        rotf.append((i, j, M @ so3_exp(rng.randn(3) * 0.002), 0.01))
    case("prior_rotf", pb, 0, 4, prior=(c, rho0, 2.0, valid), rotf=rotf)

    # 6. the inverse-depth solve (_solve_np_id) on make_problem problems converted by to_inverse_depth
    from xclean.backend.ba import _solve_np_id
    sys.path.insert(0, os.path.dirname(os.path.dirname(out_dir)))
    from xclean_amd.synth import to_inverse_depth

    def case_id(name, q, free, iters, lam0=1e-4, prior=None, rotf=None, rho_floor=1e-4):
        p, rho, mb, st = _solve_np_id(q["poses0"].copy(), q["rho0"].copy(), q["mbear0"].copy(), q["hidx"],
                                      q["hpose_ext"], q["okf"], q["olm"], q["ouv"], q["rig"], free, iters, lam0,
                                      prior=prior, rotf=rotf, rho_floor=rho_floor)
        d = dict(poses0=q["poses0"], rho0=q["rho0"], mbear0=q["mbear0"], hidx=q["hidx"], hpose_ext=q["hpose_ext"],
                 okf=q["okf"], olm=q["olm"], ouv=q["ouv"], poses_true=q["poses_true"], fx=q["rig"].fx_rect,
                 cx=q["rig"].cx, cy=q["rig"].cy, free=np.asarray(free), iters=iters, lam0=lam0,
                 rho_floor=rho_floor, out_poses=p, out_rho=rho, out_mbear=mb,
                 stats=np.array([st["chi2_0"], st["chi2_n"], st["rms_0"], st["rms_n"]]))
        if prior is not None:
            d.update(pr_m0=prior[0], pr_rho0=prior[1], pr_wbear=np.broadcast_to(prior[2], q["rho0"].shape),
                     pr_invsig=prior[3], pr_valid=prior[4])
        if rotf:
            d.update(rot_ab=np.array([[f[0], f[1]] for f in rotf]), rot_M=np.stack([f[2] for f in rotf]),
                     rot_sig=np.array([f[3] for f in rotf]))
        np.savez_compressed(os.path.join(out_dir, "solve_np_id_%s.npz" % name), **d)
        print("id", name, st)

    q = to_inverse_depth(make_problem(nk=8, nl=150, obs_per=4, seed=3), ext_every=7,
                         ext_shift=[0.01, 0.0, 0.0, 0.0, 0.002, 0.0])
    case_id("default", q, 0, 5)
    pb = make_problem(nk=7, nl=120, obs_per=3, seed=7)
    q = to_inverse_depth(pb, ext_every=5)
    rng = np.random.RandomState(4)
    nl = len(q["rho0"])
    m0 = q["mbear0"] + np.concatenate([rng.randn(nl, 2) * 0.002, np.zeros((nl, 1))], 1)
    rho_p = q["rho0"] * (1 + 0.05 * rng.randn(nl))
    invsig = 1.0 / (0.05 * q["rho0"] + 1e-3)
    valid = (rng.rand(nl) > 0.25).astype(float)
    rotf = []
    T_true = pb["poses_true"]
    from xclean.core.lie import so3_exp
    for (i, j) in ((0, 1), (2, 3), (4, 6)):
        Ri = np.linalg.inv(T_true[i])[:3, :3]
        Rj = np.linalg.inv(T_true[j])[:3, :3]
        rotf.append((i, j, Ri @ Rj.T @ so3_exp(rng.randn(3) * 0.002), 0.01))
    free = np.ones(7, bool)
    free[[0, 3]] = False
    case_id("mask_prior_rotf", q, free, 4, lam0=1e-3, prior=(m0, rho_p, np.full(nl, q["rig"].fx_rect), invsig, valid),
            rotf=rotf, rho_floor=0.15)
    q = to_inverse_depth(make_problem(nk=6, nl=80, obs_per=3, seed=11, pose_noise=0.3, pt_noise=0.8, px_noise=8.0))
    case_id("rejects", q, 0, 6)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
