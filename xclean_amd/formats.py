"""The reference engine's on-disk window format, so recorded windows replay through the GPU solver.

Writer and reader of ``win_%05d.npz`` (vio/engine_live.py:804-821 ``_dump``, written every few tracked frames when
XCLEAN_P1_DUMP names a directory).  One file holds one VO window as the engine hands it to ``window.solve``
(vio/window.py:768-792 ``Problem``, built by engine_live.py:407-539 ``_problem``): slot poses R_wb, p_wb, the
body-camera extrinsic, the landmarks (a, b, rho), their hosts and inverse-depth priors, the observations as CSR by
landmark, the camera ``[fx, fy, cx, cy, B]`` and the square-root prior (S, r, the slot index and dims of each prior
block, each block's linearisation point R0, p0).

What the dump does not carry, and the reader therefore sets: ``lm_mask`` (every landmark in, as ``_problem`` builds
it), ``huber_k`` (core/derive.py:47 HUBER_K, the engine's threshold), no VI states, IMU factors or pose-free block.  A
prior block of 15 dims would need v0, bg0, ba0, which the dump lacks: such a file is refused.  ``pb_%05d.pkl``
(engine_live.py:546-551, the pickled Problem) is not read: unpickling needs the reference's classes and runs code
from the file.
"""
from __future__ import annotations

import glob
import os

import numpy as np

from .problem import Prior, Problem

HUBER_K = 1.345  # core/derive.py:47

# the arrays of a dump, in the order engine_live.py:806-820 writes them
FIELDS = ("Rwb", "pwb", "Rbc", "pbc", "lmp", "lm_host", "rho_meas", "rho_isig", "ob_start", "ob_pose", "ob_cam",
          "ob_z", "ob_isig", "ob_valid", "cam", "prior_S", "prior_r", "prior_idx", "prior_dims", "prior_R0",
          "prior_p0")


def save_window_npz(pb, path):
    """Write ``pb`` (a VO Problem) in the dump layout.  The prior's blocks are stored by the slot index they are bound
    to; VI states, IMU factors, the pose-free block, ``lm_mask`` and ``huber_k`` are not part of the format."""
    pr = pb.prior
    if pr is not None and any(d != 6 for d in pr.dims):
        raise ValueError("the window dump format stores pose (6-dim) prior blocks only")
    np.savez_compressed(
        path, Rwb=pb.Rwb, pwb=pb.pwb, Rbc=pb.Rbc, pbc=pb.pbc, lmp=pb.lmp, lm_host=pb.lm_host, rho_meas=pb.rho_meas,
        rho_isig=pb.rho_isig, ob_start=pb.ob_start, ob_pose=pb.ob_pose, ob_cam=pb.ob_cam, ob_z=pb.ob_z,
        ob_isig=pb.ob_isig, ob_valid=pb.ob_valid, cam=np.array([pb.fx, pb.fy, pb.cx, pb.cy, pb.B]),
        prior_S=(pr.S if pr is not None else np.zeros((0, 0))),
        prior_r=(pr.r if pr is not None else np.zeros(0)),
        prior_idx=(np.asarray(pr._idx, np.int64) if pr is not None else np.zeros(0, np.int64)),
        prior_dims=(np.asarray(pr.dims, np.int64) if pr is not None else np.zeros(0, np.int64)),
        prior_R0=(np.stack([x[0] for x in pr.x0]) if pr is not None else np.zeros((0, 3, 3))),
        prior_p0=(np.stack([x[1] for x in pr.x0]) if pr is not None else np.zeros((0, 3))))


def load_window_npz(path, huber_k=HUBER_K):
    """Read one dump into a ``Problem`` (numpy arrays only; nothing in the file is executed)."""
    with np.load(path, allow_pickle=False) as z:
        missing = [k for k in FIELDS if k not in z.files]
        if missing:
            raise ValueError("%s is not a window dump: missing %s" % (path, ", ".join(missing)))
        a = {k: z[k] for k in FIELDS}
    P, L, M = a["Rwb"].shape[0], a["lmp"].shape[0], a["ob_pose"].shape[0]
    if a["Rwb"].shape != (P, 3, 3) or a["pwb"].shape != (P, 3) or a["lmp"].shape != (L, 3):
        raise ValueError("%s: inconsistent slot / landmark arrays" % path)
    if a["ob_start"].shape != (L + 1,) or int(a["ob_start"][-1]) != M or a["ob_z"].reshape(-1, 2).shape[0] != M:
        raise ValueError("%s: inconsistent observation CSR" % path)
    fx, fy, cx, cy, B = (float(v) for v in a["cam"])
    prior = None
    nb = a["prior_idx"].shape[0]
    if nb:
        dims = [int(d) for d in a["prior_dims"]]
        if any(d != 6 for d in dims):
            raise ValueError("%s: a 15-dim prior block needs v0, bg0, ba0, which the dump does not hold" % path)
        idx = [int(i) for i in a["prior_idx"]]
        if min(idx) < 0 or max(idx) >= P:
            raise ValueError("%s: prior block slot index out of range" % path)
        prior = Prior(idx, dims, [(a["prior_R0"][k], a["prior_p0"][k], None, None, None) for k in range(nb)],
                      a["prior_S"], a["prior_r"])
        prior.bind({i: i for i in idx})
    f64 = lambda x: np.ascontiguousarray(x, np.float64)
    i64 = lambda x: np.ascontiguousarray(x, np.int64)
    return Problem(Rwb=f64(a["Rwb"]), pwb=f64(a["pwb"]), Rbc=f64(a["Rbc"]), pbc=f64(a["pbc"]), lmp=f64(a["lmp"]),
                   lm_host=i64(a["lm_host"]), lm_mask=np.ones(L, np.bool_), rho_meas=f64(a["rho_meas"]),
                   rho_isig=f64(a["rho_isig"]), ob_start=i64(a["ob_start"]), ob_pose=i64(a["ob_pose"]),
                   ob_cam=i64(a["ob_cam"]), ob_z=f64(a["ob_z"]).reshape(-1, 2), ob_isig=f64(a["ob_isig"]),
                   ob_valid=np.ascontiguousarray(a["ob_valid"], np.bool_), fx=fx, fy=fy, cx=cx, cy=cy, B=B,
                   huber_k=float(huber_k), prior=prior)


def load_window_dir(directory, pattern="win_*.npz"):
    """Every dump of a directory in file-name order (the engine's ``win_%05d`` counter)."""
    return [load_window_npz(p) for p in sorted(glob.glob(os.path.join(directory, pattern)))]
