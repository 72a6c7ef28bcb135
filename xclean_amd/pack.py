"""A window Problem packed from sliding-window engine state, and the solved state written back.

  pack_problem(eng, lm_select=None, imu_sel=None, with_prior=True, posefree=False) -> (Problem, meta)
      vio/engine_live.py:407-539 (LiveEngine._problem)
  writeback(eng, pb, meta)
      vio/engine_live.py:541-553 (LiveEngine._writeback)

``eng`` is any object carrying the engine's window state under the engine's attribute names:
  slots          list of keyframe dicts: sid, R (3,3), p (3,), full (bool), lid / uv / cam / valid / sig (the
                 keyframe's observations: landmark id, pixel (2,), camera 0 left / 1 right, valid flag, the LK
                 sigma of a right-camera observation) and, in VI, v / bg / ba (3,)
  lm_alive, n_lm, lm_host (host sid per landmark id), lm_p (n, 3) [a, b, rho] in the host camera
  stereo_mode    "prior" (the inverse-depth prior rows) or "ur" (right-camera rows)
  lm_rho_meas, lm_sig_lk, lm_span, kappa_d, fB, sig_px
  vi, prior (a Prior or None), imu_factors (list of dicts i, j (sids), pre, W), Rbc, pbc, fx, fy, cx, cy, B, g_w

The packed arrays are the CSR-by-landmark wire format the solver reads: observations grouped by the landmark's
position in the kept landmark list, and inside one landmark in keyframe order, then in the keyframe's own order
(``meta["src"]`` gives every row's (keyframe, index) origin).
"""
from __future__ import annotations

import numpy as np

from .problem import PoseFree, Problem

HUBER_K = 1.345  # core/derive.py:47


def _slot_rows(slots):
    """Every observation of every keyframe, flattened in keyframe order: (lid, uv, cam, valid, sig, pose, src)."""
    parts = [(i, s) for i, s in enumerate(slots) if len(s["lid"])]
    if not parts:
        return (np.zeros(0, np.int64), np.zeros((0, 2)), np.zeros(0, np.int64), np.zeros(0, bool), np.zeros(0),
                np.zeros(0, np.int64), np.zeros((0, 2), np.int64))
    lid = np.concatenate([np.asarray(s["lid"], np.int64) for _, s in parts])
    uv = np.concatenate([np.asarray(s["uv"], np.float64).reshape(-1, 2) for _, s in parts])
    cam = np.concatenate([np.asarray(s["cam"], np.int64) for _, s in parts])
    val = np.concatenate([np.asarray(s["valid"], bool) for _, s in parts])
    sig = np.concatenate([np.asarray(s["sig"], np.float64) for _, s in parts])
    n = [len(s["lid"]) for _, s in parts]
    pose = np.repeat(np.array([i for i, _ in parts], np.int64), n)
    within = np.concatenate([np.arange(k, dtype=np.int64) for k in n])
    return lid, uv, cam, val, sig, pose, np.stack([pose, within], 1)


def _csr_start(local, L):
    start = np.zeros(L + 1, np.int64)
    if L:
        start[1:] = np.cumsum(np.bincount(local, minlength=L))
    return start


def pack_problem(eng, lm_select=None, imu_sel=None, with_prior=True, posefree=False):
    """vio/engine_live.py:407-539: the window Problem of the engine's current state and the bookkeeping that maps
    it back (meta: alive, src, sid2i, local, off, dims, pf_lids).

    lm_select(alive) -> an index or mask into the alive landmark ids (None: every alive landmark).  imu_sel: the
    indices into eng.imu_factors to include (None: all).  posefree: landmarks whose one valid observation is their
    host's left one leave the Schur arrays for ``Problem.pf`` (stereo_mode "prior" only)."""
    slots = eng.slots
    sid2i = {s["sid"]: i for i, s in enumerate(slots)}
    alive = np.flatnonzero(np.asarray(eng.lm_alive)[:eng.n_lm])
    if lm_select is not None:
        alive = alive[lm_select(alive)]
    L = len(alive)
    pos_of_lid = np.full(eng.n_lm, -1, np.int64)
    pos_of_lid[alive] = np.arange(L)

    lid, uv, cam, val, sig, pose, src = _slot_rows(slots)
    local = pos_of_lid[lid] if len(lid) else np.zeros(0, np.int64)
    kept = np.flatnonzero(local >= 0)
    # stable grouping by landmark position: ties keep keyframe order, then each keyframe's own order
    kept = kept[np.argsort(local[kept], kind="stable")]
    local, uv, cam, val, sig, pose, src = (local[kept], uv[kept], cam[kept], val[kept], sig[kept], pose[kept],
                                           src[kept])
    start = _csr_start(local, L)
    host = (np.fromiter((sid2i[int(h)] for h in np.asarray(eng.lm_host)[alive]), np.int64, count=L)
            if L else np.zeros(0, np.int64))

    if eng.stereo_mode == "prior":
        rho_meas = np.asarray(eng.lm_rho_meas, np.float64)[alive]
        # an ambiguous disparity's basin span floors its sigma
        sig_rho = np.maximum(eng.kappa_d * np.asarray(eng.lm_sig_lk, np.float64)[alive],
                             np.asarray(eng.lm_span, np.float64)[alive]) / eng.fB
        usable = np.isfinite(sig_rho) & (sig_rho > 0) & (rho_meas > 0)
        rho_isig = np.zeros(L)
        rho_isig[usable] = 1.0 / sig_rho[usable]
    else:
        rho_meas = np.zeros(L)
        rho_isig = np.zeros(L)

    pf = None
    pf_lids = alive[:0]
    if posefree and eng.stereo_mode == "prior" and L:
        on_host = pose == host[local]
        n_valid = np.bincount(local, weights=val, minlength=L)
        n_off_host = np.bincount(local, weights=val & ~on_host, minlength=L)
        host_left = val & on_host & (cam == 0)
        n_host_left = np.bincount(local, weights=host_left, minlength=L)
        is_pf = (n_off_host == 0) & (n_valid == 1) & (n_host_left == 1)
        if is_pf.any():
            z = np.zeros((L, 2))
            z[local[host_left]] = uv[host_left]
            pf_lids = alive[is_pf]
            pf = PoseFree(host[is_pf], z[is_pf], 1.0 / eng.sig_px, rho_meas[is_pf], rho_isig[is_pf],
                          np.asarray(eng.lm_p, np.float64)[pf_lids])
            keep_lm = ~is_pf
            renumber = np.cumsum(keep_lm) - 1
            rows = keep_lm[local]
            local = renumber[local[rows]]
            uv, cam, val, sig, pose, src = uv[rows], cam[rows], val[rows], sig[rows], pose[rows], src[rows]
            alive, host, rho_meas, rho_isig = alive[keep_lm], host[keep_lm], rho_meas[keep_lm], rho_isig[keep_lm]
            L = len(alive)
            start = _csr_start(local, L)

    dims = np.array([15 if (eng.vi and s["full"]) else 6 for s in slots], np.int64)
    off = np.concatenate([[0], np.cumsum(dims)]).astype(np.int64)
    prior = eng.prior if with_prior else None
    if prior is not None:
        if any(sd not in sid2i for sd in prior.sids):
            raise RuntimeError("prior refers to a slot that left the window")
        prior.bind(sid2i, off, dims)

    vel = bgs = bas = None
    imu = []
    if eng.vi:
        vel = np.ascontiguousarray(np.stack([s["v"] for s in slots]), np.float64)
        bgs = np.ascontiguousarray(np.stack([s["bg"] for s in slots]), np.float64)
        bas = np.ascontiguousarray(np.stack([s["ba"] for s in slots]), np.float64)
        chosen = None if imu_sel is None else set(int(k) for k in imu_sel)
        for k, f in enumerate(eng.imu_factors):
            if chosen is None or k in chosen:
                imu.append((sid2i[f["i"]], sid2i[f["j"]], f["pre"], f["W"]))

    # whitening of every row at the engine's current scales: sigma_px on the left camera, kappa_d times the
    # observation's own LK sigma on the right one
    ob_isig = np.where(cam == 0, 1.0 / eng.sig_px, 1.0 / np.maximum(eng.kappa_d * sig, 1e-9))
    pb = Problem(Rwb=np.ascontiguousarray(np.stack([s["R"] for s in slots]), np.float64),
                 pwb=np.ascontiguousarray(np.stack([s["p"] for s in slots]), np.float64),
                 Rbc=eng.Rbc, pbc=eng.pbc, lmp=np.ascontiguousarray(np.asarray(eng.lm_p, np.float64)[alive]),
                 lm_host=host, lm_mask=np.ones(L, np.bool_), rho_meas=np.ascontiguousarray(rho_meas),
                 rho_isig=np.ascontiguousarray(rho_isig), ob_start=start, ob_pose=np.ascontiguousarray(pose),
                 ob_cam=np.ascontiguousarray(cam), ob_z=np.ascontiguousarray(uv),
                 ob_isig=np.ascontiguousarray(ob_isig), ob_valid=np.ascontiguousarray(val),
                 fx=eng.fx, fy=eng.fy, cx=eng.cx, cy=eng.cy, B=eng.B, huber_k=HUBER_K, prior=prior,
                 dims=dims if eng.vi else None, vel=vel, bgs=bgs, bas=bas, imu=imu, g_w=eng.g_w, pf=pf)
    meta = {"alive": alive, "src": src, "sid2i": sid2i, "local": local, "off": off, "dims": dims,
            "pf_lids": pf_lids}
    return pb, meta


def writeback(eng, pb, meta):
    """vio/engine_live.py:541-553: the solved keyframe states (and, for full VI keyframes, velocity and biases)
    and landmark states back into the engine."""
    for i, s in enumerate(eng.slots):
        s["R"] = np.array(pb.Rwb[i], copy=True)
        s["p"] = np.array(pb.pwb[i], copy=True)
        if pb.vel is not None and s["full"]:
            s["v"] = np.array(pb.vel[i], copy=True)
            s["bg"] = np.array(pb.bgs[i], copy=True)
            s["ba"] = np.array(pb.bas[i], copy=True)
    eng.lm_p[meta["alive"]] = pb.lmp
    if getattr(pb, "pf", None) is not None:
        eng.lm_p[meta["pf_lids"]] = pb.pf.lmp
