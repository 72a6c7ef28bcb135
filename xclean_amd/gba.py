"""Run-end global BA in inverse depth on the GPU: the drop-in for backend/global_ba.py:694 ``global_ba_id``.

The whole LM solve runs in libxba (``xba_gba_solve``): the reduced camera system is assembled on the device row by
row (one workgroup per block row, landmarks re-linearised per visit, the Schur terms staged in LDS) and solved by
the two-level deflated PCG of backend/gba_linsolve.py (segment smoother, rigid-motion and scale deflation space,
true-residual stopping rule); the host only marshals arrays and reads the statistics.

Covered: the reprojection rows with the Huber IRLS weight and optional per-keyframe weights (``vt_weights``), the
per-observation disparity rows (``odisp``), the landmark prior (with a constant ``beta0``), the gyro rotation factors
(``rotf``, ``rotf_huber``), the loop welds (``welds``), the gauge keyframes (``fixed``), the LM accept / damping /
best-state rules, the wall-clock budget, the MAD rejection and the second pass with the re-weighted prior, the final
SO(3) projection, the sampled marginal covariances (``want_cov``, ``stats["cov"]``).  Not covered
(NotImplementedError): the inertial term, the disparity offset states and the XTEN_GBA_SCREEN / ACCEPT / ITERBUDGET /
BESTFIX modes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

DISPARITY_NOISE_PX = 0.5   # core/derive.py:60
RTOL = 1e-8                # gba_linsolve.py:57, above XBA_GBA_DIRECT_MAX unknowns
RTOL_DIRECT = 1e-12        # at or below it the reference solves directly: the PCG is driven to this instead,
RTOL_DIRECT_FLOOR = 1e-9   # and its best iterate used when the recursion stops short of it within this
MAXITER = 2000             # gba_linsolve.py:62

PCG_REASONS = ("running", "tol", "maxiter", "breakdown", "divergence", "stagnation", "zero_rhs", "nonfinite_rhs",
               "setup")

_MODES_OFF = {"XTEN_GBA_SCREEN": ("", "0", "off"), "XTEN_GBA_ACCEPT": ("", "0", "off", "default", "partial"),
              "XTEN_GBA_ITERBUDGET": ("", "0", "off"), "XTEN_GBA_BESTFIX": ("", "0"),
              "XTEN_GBA_SOLVER": ("", "0", "off", "dpcg"), "XTEN_GBA_SCALECAL": ("", "0", "off", "default"),
              "XTEN_INERTIAL_FIX": ("", "0"), "XTEN_IF_BESTFIX": ("", "0")}


def _check_modes():
    for k, ok in _MODES_OFF.items():
        v = (os.environ.get(k, "") or "").strip().lower()
        if v not in ok:
            raise NotImplementedError("%s=%r selects a global-BA mode this solver does not implement" % (k, v))


def global_ba_id(kf_ids, poses, rho, mbear, host_idx, okf, olm, ouv, rig, odisp=None, prior=None, rotf=None,
                 welds=None, iters=8, fixed=(0,), lam0=1e-4, verbose=False, time_budget_s=None, passes=2,
                 reject_sigma=3.0, want_cov=0, rho_floor=1e-4, inertial=None, scale_bias=False, beta0=0.0,
                 rotf_huber=None, vt_weights=None, dof=None, _screen=True, device=None, rtol=RTOL,
                 rtol_direct=RTOL_DIRECT, rtol_direct_floor=RTOL_DIRECT_FLOOR, maxiter=MAXITER, seg_len=None,
                 history=None, profile=False):
    """backend/global_ba.py:694: (poses T_wc (nk,4,4), rho (nl,), mbear (nl,3), stats).

    poses: (nk,4,4) T_wc; rho, mbear: the landmarks' inverse depth and bearing in their host camera; host_idx (nl,);
    okf / olm / ouv: the observing keyframe, landmark and pixel of every row; rig: fx_rect, cx, cy and, with
    ``odisp``, fB and rho_sigma; prior: (m0 (nl,2), rho0, wbear, invsig, valid); rotf: [(a, b, M (3,3), sigma)];
    welds: [(a, b, Z (4,4), w scalar or (6,) information diagonal)].  ``history``, a list, receives one
    dict per LM iteration (pass, it, lam, chi, chi2, accepted, pcg_iters, relres, pcg_reason, restarts)."""
    import torch
    for name, v in (("inertial", inertial), ("dof", dof)):
        if v:
            raise NotImplementedError("global_ba_id(%s=...) is not implemented on the GPU" % name)
    if _screen:
        _check_modes()
    _lib.require_gpu()
    L = _lib.load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    poses = np.asarray(poses, np.float64).reshape(-1, 4, 4)
    nk = len(poses)
    rho = np.asarray(rho, np.float64).ravel()
    nl = len(rho)
    mbear = np.asarray(mbear, np.float64).reshape(nl, 3)
    host_idx = np.asarray(host_idx, np.int64).ravel()
    okf = np.asarray(okf, np.int64).ravel()
    olm = np.asarray(olm, np.int64).ravel()
    N = len(okf)
    ouv = np.asarray(ouv, np.float64).reshape(N, 2)
    if len(olm) != N or len(host_idx) != nl:
        raise ValueError("okf / olm / ouv and rho / host_idx must agree in length")
    if nk > _lib.XBA_GBA_MAX_NK:
        raise ValueError("global_ba_id: %d keyframes exceed XBA_GBA_MAX_NK" % nk)
    free = np.ones(nk, np.uint8)
    for f in fixed:
        if 0 <= f < nk:
            free[f] = 0

    def t(a, dt):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=dt)).to(dev)

    keep = []
    d_poses = t(poses.reshape(nk, 16), np.float64)
    d_rho = t(rho, np.float64)
    d_mb = t(mbear, np.float64)
    q = _lib.XbaGbaProblem()
    q.nk, q.nl, q.N = nk, nl, N
    q.poses, q.rho, q.mbear = d_poses.data_ptr(), d_rho.data_ptr(), d_mb.data_ptr()
    for name, arr, dt in (("host", host_idx, np.int32), ("okf", okf, np.int32), ("olm", olm, np.int32),
                          ("ouv", ouv, np.float64), ("free_mask", free, np.uint8)):
        d = t(arr, dt)
        keep.append(d)
        setattr(q, name, d.data_ptr())
    fB = float(getattr(rig, "fB", 0.0) or 0.0) if odisp is not None else 0.0
    if odisp is not None and fB > 0.0:
        d = t(np.asarray(odisp, np.float64).ravel(), np.float64)
        keep.append(d)
        q.odisp = d.data_ptr()
    if vt_weights is not None:
        d = t(np.asarray(vt_weights, np.float64).ravel(), np.float64)
        keep.append(d)
        q.vt_w = d.data_ptr()
    if prior is not None:
        m0, rho0, wbear, invsig, valid = prior
        # beta is not a state here (no inertial term): the prior rows read rho0 (1 + beta0) throughout
        rho0 = np.asarray(rho0, np.float64) * (1.0 + float(beta0))
        for name, arr in (("pr_m0", np.asarray(m0, np.float64).reshape(nl, 2)), ("pr_rho0", rho0),
                          ("pr_wbear", wbear), ("pr_invsig", invsig), ("pr_valid", valid)):
            d = t(np.asarray(arr, np.float64), np.float64)
            keep.append(d)
            setattr(q, name, d.data_ptr())
    if rotf:
        ab = np.array([[int(f[0]), int(f[1])] for f in rotf], np.int32)
        Ms = np.stack([np.asarray(f[2], np.float64).reshape(3, 3) for f in rotf])
        sg = np.array([float(f[3]) for f in rotf], np.float64)
        for name, arr, dt in (("rot_ab", ab, np.int32), ("rot_M", Ms, np.float64), ("rot_sig", sg, np.float64)):
            d = t(arr, dt)
            keep.append(d)
            setattr(q, name, d.data_ptr())
        q.n_rot = len(rotf)
        q.rot_huber = float(rotf_huber) if rotf_huber is not None else 0.0
    if welds:
        ab = np.array([[int(f[0]), int(f[1])] for f in welds], np.int32)
        Zs = np.stack([np.asarray(f[2], np.float64).reshape(4, 4) for f in welds]).copy()
        Zs[:, :3, 3] /= (1.0 + float(beta0))     # a weld's translation carries the map's scale (global_ba.py:2018)
        ws = np.stack([np.asarray(f[3], np.float64) * np.ones(6) if np.ndim(f[3]) == 0
                       else np.asarray(f[3], np.float64).reshape(6) for f in welds])
        for name, arr, dt in (("weld_ab", ab, np.int32), ("weld_Z", Zs, np.float64), ("weld_w", ws, np.float64)):
            d = t(arr, dt)
            keep.append(d)
            setattr(q, name, d.data_ptr())
        q.n_weld = len(welds)
    q.fx, q.cx, q.cy = float(rig.fx_rect), float(rig.cx), float(rig.cy)
    q.fB = fB
    q.rho_sigma = float(getattr(rig, "rho_sigma", 0.0) or 0.0) if odisp is not None else 0.0
    q.rho_floor = float(rho_floor)
    o = _lib.XbaGbaOptions()
    o.iters, o.passes, o.maxiter = int(iters), int(passes), int(maxiter)
    o.seg_len = int(seg_len or 0)
    o.profile = 1 if profile else 0
    o.lam0, o.reject_sigma = float(lam0), float(reject_sigma)
    o.time_budget_s = float(time_budget_s) if time_budget_s else 0.0
    o.rtol, o.rtol_direct, o.rtol_direct_floor = float(rtol), float(rtol_direct), float(rtol_direct_floor)
    want = max(int(want_cov or 0), 0)
    cov_ids = np.zeros(max(want, 1), np.int32)
    cov = np.zeros((max(want, 1), 6, 6))
    o.want_cov = want
    o.cov_ids = cov_ids.ctypes.data_as(C.c_void_p)
    o.cov = cov.ctypes.data_as(C.c_void_p)
    stats = np.zeros(_lib.XBA_GBA_NSTATS)
    hcap = 2 * max(int(iters), 0) + 2
    hist = np.zeros((hcap, _lib.XBA_GBA_HIST))
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        rc = L.xba_gba_solve(C.byref(q), C.byref(o), stats.ctypes.data_as(C.c_void_p),
                             hist.ctypes.data_as(C.c_void_p), hcap, C.c_void_p(stream))
    if rc == _lib.XBA_E_RANGE:
        raise ValueError("global_ba_id: an index is out of range (okf, olm or host_idx)")
    _lib.check(rc, "xba_gba_solve")
    out_poses = d_poses.cpu().numpy().reshape(nk, 4, 4)
    out_rho = d_rho.cpu().numpy()
    out_mb = d_mb.cpu().numpy()
    del keep
    st = _stats(stats, nk, nl, beta0)
    st["cov"] = {int(cov_ids[i]): cov[i].copy() for i in range(int(stats[32]))}
    n_it = int(stats[27])
    if history is not None:
        for row in hist[:min(n_it, hcap)]:
            history.append({"pass": int(row[0]), "it": int(row[1]), "lam": float(row[2]), "chi": float(row[3]),
                            "chi2": None if np.isnan(row[4]) else float(row[4]), "accepted": bool(row[5]),
                            "pcg_iters": int(row[6]), "relres": float(row[7]),
                            "pcg_reason": PCG_REASONS[int(row[8])], "restarts": int(row[9])})
    if verbose:
        print("   [gba/id gpu] rms %.3f -> %.3f px, %d rejected, %d PCG solves (%d iterations)"
              % (st["rms_px_before"], st["rms_px_after"], st["rejected"], st["gba_solver"]["solves"],
                 st["gba_solver"]["pcg_iters"]))
    return out_poses, out_rho, out_mb, st


def _f(v):
    return None if not np.isfinite(v) else float(v)


def _stats(s, nk, nl, beta0):
    """the statistics dict of global_ba.py:1760-1784 (and 1640-1660 after a second pass) from the stats array"""
    N = int(s[2])
    lm = {"chi0_vision": float(s[6]), "chi0_stereo": float(s[7]), "chi_best": float(s[8]), "accepted": int(s[9]),
          "best_updates": int(s[10])}
    st = {"rms_px_before": float(s[0]), "rms_px_after": float(s[1]), "keyframes": nk, "landmarks": nl,
          "observations": N, "rejected": int(s[3]), "rejected_pct": 100.0 * int(s[3]) / max(N, 1),
          "stereo_rows": int(s[5]), "stereo_scale_bias": float(beta0), "scale_bias_free": False,
          "measured_disparity_sigma_px": _f(s[17]), "measured_disparity_bias_px": _f(s[18]),
          "declared_disparity_sigma_px": DISPARITY_NOISE_PX, "lm": lm, "seconds": float(s[23]),
          "gba_solver": {"mode": "dpcg", "solves": int(s[20]), "pcg_iters": int(s[21]), "unconverged": int(s[22]),
                         "segments": int(s[24]), "seg_len": int(s[25]), "blocks": int(s[26]),
                         "rows_last_pass": int(s[31]), "iters_run": int(s[27]),
                         "assembly_ms": _f(s[28]), "setup_ms": _f(s[29]), "pcg_ms": _f(s[30])}}
    if int(s[16]) == 2:
        st["reject_threshold_px"] = float(s[4])
        st["lm_pass1"] = {"chi0_vision": float(s[11]), "chi0_stereo": float(s[12]), "chi_best": float(s[13]),
                          "accepted": int(s[14]), "best_updates": int(s[15])}
        if np.isfinite(s[19]):
            st["measured_sigma_is_first_pass"] = True
    return st
