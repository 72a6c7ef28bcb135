"""The local bundle adjustment drop-in: ``_solve_np`` with the reference's signature and results, solved on the GPU.

backend/ba.py:713-869 (``_solve_np``): Euclidean world points, camera poses T_cw = T_wc^-1 perturbed on the left,
Huber IRLS on the pixel-error norm, Marquardt damping of every 3x3 landmark block, Schur complement onto the poses,
held (non-free) keyframes taken out of the solve, damped reduced system, SE(3) update projected onto SO(3), and the
accept rule on the unrobust chi2 with the lambda schedule x0.5 (floor 1e-7) / x4 (cap 1e2).  The optional
inverse-range prior (ba.py:774-794) and gyro rotation factors (ba.py:644-697, which enter the gradient) are taken in
the reference's forms.

The whole loop runs in FP64 on the device as one stream of kernels (libxba ``xba_np_solve``); the host only plans
the landmark grouping.  There is no CPU path: without the built library or a GPU this raises.

``_solve_gpu`` / ``_solve_gpu_id`` (backend/ba.py:930-1089, 1345-1497) are the reference's GPU drivers of the same two
solves: same signatures and math, but their own LM rule -- every step is applied, the next iteration rejects it when
the Huber-weighted chi2 grew (restoring the last accepted state, lambda x4), a non-finite step sets lambda x10, the
rotation factors' Hessian enters the reduced system, SO(3) is projected once at the end, and the stats are the
Huber-weighted chi2 with the pair and observation counts.  Here they run in FP64 on the same kernels.

``_solve_np_id`` is the inverse-depth twin (backend/ba.py:1212-1342, libxba ``xba_np_id_solve``): landmarks are a
bearing (x_n, y_n, 1) and inverse depth in the camera frame of a host keyframe, or of a constant out-of-window host
pose; every observation has an observer and a host pose block; the bearing prior is unrobust, the inverse-range
prior Huber; the rotation factors' Hessian enters the reduced system; updated inverse depths are floored.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

HUBER_K = 1.345      # core/derive.py:47
D_POSE = 6           # backend/ba.py:58


def _free_mask(free, nk):
    """backend/ba.py:700-710: a boolean mask, or the single held (gauge) index."""
    if np.ndim(free) == 0:
        m = np.ones(nk, bool)
        m[int(free)] = False
        return m
    return np.asarray(free, bool)


def _rig_intrinsics(rig):
    return float(rig.fx_rect), float(rig.cx), float(rig.cy)


def _solve_np(poses, pts, okf, olm, ouv, rig, free, iters, lam0=1e-4, prior=None, rotf=None, device=None,
              history=False, _policy=0, noalloc=False):
    """backend/ba.py:713 ``_solve_np``.  Returns (poses T_wc (nk,4,4), pts (nl,3), stats dict) as numpy arrays,
    stats = {chi2_0, chi2_n, rms_0, rms_n}.  ``history=True`` also returns the (iters, 6 nk) pose updates.

    ``prior`` = (c (nl,3), rho0 (nl), invsig (nl) or scalar, valid (nl)); ``rotf`` = [(a, b, R (3,3), sigma), ...].
    Inputs may be numpy arrays or torch tensors; they are copied to the GPU once per call.  ``noalloc``: solve
    through ``xba_np_solve_noalloc`` (the library's static workspace, for problems within XBA_NOALLOC_MAX_*)."""
    import torch
    _lib.require_gpu()
    L = _lib.load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    poses = np.asarray(poses, np.float64).reshape(-1, 4, 4)
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    nk, nl = len(poses), len(pts)
    okf = np.asarray(okf).astype(np.int32).ravel()
    olm = np.asarray(olm).astype(np.int32).ravel()
    ouv = np.asarray(ouv, np.float64).reshape(-1, 2)
    N = len(okf)
    if len(olm) != N or len(ouv) != N:
        raise ValueError("okf, olm and ouv must have one row per observation")
    fm = _free_mask(free, nk)
    if fm.shape != (nk,):
        raise ValueError("free mask must have one entry per keyframe")

    def dt(x, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).to(dev).contiguous()

    d_poses, d_pts = dt(poses), dt(pts)
    d_okf, d_olm, d_ouv = dt(okf, torch.int32), dt(olm, torch.int32), dt(ouv)
    d_free = dt(fm.astype(np.uint8), torch.uint8)
    keep = [d_poses, d_pts, d_okf, d_olm, d_ouv, d_free]
    q = _lib.XbaNpProblem()
    q.nk, q.nl, q.N = nk, nl, N
    q.poses, q.pts, q.okf, q.olm, q.ouv = (d_poses.data_ptr(), d_pts.data_ptr(), d_okf.data_ptr(),
                                           d_olm.data_ptr(), d_ouv.data_ptr())
    q.free_mask = d_free.data_ptr()
    q.fx, q.cx, q.cy = _rig_intrinsics(rig)
    q.huber_k = HUBER_K
    q.lm_policy = int(_policy)
    if prior is not None:
        c_, rho0, invsig, valid = prior
        pc = dt(np.asarray(c_, np.float64).reshape(nl, 3))
        pr0 = dt(np.broadcast_to(np.asarray(rho0, np.float64), (nl,)))
        pis = dt(np.broadcast_to(np.asarray(invsig, np.float64), (nl,)))
        pv = dt(np.broadcast_to(np.asarray(valid, np.float64), (nl,)))
        keep += [pc, pr0, pis, pv]
        q.pr_c, q.pr_rho0, q.pr_invsig, q.pr_valid = pc.data_ptr(), pr0.data_ptr(), pis.data_ptr(), pv.data_ptr()
    if rotf:
        ab = dt(np.array([[f[0], f[1]] for f in rotf], np.int32), torch.int32)
        Ms = dt(np.stack([np.asarray(f[2], np.float64) for f in rotf]))
        sg = dt(np.array([float(f[3]) for f in rotf], np.float64))
        keep += [ab, Ms, sg]
        q.n_rot, q.rot_ab, q.rot_M, q.rot_sig = len(rotf), ab.data_ptr(), Ms.data_ptr(), sg.data_ptr()
    stats = torch.zeros(4, dtype=torch.float64, device=dev)
    dxh = torch.zeros((max(iters, 0), D_POSE * nk), dtype=torch.float64, device=dev) if history else None
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        entry = L.xba_np_solve_noalloc if noalloc else L.xba_np_solve
        rc = entry(C.byref(q), int(iters), float(lam0), _lib.XBA_FP64, C.c_void_p(stats.data_ptr()),
                   C.c_void_p(dxh.data_ptr()) if dxh is not None else None, C.c_void_p(stream))
    _lib.check(rc, "xba_np_solve_noalloc" if noalloc else "xba_np_solve")
    s = stats.cpu().numpy()
    st = {"chi2_0": float(s[0]), "chi2_n": float(s[1]), "rms_0": float(s[2]), "rms_n": float(s[3])}
    out = (d_poses.cpu().numpy(), d_pts.cpu().numpy(), st)
    del keep
    if history:
        return out + (dxh.cpu().numpy(),)
    return out


def _solve_np_id(poses, rho, mbear, hidx, hpose_ext, okf, olm, ouv, rig, free, iters, lam0=1e-4, prior=None,
                 rotf=None, rho_floor=1e-4, device=None, history=False, _policy=0):
    """backend/ba.py:1212 ``_solve_np_id``.  Returns (poses T_wc (nk,4,4), rho (nl,), mbear (nl,3), stats dict) as
    numpy arrays, stats = {chi2_0, chi2_n, rms_0, rms_n}; ``history=True`` also returns the (iters, 6 nk) pose
    updates.  ``prior`` = (m0 (nl, >=2), rho0, wbear, invsig, valid), arrays or scalars; ``rotf`` = [(a, b, R (3,3),
    sigma), ...]; ``hpose_ext`` (nl, 12) or (nl, 3, 4): host T_wc rows for landmarks with hidx < 0."""
    import torch
    _lib.require_gpu()
    L = _lib.load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    poses = np.asarray(poses, np.float64).reshape(-1, 4, 4)
    rho = np.asarray(rho, np.float64).ravel()
    nk, nl = len(poses), len(rho)
    mbear = np.asarray(mbear, np.float64).reshape(nl, 3)
    hidx = np.asarray(hidx).astype(np.int32).ravel()
    hext = np.asarray(hpose_ext, np.float64).reshape(-1, 12) if hpose_ext is not None else np.zeros((0, 12))
    if len(hext) != nl:
        hext = np.zeros((nl, 12)) if not (hidx < 0).any() else np.broadcast_to(hext, (nl, 12))
    okf = np.asarray(okf).astype(np.int32).ravel()
    olm = np.asarray(olm).astype(np.int32).ravel()
    ouv = np.asarray(ouv, np.float64).reshape(-1, 2)
    N = len(okf)
    if len(olm) != N or len(ouv) != N or len(hidx) != nl:
        raise ValueError("okf, olm and ouv need one row per observation and hidx one entry per landmark")
    fm = _free_mask(free, nk)
    if fm.shape != (nk,):
        raise ValueError("free mask must have one entry per keyframe")

    def dt(x, dtype=torch.float64):
        return torch.as_tensor(np.array(x, copy=True), dtype=dtype).to(dev).contiguous()

    d_poses, d_rho, d_mb = dt(poses), dt(rho), dt(mbear)
    d_hidx, d_hext = dt(hidx, torch.int32), dt(np.ascontiguousarray(hext))
    d_okf, d_olm, d_ouv = dt(okf, torch.int32), dt(olm, torch.int32), dt(ouv)
    d_free = dt(fm.astype(np.uint8), torch.uint8)
    keep = [d_poses, d_rho, d_mb, d_hidx, d_hext, d_okf, d_olm, d_ouv, d_free]
    q = _lib.XbaNpidProblem()
    q.nk, q.nl, q.N = nk, nl, N
    q.poses, q.rho, q.mbear, q.hidx, q.hpose_ext = (d_poses.data_ptr(), d_rho.data_ptr(), d_mb.data_ptr(),
                                                    d_hidx.data_ptr(), d_hext.data_ptr())
    q.okf, q.olm, q.ouv, q.free_mask = d_okf.data_ptr(), d_olm.data_ptr(), d_ouv.data_ptr(), d_free.data_ptr()
    q.fx, q.cx, q.cy = _rig_intrinsics(rig)
    q.huber_k, q.rho_floor = HUBER_K, float(rho_floor)
    q.lm_policy = int(_policy)
    if prior is not None:
        m0, rho0, wbear, invsig, valid = prior
        m0 = np.asarray(m0, np.float64).reshape(nl, -1)
        pm = np.zeros((nl, 3))
        pm[:, :min(3, m0.shape[1])] = m0[:, :3]
        arrs = [dt(np.broadcast_to(np.asarray(a, np.float64), (nl,))) for a in (rho0, wbear, invsig, valid)]
        d_pm = dt(pm)
        keep += [d_pm] + arrs
        q.pr_m0 = d_pm.data_ptr()
        q.pr_rho0, q.pr_wbear, q.pr_invsig, q.pr_valid = [a.data_ptr() for a in arrs]
    if rotf:
        ab = dt(np.array([[f[0], f[1]] for f in rotf], np.int32), torch.int32)
        Ms = dt(np.stack([np.asarray(f[2], np.float64) for f in rotf]))
        sg = dt(np.array([float(f[3]) for f in rotf], np.float64))
        keep += [ab, Ms, sg]
        q.n_rot, q.rot_ab, q.rot_M, q.rot_sig = len(rotf), ab.data_ptr(), Ms.data_ptr(), sg.data_ptr()
    stats = torch.zeros(4, dtype=torch.float64, device=dev)
    dxh = torch.zeros((max(iters, 0), D_POSE * nk), dtype=torch.float64, device=dev) if history else None
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        rc = L.xba_np_id_solve(C.byref(q), int(iters), float(lam0), _lib.XBA_FP64, C.c_void_p(stats.data_ptr()),
                               C.c_void_p(dxh.data_ptr()) if dxh is not None else None, C.c_void_p(stream))
    _lib.check(rc, "xba_np_id_solve")
    s = stats.cpu().numpy()
    st = {"chi2_0": float(s[0]), "chi2_n": float(s[1]), "rms_0": float(s[2]), "rms_n": float(s[3])}
    out = (d_poses.cpu().numpy(), d_rho.cpu().numpy(), d_mb.cpu().numpy(), st)
    del keep
    if history:
        return out + (dxh.cpu().numpy(),)
    return out


def build_pairs_device(olm_sorted, nl):
    """backend/ba.py:902-927 ``build_pairs`` on device tensors: ``olm_sorted`` (int32 CUDA tensor, ascending) ->
    (offs (nl + 1,), pa (npair,), pb (npair,)) int32 CUDA tensors, built by libxba ``xba_build_pairs`` on the tensor's
    stream.  Raises ValueError for a landmark index outside [0, nl) or a list that is not ascending."""
    import torch
    _lib.require_gpu()
    L = _lib.load()
    if olm_sorted.dtype != torch.int32 or not olm_sorted.is_cuda:
        raise ValueError("olm_sorted must be an int32 CUDA tensor")
    olm_sorted = olm_sorted.contiguous().reshape(-1)
    dev = olm_sorted.device
    nl = int(nl)
    offs = torch.empty(nl + 1, dtype=torch.int32, device=dev)
    n = C.c_int64(0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        optr = C.c_void_p(olm_sorted.data_ptr()) if olm_sorted.numel() else None
        rc = L.xba_build_pairs(optr, int(olm_sorted.numel()), nl, C.c_void_p(offs.data_ptr()), None, None, 0,
                               C.byref(n), C.c_void_p(stream))
        if rc == _lib.XBA_E_RANGE:
            raise ValueError("olm_sorted must be ascending with every index in [0, %d)" % nl)
        _lib.check(rc, "xba_build_pairs")
        pa = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
        pb = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
        rc = L.xba_build_pairs(optr, int(olm_sorted.numel()), nl, C.c_void_p(offs.data_ptr()),
                               C.c_void_p(pa.data_ptr()), C.c_void_p(pb.data_ptr()), int(pa.numel()), C.byref(n),
                               C.c_void_p(stream))
        _lib.check(rc, "xba_build_pairs")
    return offs, pa[:n.value], pb[:n.value]


def build_pairs(olm_sorted, nl, device=None):
    """backend/ba.py:902-927 ``build_pairs(olm_sorted, nl)`` -> (offs, pa, pb) as int32 numpy arrays: offs[l] the
    first observation of landmark l (nl + 1 entries), and every ordered observation pair (a, b) inside a landmark,
    landmark-major, a-major then b.  The lists are built on the GPU (``build_pairs_device``)."""
    import torch
    _lib.require_gpu()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    o = torch.as_tensor(np.ascontiguousarray(np.asarray(olm_sorted).ravel(), dtype=np.int32)).to(dev)
    offs, pa, pb = build_pairs_device(o, nl)
    return offs.cpu().numpy(), pa.cpu().numpy(), pb.cpu().numpy()


def _npair(olm, nl):
    """backend/ba.py:902-927 build_pairs' count: every ordered observation pair inside a landmark, sum c_l^2."""
    c = np.bincount(np.asarray(olm, np.int64).ravel(), minlength=nl) if len(olm) else np.zeros(0, np.int64)
    return int(np.sum(c.astype(np.int64) ** 2))


def _solve_gpu(poses, pts, okf, olm, ouv, rig, free, iters, lam0=1e-4, led=None, kf_id=None, prior=None, rotf=None,
               device=None):
    """backend/ba.py:930 ``_solve_gpu``: (poses T_wc, pts, stats) with stats = {chi2_0, chi2_n, rms_0, rms_n, npair,
    n_obs}; ``led`` / ``kf_id`` (the reference's timing ledger hooks) are accepted and unused."""
    p, q, st = _solve_np(poses, pts, okf, olm, ouv, rig, free, iters, lam0, prior=prior, rotf=rotf, device=device,
                         _policy=1)
    if int(iters) == 0:
        st["rms_0"] = None  # the reference reports no initial rms without an iteration
    st.update(npair=_npair(olm, len(np.asarray(pts).reshape(-1, 3))), n_obs=int(len(np.asarray(okf).ravel())))
    return p, q, st


def _solve_gpu_id(poses, rho, mbear, hidx, hpose_ext, okf, olm, ouv, rig, free, iters, lam0=1e-4, prior=None,
                  rotf=None, rho_floor=1e-4, device=None):
    """backend/ba.py:1345 ``_solve_gpu_id``: (poses T_wc, rho, mbear, stats) with the stats of ``_solve_gpu``."""
    p, r, m, st = _solve_np_id(poses, rho, mbear, hidx, hpose_ext, okf, olm, ouv, rig, free, iters, lam0,
                               prior=prior, rotf=rotf, rho_floor=rho_floor, device=device, _policy=1)
    if int(iters) == 0:
        st["rms_0"] = None
    st.update(npair=_npair(olm, len(np.asarray(rho).ravel())), n_obs=int(len(np.asarray(okf).ravel())))
    return p, r, m, st
