"""ctypes front end of the CPU FP64 oracle (oracle/xba_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module to
check or time the product path; the product package (xclean_amd/) never imports it.

Functions and the reference code they restate (paths relative to the reference's src/xclean/):
  window_solve          vio/window.py:889-982  (solve: build_system + prior + IMU + Cholesky + LM accept/reject)
  window_build          vio/window.py:291-463  (build_system)
  window_full_system    vio/window.py:865-886  (_full_system)
  window_final          vio/window.py:985-1008 (final_linearisation)
  visual_cost           vio/window.py:466-493
  obs_residual_jac      vio/window.py:125-263  (obs_residual + obs_jacobians_ws)
  imu_factor            vio/imu_factor.py:771-875 (factor_core)
  preintegrate          vio/imu_factor.py:392-704 (_step, preint_core)
  sqrt_info             vio/imu_factor.py:707-746 (sqrt_info_core)
  solve_np              backend/ba.py:713-869  (_solve_np, the drop-in entry point)
  so3_exp_nb/so3_log_nb vio/window.py:43-105; log_quat/jr_inv/exp_jr vio/imu_factor.py:195-330
  se3_exp/so3_project   core/lie.py:301-314, 228-242
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class OwProblem(C.Structure):
    _fields_ = [("P", C.c_int32), ("L", C.c_int32), ("M", C.c_int32),
                ("Rwb", f64p), ("pwb", f64p), ("Rbc", f64p), ("pbc", f64p), ("lmp", f64p),
                ("lm_host", i32p), ("lm_mask", u8p), ("rho_meas", f64p), ("rho_isig", f64p),
                ("ob_start", i32p), ("ob_pose", i32p), ("ob_cam", u8p), ("ob_valid", u8p),
                ("ob_z", f64p), ("ob_isig", f64p),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("B", C.c_double), ("huber_k", C.c_double),
                ("dims", i32p), ("vel", f64p), ("bgs", f64p), ("bas", f64p), ("g_w", f64p),
                ("n_imu", C.c_int32), ("imu_ij", i32p), ("imu_rec", f64p), ("imu_k", C.c_double),
                ("prior_nb", C.c_int32), ("prior_m", C.c_int32), ("prior_slot", i32p), ("prior_dim", i32p),
                ("prior_x0", f64p), ("prior_S", f64p), ("prior_r", f64p),
                ("n_pf", C.c_int32), ("pf_host", i32p), ("pf_z", f64p), ("pf_isg", C.c_double),
                ("pf_rmeas", f64p), ("pf_ris", f64p), ("pf_lmp", f64p), ("ob_w", f64p)]


class OnpProblem(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nl", C.c_int32), ("N", C.c_int32),
                ("poses", f64p), ("pts", f64p), ("okf", i32p), ("olm", i32p), ("ouv", f64p),
                ("fx", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("huber_k", C.c_double),
                ("free_mask", u8p), ("pr_c", f64p), ("pr_rho0", f64p), ("pr_invsig", f64p),
                ("pr_valid", f64p), ("n_rot", C.c_int32), ("rot_ab", i32p), ("rot_M", f64p), ("rot_sig", f64p),
                ("lm_policy", C.c_int32)]


class OnpidProblem(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nl", C.c_int32), ("N", C.c_int32),
                ("poses", f64p), ("rho", f64p), ("mbear", f64p), ("hidx", i32p), ("hpose_ext", f64p),
                ("okf", i32p), ("olm", i32p), ("ouv", f64p),
                ("fx", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("huber_k", C.c_double),
                ("rho_floor", C.c_double), ("free_mask", u8p),
                ("pr_m0", f64p), ("pr_rho0", f64p), ("pr_wbear", f64p), ("pr_invsig", f64p), ("pr_valid", f64p),
                ("n_rot", C.c_int32), ("rot_ab", i32p), ("rot_M", f64p), ("rot_sig", f64p),
                ("lm_policy", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libxba_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.ow_solve.argtypes = [C.POINTER(OwProblem), C.c_int, C.c_double, f64p, f64p, u8p]
        L.ow_build.argtypes = [C.POINTER(OwProblem), f64p, f64p, f64p, C.c_double, f64p, f64p, f64p, f64p, f64p,
                               u8p, f64p]
        L.ow_build.restype = C.c_double
        L.ow_full_system.argtypes = [C.POINTER(OwProblem), f64p, f64p, f64p, f64p]
        L.ow_full_system.restype = C.c_double
        L.ow_visual_cost.argtypes = [C.POINTER(OwProblem), f64p, f64p, f64p]
        L.ow_visual_cost.restype = C.c_double
        L.ow_final_linearisation.argtypes = [C.POINTER(OwProblem), f64p, f64p, f64p, u8p, f64p, f64p, f64p, u8p]
        L.o_obs_residual_jac.argtypes = [C.POINTER(OwProblem), C.c_int, C.c_int, f64p, f64p, f64p, f64p]
        L.o_obs_residual_jac.restype = C.c_int
        L.o_imu_factor.argtypes = [f64p] * 12 + [C.c_int, f64p, f64p, f64p, f64p]
        L.onp_solve.argtypes = [C.POINTER(OnpProblem), C.c_int, C.c_double, f64p, f64p, u8p]
        L.onpid_solve.argtypes = [C.POINTER(OnpidProblem), C.c_int, C.c_double, f64p, f64p, u8p]
        L.o_id_blocks.argtypes = [f64p, f64p, C.c_double, f64p, f64p, C.c_int, C.c_double, f64p, C.c_double, f64p]
        for name in ("o_so3_exp_nb", "o_so3_log_nb", "o_log_quat", "o_jr_inv", "o_se3_exp", "o_so3_project"):
            getattr(L, name).argtypes = [f64p, f64p]
        L.o_exp_jr.argtypes = [f64p, f64p, f64p]
        L.o_chol_solve.argtypes = [f64p, C.c_int, f64p, f64p]
        L.o_preint.argtypes = [f64p, f64p, f64p, C.c_int, C.c_double, C.c_double, C.c_double, f64p, f64p,
                               C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               f64p, f64p, f64p, f64p, f64p, f64p]
        L.o_sqrt_info.argtypes = [f64p, C.c_double, C.c_double, C.c_double, f64p, f64p]
        L.o_sqrt_info.restype = C.c_int
        L.o_chol_solve.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class _Packed:
    """Keeps the contiguous arrays alive while the struct points at them."""

    def __init__(self, pb):
        from xclean_amd.problem import pack_imu
        from xclean_amd.synth import huber_k_dof
        self.pb = pb
        P, L, M = pb.Rwb.shape[0], pb.lmp.shape[0], len(pb.ob_pose)
        c = np.ascontiguousarray
        # state arrays are passed by reference so the oracle updates them in place (window.solve mutates pb)
        for k in ("Rwb", "pwb", "lmp", "vel", "bgs", "bas"):
            v = getattr(pb, k)
            if v is not None:
                if not (v.flags.c_contiguous and v.dtype == np.float64):
                    v = c(v, np.float64)
                    setattr(pb, k, v)
        self.Rbc = c(pb.Rbc, np.float64)
        self.pbc = c(pb.pbc, np.float64)
        self.lm_host = c(pb.lm_host, np.int32)
        self.lm_mask = c(pb.lm_mask, np.uint8)
        self.rho_meas = c(pb.rho_meas, np.float64)
        self.rho_isig = c(pb.rho_isig, np.float64)
        self.ob_start = c(pb.ob_start, np.int32)
        self.ob_pose = c(pb.ob_pose, np.int32)
        self.ob_cam = c(pb.ob_cam, np.uint8)
        self.ob_valid = c(pb.ob_valid, np.uint8)
        self.ob_z = c(pb.ob_z, np.float64)
        self.ob_isig = c(pb.ob_isig, np.float64)
        self.dims = c(pb.dims, np.int32) if pb.dims is not None else None
        self.g_w = c(pb.g_w, np.float64) if pb.g_w is not None else None
        self.imu_ij, self.imu_rec = pack_imu(pb.imu)
        s = OwProblem()
        s.P, s.L, s.M = P, L, M
        s.Rwb, s.pwb, s.lmp = _p(pb.Rwb, f64p), _p(pb.pwb, f64p), _p(pb.lmp, f64p)
        s.Rbc, s.pbc = _p(self.Rbc, f64p), _p(self.pbc, f64p)
        s.lm_host, s.lm_mask = _p(self.lm_host, i32p), _p(self.lm_mask, u8p)
        s.rho_meas, s.rho_isig = _p(self.rho_meas, f64p), _p(self.rho_isig, f64p)
        s.ob_start, s.ob_pose = _p(self.ob_start, i32p), _p(self.ob_pose, i32p)
        s.ob_cam, s.ob_valid = _p(self.ob_cam, u8p), _p(self.ob_valid, u8p)
        s.ob_z, s.ob_isig = _p(self.ob_z, f64p), _p(self.ob_isig, f64p)
        ob_w = getattr(pb, "ob_w", None)
        self.ob_w = c(ob_w, np.float64) if ob_w is not None else None
        s.ob_w = _p(self.ob_w, f64p)
        s.fx, s.fy, s.cx, s.cy, s.B, s.huber_k = pb.fx, pb.fy, pb.cx, pb.cy, pb.B, pb.huber_k
        s.dims = _p(self.dims, i32p)
        s.vel, s.bgs, s.bas = _p(pb.vel, f64p), _p(pb.bgs, f64p), _p(pb.bas, f64p)
        s.g_w = _p(self.g_w, f64p)
        s.n_imu = len(self.imu_ij)
        s.imu_ij, s.imu_rec = _p(self.imu_ij, i32p), _p(self.imu_rec, f64p)
        s.imu_k = huber_k_dof(15) if s.n_imu else 0.0
        pr = pb.prior
        if pr is not None:
            self.prior_slot = c(pr.slots(), np.int32)
            self.prior_dim = c(pr.dims, np.int32)
            self.prior_x0 = c(pr.x0_packed())
            self.prior_S = c(pr.S, np.float64)
            self.prior_r = c(pr.r, np.float64)
            s.prior_nb, s.prior_m = len(pr.sids), pr.S.shape[0]
            s.prior_slot, s.prior_dim = _p(self.prior_slot, i32p), _p(self.prior_dim, i32p)
            s.prior_x0, s.prior_S, s.prior_r = _p(self.prior_x0, f64p), _p(self.prior_S, f64p), _p(self.prior_r,
                                                                                                   f64p)
        pf = pb.pf
        if pf is not None and len(pf.host):
            self.pf_host = c(pf.host, np.int32)
            self.pf_z = c(pf.z, np.float64)
            self.pf_rmeas = c(pf.rmeas)
            self.pf_ris = c(pf.ris)
            s.n_pf = len(pf.host)
            s.pf_host, s.pf_z, s.pf_isg = _p(self.pf_host, i32p), _p(self.pf_z, f64p), pf.isg
            s.pf_rmeas, s.pf_ris, s.pf_lmp = _p(self.pf_rmeas, f64p), _p(self.pf_ris, f64p), _p(pf.lmp, f64p)
        self.s = s


def window_solve(pb, iters=5, lam0=1e-4, history=False):
    """vio/window.py:889-982 on ``pb`` (mutated in place).  Returns the stats dict; with ``history`` also the
    per-iteration full update vectors (iters, n) and accept flags."""
    pk = _Packed(pb)
    n = pb.n_state()
    stats = np.zeros(5)
    dxh = np.zeros((iters, n)) if history else None
    acc = np.zeros(iters, np.uint8) if history else None
    lib().ow_solve(C.byref(pk.s), iters, lam0, _p(stats, f64p), _p(dxh, f64p), _p(acc, u8p))
    st = {"cost0": stats[0], "cost": stats[1], "accepted": int(stats[2]), "iters": int(stats[3]), "lam": stats[4]}
    if history:
        return st, dxh, acc.astype(bool)
    return st


def window_build(pb, lam=0.0):
    """build_system at the problem's state: (cost, H6, b6, Hll_inv, bl, Hlp, lm_ok, ob_chi2)."""
    pk = _Packed(pb)
    P, L, M = pb.P, pb.L, pb.M
    H6 = np.zeros((6 * P, 6 * P))
    b6 = np.zeros(6 * P)
    Hi = np.zeros((max(L, 1), 3, 3))
    bl = np.zeros((max(L, 1), 3))
    Hlp = np.zeros((max(L, 1), P, 3, 6))
    ok = np.zeros(max(L, 1), np.uint8)
    chi = np.zeros(max(M, 1))
    c = lib().ow_build(C.byref(pk.s), _p(pb.Rwb, f64p), _p(pb.pwb, f64p), _p(pb.lmp, f64p), lam, _p(H6, f64p),
                       _p(b6, f64p), _p(Hi, f64p), _p(bl, f64p), _p(Hlp, f64p), _p(ok, u8p), _p(chi, f64p))
    return c, H6, b6, Hi[:L], bl[:L], Hlp[:L], ok[:L].astype(bool), chi[:M]


def window_full_system(pb, H6, b6):
    pk = _Packed(pb)
    n = pb.n_state()
    H = np.zeros((n, n))
    b = np.zeros(n)
    c = lib().ow_full_system(C.byref(pk.s), _p(np.ascontiguousarray(H6), f64p), _p(np.ascontiguousarray(b6), f64p),
                             _p(H, f64p), _p(b, f64p))
    return H, b, c


def window_final(pb):
    """final_linearisation: (H, b, ob_chi2, lm_ok, Hll_inv, imu_chi2)."""
    pk = _Packed(pb)
    n, L, M = pb.n_state(), pb.L, pb.M
    H = np.zeros((n, n))
    b = np.zeros(n)
    chi = np.zeros(max(M, 1))
    ok = np.zeros(max(L, 1), np.uint8)
    Hi = np.zeros((max(L, 1), 3, 3))
    ni = len(pb.imu or [])
    ichi = np.zeros(max(ni, 1))
    npf = len(pb.pf.host) if pb.pf is not None else 0
    pfH = np.zeros((max(npf, 1), 3, 3))
    pfok = np.zeros(max(npf, 1), np.uint8)
    lib().ow_final_linearisation(C.byref(pk.s), _p(H, f64p), _p(b, f64p), _p(chi, f64p), _p(ok, u8p), _p(Hi, f64p),
                                 _p(ichi, f64p), _p(pfH, f64p), _p(pfok, u8p))
    if npf:
        pb.pf.Hinv[:] = pfH[:npf]
        pb.pf.ok[:] = pfok[:npf].astype(bool)
    return H, b, chi[:M], ok[:L].astype(bool), Hi[:L], list(ichi[:ni])


def visual_cost(pb, Rwb, pwb, lmp):
    pk = _Packed(pb)
    c = np.ascontiguousarray
    Rwb, pwb, lmp = c(Rwb, np.float64), c(pwb, np.float64), c(lmp, np.float64)
    return lib().ow_visual_cost(C.byref(pk.s), _p(Rwb, f64p), _p(pwb, f64p), _p(lmp, f64p))


def obs_residual_jac(pb, l, k):
    """(e (nd,), Jl (nd,3), Jt (nd,6), Jh (nd,6)) of observation k of landmark l, unwhitened."""
    pk = _Packed(pb)
    e = np.zeros(2)
    Jl = np.zeros((2, 3))
    Jt = np.zeros((2, 6))
    Jh = np.zeros((2, 6))
    nd = lib().o_obs_residual_jac(C.byref(pk.s), l, k, _p(e, f64p), _p(Jl, f64p), _p(Jt, f64p), _p(Jh, f64p))
    return e[:nd], Jl[:nd], Jt[:nd], Jh[:nd]


def imu_factor(si, sj, pre, g, W=None, whiten=True):
    """factor_core: (r (15,), Ji, Jj (15,15), Jg (15,3)).  si/sj = (R, p, v, bg, ba)."""
    from xclean_amd.problem import pack_imu
    _ij, rec = pack_imu([(0, 1, pre, np.eye(15) if W is None else W)])
    args = [np.ascontiguousarray(x, np.float64) for x in (*si, *sj, g)]
    r = np.zeros(15)
    Ji = np.zeros((15, 15))
    Jj = np.zeros((15, 15))
    Jg = np.zeros((15, 3))
    lib().o_imu_factor(*[_p(a, f64p) for a in args], _p(rec[0], f64p), int(bool(whiten)), _p(r, f64p),
                       _p(Ji, f64p), _p(Jj, f64p), _p(Jg, f64p))
    return r, Ji, Jj, Jg


def solve_np(poses, pts, okf, olm, ouv, rig, free, iters, lam0=1e-4, prior=None, rotf=None, history=False,
             policy="np"):
    """backend/ba.py:713-869 (_solve_np): returns (poses T_wc, pts, stats) like the reference.  policy="gpu": the LM
    policy of the reference's GPU driver _solve_gpu (ba.py:930-1089), computed in FP64."""
    nk, nl, N = len(poses), len(pts), len(okf)
    poses = np.ascontiguousarray(poses, np.float64).copy()
    pts = np.ascontiguousarray(pts, np.float64).copy()
    okf = np.ascontiguousarray(okf, np.int32)
    olm = np.ascontiguousarray(olm, np.int32)
    ouv = np.ascontiguousarray(ouv, np.float64)
    if np.ndim(free) == 0:
        fm = np.ones(nk, np.uint8)
        fm[int(free)] = 0
    else:
        fm = np.asarray(free, bool).astype(np.uint8)
    s = OnpProblem()
    s.nk, s.nl, s.N = nk, nl, N
    s.poses, s.pts, s.okf, s.olm, s.ouv = _p(poses, f64p), _p(pts, f64p), _p(okf, i32p), _p(olm, i32p), _p(ouv, f64p)
    s.fx, s.cx, s.cy, s.huber_k = rig.fx_rect, rig.cx, rig.cy, 1.345
    s.free_mask = _p(fm, u8p)
    s.lm_policy = 1 if policy == "gpu" else 0
    keep = []
    if prior is not None:
        c_, rho0, invsig, valid = prior
        pc = np.ascontiguousarray(c_, np.float64)
        pr0 = np.ascontiguousarray(np.broadcast_to(rho0, (nl,)), np.float64)
        pv = np.ascontiguousarray(np.broadcast_to(valid, (nl,)), np.float64)
        pis = np.ascontiguousarray(np.broadcast_to(invsig, (nl,)), np.float64)
        keep += [pc, pr0, pv, pis]
        s.pr_c, s.pr_rho0, s.pr_invsig, s.pr_valid = _p(pc, f64p), _p(pr0, f64p), _p(pis, f64p), _p(pv, f64p)
    if rotf:
        ab = np.ascontiguousarray([[f[0], f[1]] for f in rotf], np.int32)
        Ms = np.ascontiguousarray(np.stack([f[2] for f in rotf]), np.float64)
        sg = np.ascontiguousarray([f[3] for f in rotf], np.float64)
        keep += [ab, Ms, sg]
        s.n_rot, s.rot_ab, s.rot_M, s.rot_sig = len(rotf), _p(ab, i32p), _p(Ms, f64p), _p(sg, f64p)
    stats = np.zeros(4)
    dxh = np.zeros((iters, 6 * nk)) if history else None
    acc = np.zeros(iters, np.uint8) if history else None
    lib().onp_solve(C.byref(s), iters, lam0, _p(stats, f64p), _p(dxh, f64p), _p(acc, u8p))
    st = {"chi2_0": stats[0], "chi2_n": stats[1], "rms_0": stats[2], "rms_n": stats[3]}
    if history:
        return poses, pts, st, dxh, acc.astype(bool)
    return poses, pts, st


def solve_np_id(poses, rho, mbear, hidx, hpose_ext, okf, olm, ouv, rig, free, iters, lam0=1e-4, prior=None,
                rotf=None, rho_floor=1e-4, history=False, policy="np"):
    """backend/ba.py:1212-1342 (_solve_np_id): returns (poses T_wc, rho, mbear, stats) like the reference
    (policy="gpu": the LM policy of the GPU driver _solve_gpu_id, ba.py:1345-1497, computed in FP64).
    prior = (m0 (nl, >=2), rho0, wbear, invsig, valid) with scalars broadcast; rotf = [(a, b, R, sigma), ...]."""
    nk, nl, N = len(poses), len(rho), len(okf)
    poses = np.ascontiguousarray(poses, np.float64).copy()
    rho = np.ascontiguousarray(rho, np.float64).copy()
    mbear = np.ascontiguousarray(mbear, np.float64).reshape(nl, 3).copy()
    hidx = np.ascontiguousarray(hidx, np.int32)
    hext = np.ascontiguousarray(np.broadcast_to(np.asarray(hpose_ext, np.float64).reshape(-1, 12), (nl, 12)))
    okf = np.ascontiguousarray(okf, np.int32)
    olm = np.ascontiguousarray(olm, np.int32)
    ouv = np.ascontiguousarray(ouv, np.float64)
    if np.ndim(free) == 0:
        fm = np.ones(nk, np.uint8)
        fm[int(free)] = 0
    else:
        fm = np.asarray(free, bool).astype(np.uint8)
    s = OnpidProblem()
    s.nk, s.nl, s.N = nk, nl, N
    s.poses, s.rho, s.mbear, s.hidx, s.hpose_ext = (_p(poses, f64p), _p(rho, f64p), _p(mbear, f64p),
                                                   _p(hidx, i32p), _p(hext, f64p))
    s.okf, s.olm, s.ouv = _p(okf, i32p), _p(olm, i32p), _p(ouv, f64p)
    s.fx, s.cx, s.cy, s.huber_k, s.rho_floor = rig.fx_rect, rig.cx, rig.cy, 1.345, rho_floor
    s.free_mask = _p(fm, u8p)
    s.lm_policy = 1 if policy == "gpu" else 0
    keep = [hext]
    if prior is not None:
        m0, rho0, wbear, invsig, valid = prior
        m0 = np.asarray(m0, np.float64).reshape(nl, -1)
        pm = np.zeros((nl, 3))
        pm[:, :min(3, m0.shape[1])] = m0[:, :3]
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (nl,))) for a in (rho0, wbear, invsig,
                                                                                              valid)]
        keep += [pm] + arrs
        s.pr_m0 = _p(pm, f64p)
        s.pr_rho0, s.pr_wbear, s.pr_invsig, s.pr_valid = [_p(a, f64p) for a in arrs]
    if rotf:
        ab = np.ascontiguousarray([[f[0], f[1]] for f in rotf], np.int32)
        Ms = np.ascontiguousarray(np.stack([f[2] for f in rotf]), np.float64)
        sg = np.ascontiguousarray([f[3] for f in rotf], np.float64)
        keep += [ab, Ms, sg]
        s.n_rot, s.rot_ab, s.rot_M, s.rot_sig = len(rotf), _p(ab, i32p), _p(Ms, f64p), _p(sg, f64p)
    stats = np.zeros(4)
    dxh = np.zeros((iters, 6 * nk)) if history else None
    acc = np.zeros(iters, np.uint8) if history else None
    lib().onpid_solve(C.byref(s), iters, lam0, _p(stats, f64p), _p(dxh, f64p), _p(acc, u8p))
    st = {"chi2_0": stats[0], "chi2_n": stats[1], "rms_0": stats[2], "rms_n": stats[3]}
    if history:
        return poses, rho, mbear, st, dxh, acc.astype(bool)
    return poses, rho, mbear, st


def id_blocks(Tcw_k, Rh, rho, Pc, Ph, host_in, fx, r, sw=1.0):
    """backend/ba.py:1138-1209 for one observation: (Jobs (2,6), Jhost (2,6), Jlm (2,3), rw (2,))."""
    J = np.zeros(32)
    a = [np.ascontiguousarray(x, np.float64) for x in (Tcw_k, Rh, Pc, Ph, r)]
    lib().o_id_blocks(_p(a[0], f64p), _p(a[1], f64p), float(rho), _p(a[2], f64p), _p(a[3], f64p), int(host_in),
                      float(fx), _p(a[4], f64p), float(sw), _p(J, f64p))
    return J[:12].reshape(2, 6), J[12:24].reshape(2, 6), J[24:30].reshape(2, 3), J[30:32]


def _unary(name, x, shape_out):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(shape_out)
    getattr(lib(), name)(_p(x, f64p), _p(out, f64p))
    return out


def so3_exp_nb(w):
    return _unary("o_so3_exp_nb", w, (3, 3))


def so3_log_nb(R):
    return _unary("o_so3_log_nb", R, (3,))


def log_quat(R):
    return _unary("o_log_quat", R, (3,))


def jr_inv(v):
    return _unary("o_jr_inv", v, (3, 3))


def se3_exp(xi):
    return _unary("o_se3_exp", xi, (4, 4))


def so3_project(R):
    return _unary("o_so3_project", R, (3, 3))


def exp_jr(v):
    v = np.ascontiguousarray(v, np.float64)
    R = np.zeros((3, 3))
    Jr = np.zeros((3, 3))
    lib().o_exp_jr(_p(v, f64p), _p(R, f64p), _p(Jr, f64p))
    return R, Jr


def chol_solve(A, b):
    A = np.ascontiguousarray(A, np.float64).copy()
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    ok = lib().o_chol_solve(_p(A, f64p), len(b), _p(b, f64p), _p(x, f64p))
    return x if ok else None


PI_STATUS, PI_NSUB, PI_NINNER, PI_EXTRAP, PI_TEND, PI_GAPMAX, PI_NREL, PI_TNOM, PI_SMAX = range(9)


def preintegrate(t, gyr, acc, t0, t1, horizon, bg, ba, sg2, sa2, swg2, swa2, tnom):
    """vio/imu_factor.py:545-704 (preint_core): dict(dR, dv, dp, J (9,6), C (15,15), info (9,)); horizon=None means
    every sample is released."""
    t = np.ascontiguousarray(t, np.float64)
    gyr = np.ascontiguousarray(gyr, np.float64).reshape(-1, 3)
    acc = np.ascontiguousarray(acc, np.float64).reshape(-1, 3)
    bg = np.ascontiguousarray(bg, np.float64)
    ba = np.ascontiguousarray(ba, np.float64)
    out = dict(dR=np.zeros((3, 3)), dv=np.zeros(3), dp=np.zeros(3), J=np.zeros((9, 6)), C=np.zeros((15, 15)),
               info=np.zeros(9))
    hz = np.inf if horizon is None else float(horizon)
    lib().o_preint(_p(t, f64p), _p(gyr, f64p), _p(acc, f64p), len(t), float(t0), float(t1), hz, _p(bg, f64p),
                   _p(ba, f64p), float(sg2), float(sa2), float(swg2), float(swa2), float(tnom),
                   *[_p(out[k], f64p) for k in ("dR", "dv", "dp", "J", "C", "info")])
    return out


def sqrt_info(Cm, kappa, sbg2, sba2, J):
    """vio/imu_factor.py:707-746 (sqrt_info_core): W (15,15) lower, or None when not positive definite."""
    Cm = np.ascontiguousarray(Cm, np.float64)
    J = np.ascontiguousarray(J, np.float64)
    W = np.zeros((15, 15))
    ok = lib().o_sqrt_info(_p(Cm, f64p), float(kappa), float(sbg2), float(sba2), _p(J, f64p), _p(W, f64p))
    return W if ok else None


# ---- vio/marg.py, restated in numpy (the marginalisation algebra of the square-root prior) ------------------------
MARG_FLOOR = 1e-12  # vio/marg.py:32


def _eig_floor(A):
    """Eigen-decomposition of the symmetrised A and the mask of eigenvalues above the floor (vio/marg.py:75-79)."""
    lam, V = np.linalg.eigh(0.5 * (A + A.T))
    top = float(lam.max()) if len(lam) else 0.0
    return lam, V, lam > MARG_FLOOR * max(top, 1e-300)


def marg_schur(H, b, keep, marg):
    """vio/marg.py:165-177: the Schur complement of the kept indices, H_mm inverted over its eigenvalues above the
    floor."""
    keep, marg = np.asarray(keep, np.int64), np.asarray(marg, np.int64)
    lam, V, ok = _eig_floor(H[np.ix_(marg, marg)])
    Hinv = (V[:, ok] / lam[ok]) @ V[:, ok].T
    Hkm = H[np.ix_(keep, marg)]
    H2 = H[np.ix_(keep, keep)] - Hkm @ Hinv @ Hkm.T
    b2 = b[keep] - Hkm @ Hinv @ b[marg]
    return 0.5 * (H2 + H2.T), b2


def prior_from_info(H, b):
    """vio/marg.py:73-84: S (rank x n, rows in ascending eigenvalue), r."""
    lam, V, ok = _eig_floor(H)
    sl = np.sqrt(lam[ok])
    return sl[:, None] * V[:, ok].T, (V[:, ok].T @ b) / sl


def boxminus(cur, x0, dim):
    """vio/marg.py:47-56: [Log(R0^T R), p - p0 (, v - v0, bg - bg0, ba - ba0)]; states (R, p, v, bg, ba)."""
    out = np.zeros(dim)
    out[0:3] = so3_log_nb(np.asarray(x0[0]).T @ np.asarray(cur[0]))
    out[3:6] = np.asarray(cur[1]) - np.asarray(x0[1])
    if dim == 15:
        for k, q in enumerate((2, 3, 4)):
            out[6 + 3 * k:9 + 3 * k] = np.asarray(cur[q]) - np.asarray(x0[q])
    return out


def combine(old, sids_new, dims_new, states_cur, H_cur, b_cur):
    """vio/marg.py:212-242.  old: None or (sids, dims, x0 list); states as (R, p, v, bg, ba) with zeros for absent
    parts.  Returns (x0 list, S, r): blocks of the old prior keep their x0 (a pose block that grew keeps its pose x0
    and takes the current v, bg, ba); b is re-expressed at x0 before the square root."""
    x0, d, o = [], np.zeros(int(sum(dims_new))), 0
    for s, dim, st in zip(sids_new, dims_new, states_cur):
        if old is not None and s in old[0]:
            j = old[0].index(s)
            xo = old[2][j]
            xb = xo if (old[1][j] == 15 or dim == 6) else (xo[0], xo[1], st[2], st[3], st[4])
            x0.append(xb)
            d[o:o + dim] = boxminus(st, xb, dim)
        else:
            x0.append(st)
        o += dim
    S, r = prior_from_info(H_cur, b_cur - H_cur @ d)
    return x0, S, r
