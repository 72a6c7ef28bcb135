"""ctypes binding of libxba.so (include/xba.h).

The shared library is built in-tree by ``xclean_amd.build`` (``__graft_entry__.build()``).  There is no fallback:
if the library is missing, or no GPU is visible, the solver raises instead of computing anything elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxba.so")
_LIB = None

XBA_FP64, XBA_MIXED = 0, 1
XBA_WPAR = 24
XBA_IMU_REC = 301
XBA_NSTATS = 10  # columns of xba_stats (include/xba.h)
XBA_E_ARG, XBA_E_HIP, XBA_E_TOO_MANY_OBS, XBA_E_RANGE, XBA_E_SIZE, XBA_E_STATE = -1, -2, -3, -4, -5, -6

vp = C.c_void_p
i32p = C.POINTER(C.c_int32)


class XbaBatch(C.Structure):
    _fields_ = [("nwin", C.c_int32),
                ("win_slot0", i32p), ("win_lm0", i32p), ("win_imu0", i32p), ("win_pblk0", i32p),
                ("win_prow0", i32p), ("slot_dim", i32p), ("ob_start", i32p), ("lm_host", i32p), ("ob_pose", i32p),
                ("imu_ij", i32p), ("prior_slot", i32p), ("prior_dim", i32p),
                ("Rwb", vp), ("pwb", vp), ("vel", vp), ("bgs", vp), ("bas", vp), ("lmp", vp),
                ("lm_mask", vp), ("rho_meas", vp), ("rho_isig", vp), ("ob_z", vp), ("ob_isig", vp),
                ("ob_cam", vp), ("ob_valid", vp), ("win_par", vp), ("imu_rec", vp), ("prior_x0", vp),
                ("prior_S", vp), ("prior_r", vp), ("ob_w", vp)]


class XbaOptions(C.Structure):
    _fields_ = [("precision", C.c_int32), ("lm_per_chunk", C.c_int32), ("slot_tile", C.c_int32),
                ("grid", C.c_int32), ("use_graph", C.c_int32), ("want_chi2", C.c_int32),
                ("deterministic", C.c_int32), ("big_n", C.c_int32), ("big_fp64", C.c_int32)]


class XbaNpProblem(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nl", C.c_int32), ("N", C.c_int32),
                ("poses", vp), ("pts", vp), ("okf", vp), ("olm", vp), ("ouv", vp), ("free_mask", vp),
                ("fx", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("huber_k", C.c_double),
                ("pr_c", vp), ("pr_rho0", vp), ("pr_invsig", vp), ("pr_valid", vp),
                ("n_rot", C.c_int32), ("rot_ab", vp), ("rot_M", vp), ("rot_sig", vp), ("lm_policy", C.c_int32)]


class XbaNpidProblem(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nl", C.c_int32), ("N", C.c_int32),
                ("poses", vp), ("rho", vp), ("mbear", vp), ("hidx", vp), ("hpose_ext", vp),
                ("okf", vp), ("olm", vp), ("ouv", vp), ("free_mask", vp),
                ("fx", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("huber_k", C.c_double),
                ("rho_floor", C.c_double),
                ("pr_m0", vp), ("pr_rho0", vp), ("pr_wbear", vp), ("pr_invsig", vp), ("pr_valid", vp),
                ("n_rot", C.c_int32), ("rot_ab", vp), ("rot_M", vp), ("rot_sig", vp), ("lm_policy", C.c_int32)]


class XbaGbaProblem(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nl", C.c_int32), ("N", C.c_int32), ("poses", vp), ("rho", vp), ("mbear", vp),
                ("host", vp), ("okf", vp), ("olm", vp), ("ouv", vp), ("odisp", vp), ("vt_w", vp), ("free_mask", vp),
                ("pr_m0", vp), ("pr_rho0", vp), ("pr_wbear", vp), ("pr_invsig", vp), ("pr_valid", vp),
                ("fx", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("fB", C.c_double),
                ("rho_sigma", C.c_double), ("rho_floor", C.c_double),
                ("n_rot", C.c_int32), ("rot_ab", vp), ("rot_M", vp), ("rot_sig", vp), ("rot_huber", C.c_double),
                ("n_weld", C.c_int32), ("weld_ab", vp), ("weld_Z", vp), ("weld_w", vp)]


class XbaGbaOptions(C.Structure):
    _fields_ = [("iters", C.c_int32), ("passes", C.c_int32), ("maxiter", C.c_int32), ("seg_len", C.c_int32),
                ("profile", C.c_int32), ("lam0", C.c_double), ("reject_sigma", C.c_double), ("time_budget_s", C.c_double),
                ("rtol", C.c_double), ("rtol_direct", C.c_double), ("rtol_direct_floor", C.c_double),
                ("want_cov", C.c_int32), ("cov_ids", vp), ("cov", vp)]


XBA_GBA_NSTATS = 33
XBA_GBA_HIST = 10
XBA_GBA_DIRECT_MAX = 30000
XBA_GBA_MAX_NK = 65536


class XbaError(RuntimeError):
    pass


# every symbol include/xba.h declares
EXPORTS = ("xba_create", "xba_destroy", "xba_upload", "xba_iterate", "xba_download", "xba_stats", "xba_last_dx",
           "xba_linearize", "xba_sizes", "xba_profile_iteration", "xba_plan_info", "xba_gn_step", "xba_backward",
           "xba_set_assembly_timing", "xba_assembly_times",
           "xba_np_solve", "xba_np_solve_noalloc", "xba_np_id_solve", "xba_preintegrate", "xba_sqrt_information",
           "xba_marg_schur", "xba_prior_from_info", "xba_build_pairs", "xba_gba_solve",
           "xba_copy_kernel", "xba_strerror", "xba_version")


def load():
    """Load libxba.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise XbaError("libxba.so is not built (%s): run __graft_entry__.build() or python -m xclean_amd.build"
                       % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.xba_create.argtypes = [C.POINTER(vp), C.POINTER(XbaBatch), C.POINTER(XbaOptions)]
    L.xba_destroy.argtypes = [vp]
    L.xba_upload.argtypes = [vp, C.c_double, vp]
    L.xba_iterate.argtypes = [vp, C.c_int32, vp]
    L.xba_download.argtypes = [vp, vp]
    L.xba_stats.argtypes = [vp, vp, vp]
    L.xba_last_dx.argtypes = [vp, vp, vp]
    L.xba_linearize.argtypes = [vp, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp]
    L.xba_sizes.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.xba_profile_iteration.argtypes = [vp, vp, vp]
    L.xba_set_assembly_timing.argtypes = [vp, C.c_int32]
    L.xba_assembly_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.xba_plan_info.argtypes = [vp, vp]
    L.xba_gn_step.argtypes = [vp, vp, vp]
    L.xba_backward.argtypes = [vp, vp, vp, vp, vp]
    L.xba_np_solve.argtypes = [C.POINTER(XbaNpProblem), C.c_int32, C.c_double, C.c_int32, vp, vp, vp]
    L.xba_np_solve_noalloc.argtypes = [C.POINTER(XbaNpProblem), C.c_int32, C.c_double, C.c_int32, vp, vp, vp]
    L.xba_np_id_solve.argtypes = [C.POINTER(XbaNpidProblem), C.c_int32, C.c_double, C.c_int32, vp, vp, vp]
    L.xba_copy_kernel.argtypes = [vp, vp, C.c_int64, vp]
    L.xba_preintegrate.argtypes = [vp, vp]
    L.xba_sqrt_information.argtypes = [C.c_int32, vp, vp, vp, vp, vp, vp]
    L.xba_marg_schur.argtypes = [vp, vp]
    L.xba_prior_from_info.argtypes = [vp, vp]
    L.xba_gba_solve.argtypes = [C.POINTER(XbaGbaProblem), C.POINTER(XbaGbaOptions), vp, vp, C.c_int32, vp]
    L.xba_build_pairs.argtypes = [vp, C.c_int64, C.c_int32, vp, vp, vp, C.c_int64, C.POINTER(C.c_int64), vp]
    L.xba_strerror.restype = C.c_char_p
    L.xba_strerror.argtypes = [C.c_int]
    L.xba_version.restype = C.c_char_p
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is C.c_int:
            pass
    _LIB = L
    return L


def check(code, what):
    if code != 0:
        msg = load().xba_strerror(int(code)).decode()
        raise XbaError("%s failed: %s (%d)" % (what, msg, code))


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise XbaError("xclean_amd needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path")
    load()
