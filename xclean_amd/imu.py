"""IMU preintegration on the GPU (vio/imu_factor.py's `preintegrate`, `sqrt_information`, `ImuNoise`, `Preint`),
many intervals per launch through `xba_preintegrate` (include/xba.h).

    from xclean_amd.imu import ImuNoise, preintegrate, preintegrate_many
    pre = preintegrate(t, gyro, acc, t_i, t_j, horizon, bg_lin, ba_lin, noise)        # one interval, as the reference
    pres = preintegrate_many([(t, gyro, acc)], [(0, t_i, t_j, horizon, bg, ba), ...], noise)   # a batch

The arithmetic (midpoint rule, exact bias Jacobians, the 15x15 covariance with the sample-support scale, the causal
hold, the refusal codes) runs in FP64 in `k_preint` (csrc/xba_preint.hip).  Host code here resolves the nominal
sample period exactly as vio/imu_factor.py:958-970 `_t_nom` does and packs the per-interval parameters.  There is no
CPU path: without the HIP library the calls raise.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

# Status codes of a preintegration request (the reference's values, vio/imu_factor.py:178-184).
OK, REFUSED_BAD_INTERVAL, REFUSED_NO_DATA, REFUSED_BEFORE_FIRST_SAMPLE, REFUSED_NONFINITE, REFUSED_NO_RATE = range(6)
STATUS_TEXT = {
    OK: "ok",
    REFUSED_BAD_INTERVAL: "refused: the interval is empty, reversed or not finite",
    REFUSED_NO_DATA: "refused: the horizon releases no sample",
    REFUSED_BEFORE_FIRST_SAMPLE: "refused: the interval starts before the stream's first sample",
    REFUSED_NONFINITE: "refused: a sample the interval uses is not finite",
    REFUSED_NO_RATE: "refused: no sample period for the covariance (no declared rate, under two released stamps)",
}
I_STATUS, I_NSUB, I_NINNER, I_EXTRAP, I_TEND, I_GAPMAX, I_NREL, I_TNOM, I_SMAX = range(9)
NPAR = 17

# calib["imu_noise"] keys in ImuNoise's argument order (Kalibr/EuRoC names)
CALIB_KEYS = ("gyroscope_noise_density", "accelerometer_noise_density", "gyroscope_random_walk",
              "accelerometer_random_walk")


def _finite_positive(label, value):
    x = float(value)
    if math.isfinite(x) and x > 0.0:
        return x
    raise ValueError("%s must be a finite positive number (got %r)" % (label, value))


def _median_spacing(stamps):
    """The median gap of a stamp sequence, or 0.0 when there are fewer than two stamps or it is not a positive
    finite number."""
    t = np.asarray(stamps, dtype=np.float64).reshape(-1)
    if t.size < 2:
        return 0.0
    gap = float(np.median(np.diff(t)))
    return gap if (math.isfinite(gap) and gap > 0.0) else 0.0


class ImuNoise:
    """Continuous-time IMU noise of vio/imu_factor.py:897-940: white gyro and accelerometer densities sigma_g,
    sigma_a and bias random walks rw_g, rw_a (Kalibr units), plus the sample rate of the discrete noise model
    (None: measured from the stamps of each request).  Every density is required and must be positive and finite;
    nothing is substituted for a missing one."""
    __slots__ = ("sigma_g", "sigma_a", "rw_g", "rw_a", "rate_hz")

    def __init__(self, sigma_g, sigma_a, rw_g, rw_a, rate_hz=None):
        self.sigma_g = _finite_positive("gyro noise density sigma_g", sigma_g)
        self.sigma_a = _finite_positive("accelerometer noise density sigma_a", sigma_a)
        self.rw_g = _finite_positive("gyro bias random walk rw_g", rw_g)
        self.rw_a = _finite_positive("accelerometer bias random walk rw_a", rw_a)
        self.rate_hz = None if rate_hz is None else _finite_positive("IMU sample rate rate_hz", rate_hz)

    @classmethod
    def from_calib(cls, calib, t=None):
        """The four densities of calib["imu_noise"] (all required); the rate is calib["imu_rate_hz"] when that is
        positive, else 1 / the median spacing of the stamps `t` when given, else left to be measured per request."""
        calib = calib or {}
        table = calib.get("imu_noise") or {}
        absent = [k for k in CALIB_KEYS if k not in table]
        if absent:
            raise ValueError("IMU noise calibration is incomplete: %s not given" % ", ".join(absent))
        declared = float(calib.get("imu_rate_hz") or 0.0)
        rate = declared if declared > 0.0 else 0.0
        if rate == 0.0 and t is not None:
            gap = _median_spacing(t)
            rate = 1.0 / gap if gap > 0.0 else 0.0
        return cls(*(table[k] for k in CALIB_KEYS), rate_hz=rate if rate > 0.0 else None)

    def squares(self):
        return self.sigma_g ** 2, self.sigma_a ** 2, self.rw_g ** 2, self.rw_a ** 2


def _as_noise(noise):
    """An ImuNoise from an ImuNoise, a calib imu_noise table, a (sigma_g, sigma_a, rw_g, rw_a[, rate]) sequence or
    None (mean only)."""
    if noise is None or isinstance(noise, ImuNoise):
        return noise
    if isinstance(noise, dict):
        return ImuNoise.from_calib({"imu_noise": noise})
    return ImuNoise(*noise)


def t_nom(noise, t, horizon):
    """The nominal sample period of the sample-support rule (vio/imu_factor.py:958-970): 1 / the declared rate,
    else the median spacing of the stamps released by `horizon` (t <= horizon), else 0 (no covariance then)."""
    if noise is None:
        return 0.0
    if noise.rate_hz is not None:
        return 1.0 / noise.rate_hz
    stamps = np.asarray(t, np.float64)
    return _median_spacing(stamps[:int(np.searchsorted(stamps, horizon, side="right"))])


def _block(r0, c0):
    return property(lambda self: self.J[r0:r0 + 3, c0:c0 + 3])


class Preint:
    """One preintegrated interval [t0, t1] in the body frame at t0, with the fields of vio/imu_factor.py:973-1005.
    When `ok` is False the request was refused (`reason` says why) and only the status fields mean anything.  `W`
    is set by a whitened launch (None when the information is not positive definite)."""
    __slots__ = ("t0", "t1", "dt", "dR", "dv", "dp", "J", "bg_lin", "ba_lin", "cov", "status", "n_sub", "n_inner",
                 "extrap_span", "t_data_end", "gap_max", "n_released", "t_nom", "white_scale_max", "W")

    ok = property(lambda self: self.status == OK)
    reason = property(lambda self: STATUS_TEXT.get(self.status, "unknown status %d" % self.status))
    extrapolated = property(lambda self: self.extrap_span > 0.0)    # ran past the last released sample
    single_sample = property(lambda self: self.n_inner == 0)        # no released sample strictly inside
    # bias Jacobian blocks: rows (phi, v, p) x columns (bg, ba)
    JRg, Jvg, Jva, Jpg, Jpa = _block(0, 0), _block(3, 0), _block(3, 3), _block(6, 0), _block(6, 3)
    cov9 = property(lambda self: self.cov[:9, :9])


class XbaPreintBatch(C.Structure):
    vp = C.c_void_p
    _fields_ = [("nf", C.c_int32), ("t", vp), ("gyr", vp), ("acc", vp), ("s0", vp), ("ns", vp), ("par", vp),
                ("dR", vp), ("dv", vp), ("dp", vp), ("J", vp), ("C", vp), ("info", vp), ("W", vp), ("w_ok", vp),
                ("rec", vp)]


def launch(t, gyr, acc, s0, ns, par, want_w=False, want_rec=False, stream=None, check=True):
    """One `xba_preintegrate` launch over device tensors: t (S,), gyr / acc (S, 3) FP64, s0 (nf,) int64, ns (nf,)
    int32, par (nf, 17) FP64 (t0, t1, horizon, bg_lin, ba_lin, sg2, sa2, swg2, swa2, t_nom, kappa, sbg2, sba2).
    Returns a dict of device tensors (dR (nf,3,3), dv, dp, J (nf,9,6), C (nf,15,15), info (nf,9) [, W, w_ok]
    [, rec (nf, 301)]).  Asynchronous on `stream` (default: torch's current stream), except for the bounds check
    of the streams against the sample arrays (a device reduction read back on the host), which `check=False` skips
    for inputs already checked."""
    import torch
    L = _lib.load()
    nf = int(par.shape[0])
    dev = par.device
    for x, dt, shp in ((t, torch.float64, None), (gyr, torch.float64, 3), (acc, torch.float64, 3),
                       (s0, torch.int64, None), (ns, torch.int32, None), (par, torch.float64, NPAR)):
        if x.dtype != dt or not x.is_contiguous() or x.device != dev or (shp and x.shape[-1] != shp):
            raise _lib.XbaError("preintegration inputs: contiguous device tensors of the stated dtypes and shapes")
    if s0.shape[0] != nf or ns.shape[0] != nf or gyr.shape[0] != t.shape[0] or acc.shape[0] != t.shape[0]:
        raise _lib.XbaError("preintegration inputs: inconsistent lengths")
    if check and nf and bool(((s0 < 0) | (ns < 0) | (s0 + ns.to(torch.int64) > t.shape[0])).any()):
        raise _lib.XbaError("preintegration inputs: a stream runs outside the sample arrays")
    out = dict(dR=torch.empty((nf, 3, 3), dtype=torch.float64, device=dev),
               dv=torch.empty((nf, 3), dtype=torch.float64, device=dev),
               dp=torch.empty((nf, 3), dtype=torch.float64, device=dev),
               J=torch.empty((nf, 9, 6), dtype=torch.float64, device=dev),
               C=torch.empty((nf, 15, 15), dtype=torch.float64, device=dev),
               info=torch.empty((nf, 9), dtype=torch.float64, device=dev))
    if want_w or want_rec:
        out["W"] = torch.empty((nf, 15, 15), dtype=torch.float64, device=dev)
        out["w_ok"] = torch.empty(nf, dtype=torch.int32, device=dev)
    if want_rec:
        out["rec"] = torch.empty((nf, 301), dtype=torch.float64, device=dev)
    b = XbaPreintBatch(nf, t.data_ptr(), gyr.data_ptr(), acc.data_ptr(), s0.data_ptr(), ns.data_ptr(),
                       par.data_ptr(), out["dR"].data_ptr(), out["dv"].data_ptr(), out["dp"].data_ptr(),
                       out["J"].data_ptr(), out["C"].data_ptr(), out["info"].data_ptr(),
                       out["W"].data_ptr() if "W" in out else None, out["w_ok"].data_ptr() if "w_ok" in out else None,
                       out["rec"].data_ptr() if "rec" in out else None)
    st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
    _lib.check(L.xba_preintegrate(C.cast(C.byref(b), C.c_void_p), C.c_void_p(st)), "xba_preintegrate")
    return out


def pack(streams, intervals, noise, kappa=1.0, bias_sigma_g=0.0, bias_sigma_a=0.0):
    """Host packing: streams [(t, gyro, acc)], intervals [(stream index, t0, t1, horizon or None, bg_lin, ba_lin)]
    -> (t, gyr, acc, s0, ns, par) numpy arrays for `launch`."""
    noise = _as_noise(noise)
    sg2, sa2, swg2, swa2 = (0.0, 0.0, 0.0, 0.0) if noise is None else noise.squares()
    ts = [np.ascontiguousarray(s[0], np.float64).reshape(-1) for s in streams]
    gs = [np.ascontiguousarray(s[1], np.float64).reshape(-1, 3) for s in streams]
    as_ = [np.ascontiguousarray(s[2], np.float64).reshape(-1, 3) for s in streams]
    for t, g, a in zip(ts, gs, as_):
        if g.shape[0] != t.shape[0] or a.shape[0] != t.shape[0]:
            raise ValueError("t (n,), gyro (n,3), acc (n,3) expected")
    off = np.concatenate([[0], np.cumsum([len(t) for t in ts])]).astype(np.int64)
    nf = len(intervals)
    s0 = np.zeros(nf, np.int64)
    ns = np.zeros(nf, np.int32)
    par = np.zeros((nf, NPAR))
    for f, (si, t0, t1, horizon, bg, ba) in enumerate(intervals):
        hz = np.inf if horizon is None else float(horizon)
        s0[f] = off[si]
        ns[f] = len(ts[si])
        par[f, 0:3] = (float(t0), float(t1), hz)
        par[f, 3:6] = np.asarray(bg, np.float64).reshape(3)
        par[f, 6:9] = np.asarray(ba, np.float64).reshape(3)
        par[f, 9:13] = (sg2, sa2, swg2, swa2)
        par[f, 13] = t_nom(noise, ts[si], hz)
        par[f, 14:17] = (float(kappa), float(bias_sigma_g) ** 2, float(bias_sigma_a) ** 2)
    cat = lambda xs, shape: np.concatenate(xs) if xs else np.zeros(shape)
    return cat(ts, (0,)), cat(gs, (0, 3)), cat(as_, (0, 3)), s0, ns, par


def preintegrate_many(streams, intervals, noise, kappa=1.0, bias_sigma_g=0.0, bias_sigma_a=0.0, whiten=False,
                      device=None):
    """Preintegrate every interval of `intervals` ([(stream index, t0, t1, horizon or None, bg_lin, ba_lin)]) over
    `streams` ([(t, gyro, acc)]) in one launch; returns a list of `Preint` (W set when `whiten`)."""
    import torch
    if not intervals:
        return []
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    t, g, a, s0, ns, par = pack(streams, intervals, noise, kappa, bias_sigma_g, bias_sigma_a)
    T = lambda x, dt: torch.as_tensor(x, dtype=dt, device=dev).contiguous()
    out = launch(T(t, torch.float64), T(g, torch.float64), T(a, torch.float64), T(s0, torch.int64),
                 T(ns, torch.int32), T(par, torch.float64), want_w=whiten)
    host = {k: v.cpu().numpy() for k, v in out.items()}
    res = []
    for f, (_si, t0, t1, _hz, bg, ba) in enumerate(intervals):
        p = Preint()
        p.t0, p.t1, p.dt = float(t0), float(t1), float(t1) - float(t0)
        p.dR, p.dv, p.dp = host["dR"][f], host["dv"][f], host["dp"][f]
        p.J, p.cov = host["J"][f], host["C"][f]
        p.bg_lin = np.asarray(bg, np.float64).reshape(3).copy()
        p.ba_lin = np.asarray(ba, np.float64).reshape(3).copy()
        info = host["info"][f]
        p.status, p.n_sub, p.n_inner = int(info[I_STATUS]), int(info[I_NSUB]), int(info[I_NINNER])
        p.extrap_span, p.t_data_end, p.gap_max = float(info[I_EXTRAP]), float(info[I_TEND]), float(info[I_GAPMAX])
        p.n_released, p.t_nom, p.white_scale_max = int(info[I_NREL]), float(info[I_TNOM]), float(info[I_SMAX])
        p.W = (host["W"][f] if host["w_ok"][f] else None) if whiten else None
        res.append(p)
    return res


def preintegrate(t, gyro, acc, t0, t1, horizon, bg_lin, ba_lin, noise):
    """vio/imu_factor.py:1008-1048: preintegrate [t0, t1] from the samples with t <= horizon (None = +inf)."""
    return preintegrate_many([(t, gyro, acc)], [(0, t0, t1, horizon, bg_lin, ba_lin)], noise)[0]


def sqrt_information_many(pres, kappa=1.0, bias_sigma_g=0.0, bias_sigma_a=0.0, device=None):
    """`sqrt_information` of many preintegrated intervals in one `xba_sqrt_information` launch: a list of W (15 x 15
    lower triangular) with W^T W = (kappa C + BIASCOV)^-1, None where that matrix is not positive definite.
    Raises ValueError for a refused interval or a kappa that is not finite and positive."""
    import torch
    pres = list(pres)
    kappa = _finite_positive("kappa", kappa)
    for p in pres:
        if not p.ok:
            raise ValueError("no information for a refused preintegration (%s)" % p.reason)
    if not pres:
        return []
    L = _lib.load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    nf = len(pres)
    f64 = lambda x: torch.as_tensor(np.ascontiguousarray(x, np.float64), device=dev).contiguous()
    Cm = f64(np.stack([p.cov for p in pres]))
    Jm = f64(np.stack([p.J for p in pres]))
    par = f64(np.tile([kappa, float(bias_sigma_g) ** 2, float(bias_sigma_a) ** 2], (nf, 1)))
    W = torch.empty((nf, 15, 15), dtype=torch.float64, device=dev)
    ok = torch.empty(nf, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(L.xba_sqrt_information(nf, Cm.data_ptr(), Jm.data_ptr(), par.data_ptr(), W.data_ptr(), ok.data_ptr(),
                                      C.c_void_p(st)), "xba_sqrt_information")
    Wh, okh = W.cpu().numpy(), ok.cpu().numpy()
    return [Wh[f] if okh[f] else None for f in range(nf)]


def sqrt_information(pre, kappa=1.0, bias_sigma_g=0.0, bias_sigma_a=0.0):
    """vio/imu_factor.py:1051-1063: W (15 x 15, lower triangular) with W^T W = (kappa C + BIASCOV)^-1 for this
    kappa and these bias sigmas, computed on the GPU from the interval's covariance and bias Jacobians.  Raises
    ValueError for a refused interval, a bad kappa or an information matrix that is not positive definite."""
    W = sqrt_information_many([pre], kappa, bias_sigma_g, bias_sigma_a)[0]
    if W is None:
        raise ValueError("kappa C + BIASCOV of this interval is not positive definite (%s)" % pre.reason)
    return W
