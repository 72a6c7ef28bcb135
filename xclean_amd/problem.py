"""Host-side containers of one sliding-window BA problem.

They mirror the reference's window arrays field for field so a caller can hand the same object to either solver:
  - ``Problem``  <- vio/window.py:768-792  (CSR-by-landmark observation arrays, VO or VI states, prior, IMU list)
  - ``PoseFree`` <- vio/window.py:795-810  (landmarks whose only observation is their host's left one)
  - ``Prior``    <- vio/marg.py:57-127     (square-root prior 1/2 ||S d + r||^2, d = x [-] x0 per block)
  - ``Preint``   <- vio/imu_factor.py:951-1005 (the fields of a preintegration record the factor reads)

Only data lives here; the arithmetic runs in the HIP library (xclean_amd/csrc).
"""
from __future__ import annotations

import numpy as np

IMU_REC = 301  # dR 9, dv 3, dp 3, dt 1, J 54, bg_lin 3, ba_lin 3, W 225 (packed per IMU factor)


class Problem:
    """One window's arrays in solver form (field names and meanings of vio/window.py:768-792).

    VO: every slot is a 6-dim pose block [dphi, dp] (R_wb <- R_wb Exp(dphi), p_wb <- p_wb + dp).  VI: ``dims`` (P,)
    gives 6 or 15 per slot and ``vel``/``bgs``/``bas`` (P,3) hold the extra state; ``imu`` is a list of
    (i, j, Preint, W) factors between 15-dim slots, ``g_w`` the window's gravity (held fixed).  ``ob_w`` (M,),
    optional, multiplies each observation row's robust weight and cost (1 when absent; not a field of the
    reference: it is what the adjoint differentiates with respect to)."""

    __slots__ = ("Rwb", "pwb", "Rbc", "pbc", "lmp", "lm_host", "lm_mask", "rho_meas", "rho_isig", "ob_start",
                 "ob_pose", "ob_cam", "ob_z", "ob_isig", "ob_valid", "fx", "fy", "cx", "cy", "B", "huber_k",
                 "prior", "dims", "vel", "bgs", "bas", "imu", "g_w", "pf", "ob_w")

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))

    @property
    def P(self):
        return int(self.Rwb.shape[0])

    @property
    def L(self):
        return int(self.lmp.shape[0])

    @property
    def M(self):
        return int(self.ob_pose.shape[0])

    def n_state(self):
        return int(6 * self.P if self.dims is None else np.sum(self.dims))

    def copy(self):
        out = Problem()
        for k in self.__slots__:
            v = getattr(self, k)
            if isinstance(v, np.ndarray):
                v = v.copy()
            elif isinstance(v, (Prior, PoseFree)):
                v = v.copy()
            elif isinstance(v, list):
                v = list(v)
            setattr(out, k, v)
        return out


class PoseFree:
    """Landmarks kept out of the Schur system (vio/window.py:795-810): host slot, host left pixel, the common left
    whitening ``isg``, the inverse-depth prior and the state ``lmp`` (updated on every accepted LM iteration)."""

    def __init__(self, host, z, isg, rmeas, ris, lmp):
        self.host = np.ascontiguousarray(host, np.int64)
        self.z = np.ascontiguousarray(np.asarray(z, np.float64).reshape(-1, 2))
        self.isg = float(isg)
        self.rmeas = np.ascontiguousarray(rmeas, np.float64)
        self.ris = np.ascontiguousarray(ris, np.float64)
        self.lmp = np.ascontiguousarray(np.asarray(lmp, np.float64).reshape(-1, 3)).copy()
        n = len(self.host)
        self.Hinv = np.zeros((n, 3, 3))
        self.ok = np.zeros(n, np.bool_)

    def copy(self):
        c = PoseFree(self.host, self.z, self.isg, self.rmeas, self.ris, self.lmp)
        c.Hinv[:] = self.Hinv
        c.ok[:] = self.ok
        return c


class Prior:
    """Square-root prior over slot blocks (vio/marg.py:57-127): cost 1/2 ||S d + r||^2, d = per block
    [Log(R0^T R), p - p0 (, v - v0, bg - bg0, ba - ba0)].  ``bind`` maps the prior's slot ids to problem slots."""

    def __init__(self, sids, dims, x0, S, r):
        self.sids = list(sids)
        self.dims = [int(d) for d in dims]
        z = np.zeros(3)
        self.x0 = []
        for x in x0:
            R, p, v, bg, ba = (tuple(x) + (None,) * 5)[:5]
            self.x0.append((np.array(R, float), np.array(p, float), np.array(z if v is None else v, float),
                            np.array(z if bg is None else bg, float), np.array(z if ba is None else ba, float)))
        self.S = np.asarray(S, float).copy()
        self.r = np.asarray(r, float).copy()
        self._idx = list(range(len(self.sids)))

    @property
    def n(self):
        return int(sum(self.dims))

    def bind(self, index_of_sid, off=None, slot_dims=None):
        """Map the prior's slot ids to problem slots.  A prior block may cover the first 6 dims of a 15-dim slot,
        never more dims than the slot has (RuntimeError)."""
        idx = [int(index_of_sid[s]) for s in self.sids]
        if slot_dims is not None:
            for s, d, i in zip(self.sids, self.dims, idx):
                if d > int(slot_dims[i]):
                    raise RuntimeError("prior block of slot %s spans %d dims; the slot has %d" % (s, d, slot_dims[i]))
        self._idx = idx

    def slots(self):
        return np.asarray(self._idx, np.int32)

    def x0_packed(self):
        out = np.zeros((len(self.sids), 21))
        for k, (R, p, v, bg, ba) in enumerate(self.x0):
            out[k] = np.concatenate([R.reshape(9), p, v, bg, ba])
        return out

    def info(self):
        return self.S.T @ self.S, self.S.T @ self.r

    def copy(self):
        c = Prior(self.sids, self.dims, self.x0, self.S, self.r)
        c._idx = list(self._idx)
        return c


def gauge_prior(sid, R, p, info=1e10):
    """The VO gauge (vio/marg.py:205-209): a 6-DoF prior of information ``info`` on one keyframe."""
    return Prior([sid], [6], [(R, p, None, None, None)], np.sqrt(info) * np.eye(6), np.zeros(6))


class Preint:
    """The fields of a preintegrated interval that the IMU factor reads (vio/imu_factor.py:951-1005)."""

    __slots__ = ("t0", "t1", "dt", "dR", "dv", "dp", "J", "bg_lin", "ba_lin", "cov")

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))


def pack_imu(imu):
    """(n_imu,2) int32 slot pairs and (n_imu, IMU_REC) float64 records of a Problem.imu list."""
    n = len(imu or [])
    ij = np.zeros((n, 2), np.int32)
    rec = np.zeros((n, IMU_REC))
    for f, (i, j, pre, W) in enumerate(imu or []):
        ij[f] = (i, j)
        rec[f] = np.concatenate([np.asarray(pre.dR).reshape(9), pre.dv, pre.dp, [pre.dt],
                                 np.asarray(pre.J).reshape(54), pre.bg_lin, pre.ba_lin, np.asarray(W).reshape(225)])
    return ij, rec
