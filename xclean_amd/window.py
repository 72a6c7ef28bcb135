"""The sliding-window BA solve on the GPU: the host mirror of vio/window.py.

Drop-in functions (same arguments, same return values, same in-place semantics as the reference):
  ``solve(pb, iters=5, lam0=1e-4)``      <- vio/window.py:889-982 (LM: build_system + prior + IMU + Cholesky +
                                             back-substitution + update + trial cost + accept/reject)
  ``final_linearisation(pb)``           <- vio/window.py:985-1008
plus the batched form ``BatchSolver`` / ``solve_batch`` (many windows of uneven size in one launch sequence).

Every number is computed by libxba.so on the GPU.  This module only packs ``Problem`` objects into the flat arrays
of include/xba.h, moves them to device memory (torch tensors) and writes the results back.  Pose-free landmarks
(``pb.pf``, vio/window.py:550-765) are solved as ordinary landmarks whose only observation is their host's left one:
their pose Jacobians are zero, so they add nothing to the reduced system and take their own damped 3x3 step, which
is the reference's exact treatment of them (vio/CONVENTIONS.md, "Pose-free landmarks").
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .problem import pack_imu

_PREC = {"fp64": _lib.XBA_FP64, "mixed": _lib.XBA_MIXED}
NARROW_SLOTS = 16   # the solver's default narrow slot tile


def _huber_k_dof15():
    from .synth import huber_k_dof
    return huber_k_dof(15)


class _Packed:
    """One window in solver layout: landmarks sorted by host (pose-free ones appended), observations CSR."""

    def __init__(self, pb):
        P = pb.P
        pf = pb.pf if (pb.pf is not None and len(pb.pf.host) > 0) else None
        L0, M0 = pb.L, pb.M
        host = np.asarray(pb.lm_host, np.int64)
        lmp = np.asarray(pb.lmp, np.float64)
        mask = np.asarray(pb.lm_mask, bool)
        rm = np.asarray(pb.rho_meas, np.float64)
        ri = np.asarray(pb.rho_isig, np.float64)
        start = np.asarray(pb.ob_start, np.int64)
        cnt = np.diff(start)
        ob_pose = np.asarray(pb.ob_pose, np.int64)
        ob_cam = np.asarray(pb.ob_cam, np.int64)
        ob_z = np.asarray(pb.ob_z, np.float64).reshape(-1, 2)
        ob_isig = np.asarray(pb.ob_isig, np.float64)
        ob_valid = np.asarray(pb.ob_valid, bool)
        has_w = getattr(pb, "ob_w", None) is not None
        ob_w = np.asarray(pb.ob_w, np.float64) if has_w else np.ones(M0)
        npf = 0
        if pf is not None:
            npf = len(pf.host)
            host = np.concatenate([host, pf.host])
            lmp = np.concatenate([lmp, pf.lmp])
            mask = np.concatenate([mask, np.ones(npf, bool)])
            rm = np.concatenate([rm, pf.rmeas])
            ri = np.concatenate([ri, pf.ris])
            cnt = np.concatenate([cnt, np.ones(npf, np.int64)])
            ob_pose = np.concatenate([ob_pose, pf.host])
            ob_cam = np.concatenate([ob_cam, np.zeros(npf, np.int64)])
            ob_z = np.concatenate([ob_z, pf.z])
            ob_isig = np.concatenate([ob_isig, np.full(npf, pf.isg)])
            ob_valid = np.concatenate([ob_valid, np.ones(npf, bool)])
            ob_w = np.concatenate([ob_w, np.ones(npf)])
            start = np.concatenate([[0], np.cumsum(cnt)])
        self.P, self.L0, self.M0, self.npf = P, L0, M0, npf
        self.L, self.M = L0 + npf, M0 + npf
        # landmark order: those coupling at most NARROW_SLOTS consecutive slots first, then by the first and last
        # slot they couple (the solver's work items are contiguous landmark ranges inside a slot tile, so this
        # keeps neighbouring landmarks in the same tile and the work items long)
        owner = np.repeat(np.arange(self.L), cnt)
        lo = host.copy()
        hi = host.copy()
        if len(owner):
            np.minimum.at(lo, owner, ob_pose)
            np.maximum.at(hi, owner, ob_pose)
        wide = (hi - lo + 1) > NARROW_SLOTS
        perm = np.lexsort((hi, lo, wide))
        self.perm = perm
        cnt_s = cnt[perm]
        st_s = np.zeros(self.L + 1, np.int64)
        st_s[1:] = np.cumsum(cnt_s)
        # source observation of every sorted observation
        if self.M:
            rep = np.repeat(np.arange(self.L), cnt_s)
            self.obs_src = (start[perm][rep] + (np.arange(self.M) - st_s[rep])).astype(np.int64)
        else:
            self.obs_src = np.zeros(0, np.int64)
        o = self.obs_src
        self.lm_host = host[perm].astype(np.int32)
        self.lmp = np.ascontiguousarray(lmp[perm])
        self.lm_mask = mask[perm].astype(np.uint8)
        self.rho_meas = rm[perm]
        self.rho_isig = ri[perm]
        self.ob_start = st_s
        self.ob_pose = ob_pose[o].astype(np.int32)
        self.ob_cam = ob_cam[o].astype(np.uint8)
        self.ob_z = np.ascontiguousarray(ob_z[o])
        self.ob_isig = ob_isig[o]
        self.ob_valid = ob_valid[o].astype(np.uint8)
        self.has_w = has_w
        self.ob_w = ob_w[o]
        self.dims = (np.full(P, 6, np.int32) if pb.dims is None else np.asarray(pb.dims, np.int32))
        self.n = int(self.dims.sum())
        self.vi = pb.vel is not None
        imu_ij, imu_rec = pack_imu(pb.imu)
        self.imu_ij, self.imu_rec = imu_ij, imu_rec
        wp = np.zeros(_lib.XBA_WPAR)
        wp[0:7] = [pb.fx, pb.fy, pb.cx, pb.cy, pb.B, pb.huber_k, _huber_k_dof15() if len(imu_ij) else 0.0]
        wp[8:17] = np.asarray(pb.Rbc, np.float64).reshape(9)
        wp[17:20] = np.asarray(pb.pbc, np.float64).reshape(3)
        if pb.g_w is not None:
            wp[20:23] = np.asarray(pb.g_w, np.float64).reshape(3)
        self.win_par = wp
        pr = pb.prior
        if pr is not None:
            self.prior_slot = pr.slots().astype(np.int32)
            self.prior_dim = np.asarray(pr.dims, np.int32)
            self.prior_x0 = pr.x0_packed()
            self.prior_S = np.ascontiguousarray(pr.S, np.float64)
            self.prior_r = np.asarray(pr.r, np.float64)
        else:
            self.prior_slot = np.zeros(0, np.int32)
            self.prior_dim = np.zeros(0, np.int32)
            self.prior_x0 = np.zeros((0, 21))
            self.prior_S = np.zeros((0, 0))
            self.prior_r = np.zeros(0)

    def unpermute_lm(self, a):
        out = np.empty_like(a)
        out[self.perm] = a
        return out

    def unpermute_obs(self, a):
        out = np.empty_like(a)
        out[self.obs_src] = a
        return out


class BatchSolver:
    """Solve many windows at once on the GPU.  ``problems`` is a list of ``Problem``s (mutated by ``download``)."""

    def __init__(self, problems, precision="fp64", lm_per_chunk=0, slot_tile=0, grid=0, use_graph=True,
                 device="cuda", deterministic=False, with_weights=False, big_n=0, big_fp64=0, tuned=None):
        """``tuned``: a dict of plan options (lm_per_chunk, slot_tile, grid, big_n, big_fp64) chosen by
        ``xclean_amd.autotune``, overriding the individual arguments.  ``big_fp64`` (mixed precision): 1, the windows
        above ``big_n`` are factored in FP64 instead of FP32 with FP64 refinement; 2, every window."""
        import torch
        _lib.require_gpu()
        self.torch = torch
        self.L = _lib.load()
        self.problems = list(problems)
        self.prec = _PREC[precision]
        self.dev = torch.device(device)
        W = [_Packed(pb) for pb in self.problems]
        self.packed = W
        nw = len(W)
        cat = np.concatenate
        off = lambda xs: np.concatenate([[0], np.cumsum(xs)]).astype(np.int32)
        self.win_slot0 = off([w.P for w in W])
        self.win_lm0 = off([w.L for w in W])
        self.win_imu0 = off([len(w.imu_ij) for w in W])
        self.win_pblk0 = off([len(w.prior_slot) for w in W])
        self.win_prow0 = off([w.prior_S.shape[0] for w in W])
        self.win_obs0 = off([w.M for w in W])
        self.win_n0 = off([w.n for w in W])
        self.slot_dim = cat([w.dims for w in W]).astype(np.int32)
        self.ob_start = cat([w.ob_start[:-1] + self.win_obs0[i] for i, w in enumerate(W)] +
                            [[self.win_obs0[-1]]]).astype(np.int32)
        self.lm_host = cat([w.lm_host for w in W]).astype(np.int32)
        self.ob_pose = cat([w.ob_pose for w in W]).astype(np.int32)
        self.imu_ij = cat([w.imu_ij for w in W]).astype(np.int32).reshape(-1, 2) if nw else np.zeros((0, 2), np.int32)
        self.prior_slot = cat([w.prior_slot for w in W]).astype(np.int32)
        self.prior_dim = cat([w.prior_dim for w in W]).astype(np.int32)
        vi = any(w.vi for w in W)
        self.vi = vi
        nslot = int(self.win_slot0[-1])

        def dev(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a, dt)).to(self.dev)

        def slot_arr(attr):
            out = np.zeros((nslot, 3))
            for i, pb in enumerate(self.problems):
                v = getattr(pb, attr)
                if v is not None:
                    out[self.win_slot0[i]:self.win_slot0[i + 1]] = v
            return out

        self.d = dict(
            Rwb=dev(cat([np.asarray(pb.Rwb, np.float64).reshape(-1, 9) for pb in self.problems]), np.float64),
            pwb=dev(cat([np.asarray(pb.pwb, np.float64).reshape(-1, 3) for pb in self.problems]), np.float64),
            vel=dev(slot_arr("vel"), np.float64), bgs=dev(slot_arr("bgs"), np.float64),
            bas=dev(slot_arr("bas"), np.float64),
            lmp=dev(cat([w.lmp for w in W]), np.float64),
            lm_mask=dev(cat([w.lm_mask for w in W]), np.uint8),
            rho_meas=dev(cat([w.rho_meas for w in W]), np.float64),
            rho_isig=dev(cat([w.rho_isig for w in W]), np.float64),
            ob_z=dev(cat([w.ob_z for w in W]), np.float64),
            ob_isig=dev(cat([w.ob_isig for w in W]), np.float64),
            ob_cam=dev(cat([w.ob_cam for w in W]), np.uint8),
            ob_valid=dev(cat([w.ob_valid for w in W]), np.uint8),
            win_par=dev(np.stack([w.win_par for w in W]), np.float64),
            imu_rec=dev(cat([w.imu_rec for w in W]).reshape(-1, _lib.XBA_IMU_REC), np.float64),
            prior_x0=dev(cat([w.prior_x0 for w in W]).reshape(-1, 21), np.float64),
            prior_S=dev(cat([w.prior_S.reshape(-1) for w in W]), np.float64),
            prior_r=dev(cat([w.prior_r for w in W]), np.float64),
        )
        if with_weights or any(w.has_w for w in W):
            self.d["ob_w"] = dev(cat([w.ob_w for w in W]), np.float64)
        # packed row -> row of the concatenated original observations (pose-free rows have none)
        orig0 = np.concatenate([[0], np.cumsum([w.M0 for w in W])]).astype(np.int64)
        rows_p, rows_o = [], []
        for i, w in enumerate(W):
            keep = np.nonzero(w.obs_src < w.M0)[0]
            rows_p.append(self.win_obs0[i] + keep)
            rows_o.append(orig0[i] + w.obs_src[keep])
        self.rows_packed = torch.from_numpy(cat(rows_p).astype(np.int64)).to(self.dev)
        self.rows_orig = torch.from_numpy(cat(rows_o).astype(np.int64)).to(self.dev)
        self.n_orig_obs = int(orig0[-1])
        b = _lib.XbaBatch()
        b.nwin = nw
        i32 = lambda a: a.ctypes.data_as(_lib.i32p)
        self._host = [self.win_slot0, self.win_lm0, self.win_imu0, self.win_pblk0, self.win_prow0, self.slot_dim,
                      self.ob_start, self.lm_host, self.ob_pose, self.imu_ij, self.prior_slot, self.prior_dim]
        (b.win_slot0, b.win_lm0, b.win_imu0, b.win_pblk0, b.win_prow0, b.slot_dim, b.ob_start, b.lm_host,
         b.ob_pose, b.imu_ij, b.prior_slot, b.prior_dim) = [i32(a) for a in self._host]
        for k, t in self.d.items():
            setattr(b, k, t.data_ptr())
        if not vi:
            b.vel = b.bgs = b.bas = None
        self.batch = b
        self.use_graph = bool(use_graph)
        self.deterministic = bool(deterministic)
        self.ctx = None
        plan = dict(lm_per_chunk=lm_per_chunk, slot_tile=slot_tile, grid=grid, big_n=big_n, big_fp64=big_fp64)
        plan.update(tuned or {})
        self.replan(**plan)
        ns, nh = C.c_int64(), C.c_int64()
        self.L.xba_sizes(self.ctx, C.byref(ns), C.byref(nh))
        self.n_state, self.n_h = ns.value, nh.value
        self.nobs = int(self.win_obs0[-1])
        self.nlm = int(self.win_lm0[-1])

    def replan(self, lm_per_chunk=0, slot_tile=0, grid=0, big_n=0, big_fp64=0):
        """(Re)create the solver context with these plan options over the same device data (the state is
        reloaded by the next ``upload``)."""
        self._destroy()
        o = _lib.XbaOptions()
        o.precision, o.use_graph, o.deterministic = self.prec, int(self.use_graph), int(self.deterministic)
        o.lm_per_chunk, o.slot_tile, o.grid, o.big_n = int(lm_per_chunk), int(slot_tile), int(grid), int(big_n)
        o.big_fp64 = int(big_fp64)
        self.opts = o
        self.plan_options = dict(lm_per_chunk=int(lm_per_chunk), slot_tile=int(slot_tile), grid=int(grid),
                                 big_n=int(big_n), big_fp64=int(big_fp64))
        ctx = C.c_void_p()
        _lib.check(self.L.xba_create(C.byref(ctx), C.byref(self.batch), C.byref(o)), "xba_create")
        self.ctx = ctx

    def _destroy(self):
        ctx = getattr(self, "ctx", None)
        if ctx is not None and ctx.value:
            self.torch.cuda.synchronize()
            self.L.xba_destroy(ctx)
        self.ctx = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    @property
    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    # ---- the LM solve
    def upload(self, lam0=1e-4):
        _lib.check(self.L.xba_upload(self.ctx, float(lam0), self.stream), "xba_upload")

    def iterate(self, iters):
        _lib.check(self.L.xba_iterate(self.ctx, int(iters), self.stream), "xba_iterate")

    def download(self, write_back=True):
        _lib.check(self.L.xba_download(self.ctx, self.stream), "xba_download")
        if not write_back:
            return
        self.torch.cuda.synchronize(self.dev)
        R = self.d["Rwb"].cpu().numpy().reshape(-1, 3, 3)
        p = self.d["pwb"].cpu().numpy()
        lm = self.d["lmp"].cpu().numpy()
        vel, bgs, bas = (self.d[k].cpu().numpy() for k in ("vel", "bgs", "bas"))
        for i, (pb, w) in enumerate(zip(self.problems, self.packed)):
            s0, s1 = self.win_slot0[i], self.win_slot0[i + 1]
            pb.Rwb[:] = R[s0:s1]
            pb.pwb[:] = p[s0:s1]
            if pb.vel is not None:
                pb.vel[:] = vel[s0:s1]
                pb.bgs[:] = bgs[s0:s1]
                pb.bas[:] = bas[s0:s1]
            lmw = w.unpermute_lm(lm[self.win_lm0[i]:self.win_lm0[i + 1]])
            pb.lmp[:] = lmw[:w.L0]
            if w.npf:
                pb.pf.lmp[:] = lmw[w.L0:]

    def stats(self):
        out = self.torch.zeros((len(self.packed), _lib.XBA_NSTATS), dtype=self.torch.float64, device=self.dev)
        _lib.check(self.L.xba_stats(self.ctx, out.data_ptr(), self.stream), "xba_stats")
        a = out.cpu().numpy()
        return [{"cost0": r[0], "cost": r[1], "accepted": int(r[2]), "iters": int(r[3]), "lam": r[4],
                 "chol_ok": bool(r[5]), "accept_last": bool(r[6]), "refine_steps": int(r[7]),
                 "fallback": int(r[8]), "refine_converged": bool(r[9])} for r in a]

    def last_dx(self):
        out = self.torch.zeros(max(self.n_state, 1), dtype=self.torch.float64, device=self.dev)
        _lib.check(self.L.xba_last_dx(self.ctx, out.data_ptr(), self.stream), "xba_last_dx")
        a = out.cpu().numpy()
        return [a[self.win_n0[i]:self.win_n0[i + 1]] for i in range(len(self.packed))]

    def linearize(self, lam=0.0, want_chi2=True, want_marginals=True):
        """One build_system at damping ``lam`` (final_linearisation at lam = 0): per window (H, b, cost, ob_chi2,
        lm_ok, Hll_inv, imu_chi2) in each window's original landmark / observation order."""
        t = self.torch
        H = t.zeros(max(self.n_h, 1), dtype=t.float64, device=self.dev)
        b = t.zeros(max(self.n_state, 1), dtype=t.float64, device=self.dev)
        cost = t.zeros(len(self.packed), dtype=t.float64, device=self.dev)
        chi = t.zeros(max(self.nobs, 1), dtype=t.float64, device=self.dev) if want_chi2 else None
        Hinv = t.zeros((max(self.nlm, 1), 9), dtype=t.float64, device=self.dev) if want_marginals else None
        ok = t.zeros(max(self.nlm, 1), dtype=t.uint8, device=self.dev) if want_marginals else None
        nimu = int(self.win_imu0[-1])
        ichi = t.zeros(max(nimu, 1), dtype=t.float64, device=self.dev)
        p = lambda x: x.data_ptr() if x is not None else None
        _lib.check(self.L.xba_linearize(self.ctx, float(lam), p(H), p(b), p(cost), p(chi), p(Hinv), p(ok), p(ichi),
                                        self.stream), "xba_linearize")
        H, b, cost = H.cpu().numpy(), b.cpu().numpy(), cost.cpu().numpy()
        chi = chi.cpu().numpy() if chi is not None else None
        Hinv = Hinv.cpu().numpy().reshape(-1, 3, 3) if Hinv is not None else None
        ok = ok.cpu().numpy().astype(bool) if ok is not None else None
        ichi = ichi.cpu().numpy()
        out = []
        h0 = 0
        for i, w in enumerate(self.packed):
            n = w.n
            Hw = H[h0:h0 + n * n].reshape(n, n)
            h0 += n * n
            bw = b[self.win_n0[i]:self.win_n0[i + 1]]
            cw = chi[self.win_obs0[i]:self.win_obs0[i + 1]] if chi is not None else None
            cw = w.unpermute_obs(cw) if cw is not None else None
            hw = w.unpermute_lm(Hinv[self.win_lm0[i]:self.win_lm0[i + 1]]) if Hinv is not None else None
            kw = w.unpermute_lm(ok[self.win_lm0[i]:self.win_lm0[i + 1]]) if ok is not None else None
            out.append(dict(H=Hw, b=bw, cost=cost[i], ob_chi2=cw, lm_ok=kw, Hll_inv=hw,
                            imu_chi2=list(ichi[self.win_imu0[i]:self.win_imu0[i + 1]]), packed=w))
        return out

    KERNELS = ("lin_schur", "lin_schur_wide", "reduce", "terms", "chol_solve", "backsub_trial", "backsub_trial_wide",
               "trial_terms", "accept", "commit")

    def profile_iteration(self):
        """Run one LM iteration eagerly with HIP events around every kernel; returns {kernel: ms}."""
        out = np.zeros(len(self.KERNELS))
        _lib.check(self.L.xba_profile_iteration(self.ctx, out.ctypes.data, self.stream), "xba_profile_iteration")
        return dict(zip(self.KERNELS, out.tolist()))

    def set_assembly_timing(self, on=True):
        """Record HIP events around the assembly launch of every LM iteration ``iterate`` runs from now on (the
        iteration's graph is split into the assembly and the rest); ``assembly_times`` reads them."""
        _lib.check(self.L.xba_set_assembly_timing(self.ctx, int(bool(on))), "xba_set_assembly_timing")

    def assembly_times(self):
        """(mean ms, count) of the assembly launches recorded since the last call (waits for them)."""
        ms, n = C.c_double(), C.c_int32()
        _lib.check(self.L.xba_assembly_times(self.ctx, C.byref(ms), C.byref(n)), "xba_assembly_times")
        return (ms.value / n.value if n.value else float("nan")), n.value

    def plan_info(self):
        out = np.zeros(8, np.int64)
        _lib.check(self.L.xba_plan_info(self.ctx, out.ctypes.data), "xba_plan_info")
        return dict(zip(("chunks", "grid", "slot_tile", "lin_lds", "solve_lds", "partH", "partb", "max_n"),
                        out.tolist()))

    def set_measurements(self, z=None, w=None):
        """Overwrite the observed pixels z (M, 2) and/or the per-observation weights w (M,), given in the
        concatenated original observation order of ``problems`` (device tensors or arrays)."""
        t = self.torch
        if z is not None:
            z = t.as_tensor(z, dtype=t.float64, device=self.dev).reshape(-1, 2)
            self.d["ob_z"].view(-1, 2)[self.rows_packed] = z[self.rows_orig]
        if w is not None:
            if "ob_w" not in self.d:
                raise _lib.XbaError("this solver was built without per-observation weights (with_weights=True)")
            w = t.as_tensor(w, dtype=t.float64, device=self.dev).reshape(-1)
            self.d["ob_w"][self.rows_packed] = w[self.rows_orig]

    def gn_step(self):
        """The Gauss-Newton step of every window at the current state (lambda = 0), FP64 factor kept for
        ``adjoint``.  Returns the concatenated step (sum n,) as a device tensor."""
        t = self.torch
        dx = t.empty(max(self.n_state, 1), dtype=t.float64, device=self.dev)
        _lib.check(self.L.xba_gn_step(self.ctx, dx.data_ptr(), self.stream), "xba_gn_step")
        return dx[:self.n_state]

    def adjoint(self, g):
        """dL/dz (M, 2) and dL/dw (M,) in the original observation order, for g = dL/d(dx) of the last
        ``gn_step``."""
        t = self.torch
        g = t.as_tensor(g, dtype=t.float64, device=self.dev).reshape(-1).contiguous()
        gz = t.empty((max(self.nobs, 1), 2), dtype=t.float64, device=self.dev)
        gw = t.empty(max(self.nobs, 1), dtype=t.float64, device=self.dev)
        _lib.check(self.L.xba_backward(self.ctx, g.data_ptr(), gz.data_ptr(), gw.data_ptr(), self.stream),
                   "xba_backward")
        oz = t.zeros((self.n_orig_obs, 2), dtype=t.float64, device=self.dev)
        ow = t.zeros(self.n_orig_obs, dtype=t.float64, device=self.dev)
        oz[self.rows_orig] = gz[self.rows_packed]
        ow[self.rows_orig] = gw[self.rows_packed]
        return oz, ow

    def solve(self, iters=5, lam0=1e-4, write_back=True):
        self.upload(lam0)
        self.iterate(iters)
        self.download(write_back)
        return self.stats()


def solve_batch(problems, iters=5, lam0=1e-4, precision="fp64", **kw):
    """Solve every window of ``problems`` (mutated in place); returns the per-window stats dicts."""
    return BatchSolver(problems, precision=precision, **kw).solve(iters, lam0)


def solve(pb, iters=5, lam0=1e-4, precision="fp64", **kw):
    """vio/window.py:889-982: fixed-iteration LM on one window, ``pb`` mutated in place on accepted steps.
    Returns dict(cost0, cost, accepted, iters, lam)."""
    st = solve_batch([pb], iters, lam0, precision, **kw)[0]
    return {k: st[k] for k in ("cost0", "cost", "accepted", "iters", "lam")}


def final_linearisation(pb, precision="fp64"):
    """vio/window.py:985-1008: (H, b, ob_chi2, lm_ok, Hll_inv, imu_chi2, off) at lambda = 0; the pose-free block's
    marginals go to pb.pf.Hinv / pb.pf.ok."""
    bs = BatchSolver([pb], precision=precision, use_graph=False)
    bs.upload(0.0)
    r = bs.linearize(0.0)[0]
    w = r["packed"]
    dims = w.dims.astype(np.int64)
    off = np.zeros(len(dims) + 1, np.int64)
    np.cumsum(dims, out=off[1:])
    M0 = w.M0
    ob_chi2 = r["ob_chi2"][:M0]
    lm_ok = r["lm_ok"][:w.L0]
    Hinv = r["Hll_inv"][:w.L0]
    if w.npf:
        pb.pf.Hinv[:] = r["Hll_inv"][w.L0:]
        pb.pf.ok[:] = r["lm_ok"][w.L0:]
    return r["H"], r["b"], ob_chi2, lm_ok, Hinv, r["imu_chi2"], off
