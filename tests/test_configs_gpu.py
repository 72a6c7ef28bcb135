"""The benchmark configurations at their stated sizes (synth.make_batch / make_config: configurations 2, 3 and 4),
against the CPU FP64 oracle.

- Configuration 2, the 256-window batch (10.4 M observation rows), with the autotuned plan and graph replay: the
  first FP64 update of every window within 1e-9 relative of the oracle's (cost within 1e-9); then 5 LM iterations in
  mixed precision against 5 in FP64: the same accept decisions and lambdas, and the mixed state within 1e-6 of the
  FP64 state relative to the distance the FP64 solve moved it.
- Configuration 4 (60 visual-inertial keyframes, n = 900, 200,000 landmarks, ~1.43 M rows): the first update in FP64
  within 1e-9 and in mixed precision within 1e-6 of the oracle's; 10 graph-replayed mixed iterations with the
  oracle's accept sequence, a non-increasing accepted cost and the final cost within 1e-6.
- Configuration 3 (64 windows x 15 keyframes x 5,000 landmarks, ~30,000 rows each): mixed LM to convergence (10
  iterations), the implicit-function adjoint of the converged step against FP64 central differences at 200
  coordinates per window (the GPU's FP64 step) and at 24 coordinates per window against central differences of the
  oracle's own FP64 Gauss-Newton step (vio/window.py:291-463, 865-886, 933-941 at lambda = 0), all within 1e-5.
  Beside it, the gap between this one-step adjoint and the derivative of the re-converged LM solution (central
  differences of 10-iteration oracle re-solves) is measured and printed.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from xclean_amd import synth
from xclean_amd.problem import Problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _oracle_many(pbs, **kw):
    """The oracle's window solve of every problem (copies), in parallel threads (one C call per window)."""
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(lambda pb: O.window_solve(pb.copy(), **kw), pbs))


def _state_vec(pb):
    return np.concatenate([pb.Rwb.reshape(-1), pb.pwb.reshape(-1), pb.lmp.reshape(-1)])


def test_config2_batch_at_size():
    from xclean_amd import autotune
    from xclean_amd.window import BatchSolver
    pbs = synth.make_batch(256, seed=0)
    M = sum(pb.M for pb in pbs)
    print("config 2: %d windows, %d rows" % (len(pbs), M), flush=True)
    assert len(pbs) == 256 and M > 10_000_000
    outs = _oracle_many(pbs, iters=1, history=True)
    b64 = BatchSolver([pb.copy() for pb in pbs], precision="fp64", use_graph=True)
    t64 = autotune.tune(b64, repeats=1)
    b64.upload(1e-4)
    b64.iterate(1)
    st64 = b64.stats()
    errs = [_rel(d, o[1][0]) for d, o in zip(b64.last_dx(), outs)]
    print("config 2 FP64 first update vs oracle: max %.2e median %.2e (plan %s)"
          % (max(errs), float(np.median(errs)), t64["options"]), flush=True)
    assert max(errs) < 1e-9
    for s, (sr, _dx, acc) in zip(st64, outs):
        assert s["accepted"] == sr["accepted"]
        assert abs(s["cost"] - sr["cost"]) <= 1e-9 * abs(sr["cost"])
    # five iterations, mixed against FP64, each with its autotuned plan, replayed from the captured graph
    qmx = [pb.copy() for pb in pbs]
    q64 = [pb.copy() for pb in pbs]
    bmx = BatchSolver(qmx, precision="mixed", use_graph=True)
    autotune.tune(bmx, repeats=1)
    b64 = BatchSolver(q64, precision="fp64", use_graph=True)
    b64.replan(**t64["options"])
    for b in (bmx, b64):
        b.upload(1e-4)
        b.iterate(5)
        b.download()
    smx, s64 = bmx.stats(), b64.stats()
    worst = 0.0
    for pb, a, b, sa, sb in zip(pbs, qmx, q64, smx, s64):
        assert sa["accepted"] == sb["accepted"] and abs(sa["lam"] - sb["lam"]) <= 1e-12 * sb["lam"]
        assert sa["fallback"] == 0
        moved = np.linalg.norm(_state_vec(b) - _state_vec(pb))
        worst = max(worst, float(np.linalg.norm(_state_vec(a) - _state_vec(b)) / moved))
    print("config 2: 5 mixed iterations vs FP64: worst state difference / FP64 motion %.2e" % worst, flush=True)
    assert worst < 1e-6


def test_config4_at_size():
    from xclean_amd.window import BatchSolver
    pb, _ = synth.make_config(4)
    print("config 4: P=%d n=%d L=%d M=%d" % (pb.P, pb.n_state(), pb.L, pb.M), flush=True)
    assert pb.n_state() == 900 and pb.L == 200_000 and pb.M > 1_400_000
    st1, dx1, acc1 = O.window_solve(pb.copy(), iters=1, history=True)
    for precision, bar in (("fp64", 1e-9), ("mixed", 1e-6)):
        bs = BatchSolver([pb.copy()], precision=precision, use_graph=True)
        bs.upload(1e-4)
        bs.iterate(1)
        e = _rel(bs.last_dx()[0], dx1[0])
        s = bs.stats()[0]
        print("config 4 %s first update vs oracle %.2e (refinement %d, fallback %d)"
              % (precision, e, s["refine_steps"], s["fallback"]), flush=True)
        assert e < bar and s["chol_ok"] and s["accepted"] == st1["accepted"]
    # ten iterations from the captured graph against the oracle's ten
    st10, _dx10, acc10 = O.window_solve(pb.copy(), iters=10, history=True)
    q = pb.copy()
    bs = BatchSolver([q], precision="mixed", use_graph=True)
    bs.upload(1e-4)
    costs = [None]
    for it in range(10):
        bs.iterate(1)
        s = bs.stats()[0]
        assert bool(s["accept_last"]) == bool(acc10[it]), it
        costs.append(s["cost"])
    assert all(b <= a for a, b in zip(costs[1:], costs[2:]))
    assert s["accepted"] == st10["accepted"]
    assert abs(s["cost"] - st10["cost"]) <= 1e-6 * abs(st10["cost"])
    print("config 4: 10 iterations, %d accepted, cost %.6e (oracle %.6e)" % (s["accepted"], s["cost"], st10["cost"]))


# ---- configuration 3: the adjoint against the oracle ---------------------------------------------------------
def _one_landmark(pb, l):
    """The problem restricted to landmark l (all its rows, no prior): its build_system is landmark l's share of the
    reduced system (the reduced system is a sum over landmarks plus the prior)."""
    a, b = int(pb.ob_start[l]), int(pb.ob_start[l + 1])
    q = Problem(Rwb=pb.Rwb, pwb=pb.pwb, Rbc=pb.Rbc, pbc=pb.pbc, lmp=pb.lmp[l:l + 1].copy(),
                lm_host=pb.lm_host[l:l + 1].copy(), lm_mask=np.ones(1, bool), rho_meas=pb.rho_meas[l:l + 1].copy(),
                rho_isig=pb.rho_isig[l:l + 1].copy(), ob_start=np.array([0, b - a], np.int64),
                ob_pose=pb.ob_pose[a:b].copy(), ob_cam=pb.ob_cam[a:b].copy(), ob_z=pb.ob_z[a:b].copy(),
                ob_isig=pb.ob_isig[a:b].copy(), ob_valid=pb.ob_valid[a:b].copy(), fx=pb.fx, fy=pb.fy, cx=pb.cx,
                cy=pb.cy, B=pb.B, huber_k=pb.huber_k, ob_w=pb.ob_w[a:b].copy())
    return q, a


def _oracle_fd(pb, g, coords, h=1e-6):
    """Central differences of L = g . dx_GN with the oracle's FP64 step dx_GN = -(H + 1e-9 I)^-1 b at lambda = 0,
    perturbing one coordinate (row k, component d: 0/1 pixel, 2 weight) at a time; only that row's landmark share of
    H and b is rebuilt (VO window: the reduced system is the pose block itself)."""
    _c, H6, b6, *_ = O.window_build(pb, 0.0)
    H, b, _ = O.window_full_system(pb, H6, b6)
    n = H.shape[0]
    out = []
    for k, d in coords:
        l = int(np.searchsorted(pb.ob_start, k, side="right") - 1)
        q, a = _one_landmark(pb, l)
        _c, H6l, b6l, *_ = O.window_build(q, 0.0)
        vals = []
        for sgn in (1.0, -1.0):
            qq = q.copy()
            if d < 2:
                qq.ob_z[k - a, d] += sgn * h
            else:
                qq.ob_w[k - a] += sgn * h
            _c, H6p, b6p, *_ = O.window_build(qq, 0.0)
            dx = -np.linalg.solve(H + (H6p - H6l) + 1e-9 * np.eye(n), b + (b6p - b6l))
            vals.append(float(g @ dx))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def _tangent(pb, ref):
    """[Log(R_ref^T R), p - p_ref] per slot: the solver's right-perturbation coordinates of pb's poses about ref."""
    return np.concatenate([np.concatenate([Rotation.from_matrix(ref.Rwb[i].T @ pb.Rwb[i]).as_rotvec(),
                                           pb.pwb[i] - ref.pwb[i]]) for i in range(pb.P)])


def test_config3_adjoint_against_the_oracle():
    from xclean_amd import gradcheck
    nwin, per_window, n_oracle = 64, 200, 24
    init = [synth.make_config(3, seed=i)[0] for i in range(nwin)]
    pbs = [pb.copy() for pb in init]
    print("config 3: %d windows x %d rows (mean)" % (nwin, int(np.mean([pb.M for pb in pbs]))), flush=True)
    r = gradcheck.differentiation_check(pbs, iters=10, per_window=per_window, seed=0)
    print("config 3 adjoint vs GPU FP64 FD: max %.2e median %.2e; vs FP64 adjoint %.2e"
          % (max(r["rel_err_vs_fd"]), float(np.median(r["rel_err_vs_fd"])), max(r["rel_err_vs_fp64_adjoint"])),
          flush=True)
    assert min(r["coords_per_window"]) == per_window
    assert max(r["rel_err_vs_fd"]) < 1e-5
    g = r["g"]
    n0 = np.concatenate([[0], np.cumsum([pb.n_state() for pb in pbs])])

    def one(w):
        c = r["coords"][w][:n_oracle]
        return _oracle_fd(pbs[w], g[n0[w]:n0[w + 1]], c), np.asarray(r["adjoint"][w][:n_oracle]), \
            np.asarray(r["adjoint_fp64"][w][:n_oracle])

    with ThreadPoolExecutor(max_workers=16) as ex:
        res = list(ex.map(one, range(nwin)))
    e_mx = [_rel(a, f) for f, a, _ in res]
    e_64 = [_rel(a, f) for f, _, a in res]
    print("config 3 adjoint vs oracle FD (%d coords x %d windows): mixed max %.2e median %.2e, FP64 max %.2e"
          % (n_oracle, nwin, max(e_mx), float(np.median(e_mx)), max(e_64)), flush=True)
    assert max(e_mx) < 1e-5 and max(e_64) < 1e-5

    # the gap to the derivative of the re-converged solution: L*(z) = g . (x*(z) [-] x*(z0)), x* from 10 oracle LM
    # iterations from the initial state, central differences at h = 1e-4 on a few coordinates of two windows
    gaps = []
    for w in (0, 1):
        gw = g[n0[w]:n0[w + 1]]
        base = init[w].copy()
        base.ob_w = pbs[w].ob_w.copy()
        ref = base.copy()
        O.window_solve(ref, iters=10)
        fd, ad = [], []
        for j, (k, d) in enumerate(r["coords"][w][:6]):
            vals = []
            for sgn in (1.0, -1.0):
                q = base.copy()
                if d < 2:
                    q.ob_z[k, d] += sgn * 1e-4
                else:
                    q.ob_w[k] += sgn * 1e-4
                O.window_solve(q, iters=10)
                vals.append(float(gw @ _tangent(q, ref)))
            fd.append((vals[0] - vals[1]) / 2e-4)
            ad.append(r["adjoint_fp64"][w][j])
        gaps.append(_rel(np.array(ad), np.array(fd)))
    print("config 3: one-step adjoint vs derivative of the re-converged solution (10 LM iterations): rel gap %s"
          % ["%.2e" % x for x in gaps], flush=True)
    assert max(gaps) < 0.5
