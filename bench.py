"""Benchmark: BA observations/sec per LM iteration of the batched mixed-precision window solve on one MI355X.

Workload (one "step"): the batched configuration -- 256 independent VO windows (5-30 keyframes, 500-30,000
landmarks log-uniform, 2..min(10, P) views, the camera / noise / 5 % outlier model of config 0), solved from their
initial state with 5 LM iterations each, every window in the same launch sequence (one captured hipGraph replayed
per LM iteration), assembly + Schur computed in FP64 from an FP32 per-observation Jacobian record, FP32 Cholesky of
the Jacobi-scaled reduced system with FP64 iterative refinement, FP64 state and update.  value = sum over windows of
observation rows x 5 / seconds per step.

Also reported: the pose-update error of the mixed path against the FP64 path over the whole batch, the plain FP64
GPU path on config 0 (one window, one LM iteration), the per-kernel durations of one iteration (HIP events on the
solver's stream), the roofline of the dominant kernel, and the CPU oracle (the FP64 C restatement of the
reference's window solve, one thread) timed on a bounded sample of the same windows.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--windows 256]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LM_ITERS = 5
DX_ERR_LIMIT = 1e-6  # pose-update error bound of the mixed path against FP64
NARROW_SLOTS = 16
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_lin_fast_narrow_batched256_traffic.md")


def _measured_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (each counter scaled by its reading on a plain copy of known size); None when that summary is absent."""
    import re
    try:
        txt = open(TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


CFG4_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "assembly_config4_traffic.md")


def _measured_config4_assembly_bytes():
    """HBM bytes of configuration 4's assembly launches (k_lin_fast narrow + wide, one LM iteration) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); None when that summary is absent."""
    import re
    try:
        txt = open(CFG4_TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"assembly_bytes_per_iteration\s*=\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def single_window_configs(torch, synth, BatchSolver, log):
    """Configurations 1 (20 visual-inertial keyframes, n = 300) and 4 (60 visual-inertial keyframes, n = 900,
    200,000 landmarks): one LM iteration from the initial state replayed from its graph, mixed precision and FP64,
    median of 7; the mixed-vs-FP64 first-update error; the damped solve's share (tiled Cholesky + refinement)."""
    out = {}
    for cfg in (1, 4):
        pb, _ = synth.make_config(cfg)
        rec = {"keyframes": pb.P, "n": pb.n_state(), "landmarks": pb.L, "obs_rows": pb.M,
               "imu_factors": len(pb.imu or [])}
        dxs = {}
        for prec in ("mixed", "fp64"):
            b = BatchSolver([pb.copy()], precision=prec, use_graph=True)
            from xclean_amd import autotune
            tune = autotune.tune(b, repeats=5)
            b.upload(1e-4)
            b.iterate(1)
            dxs[prec] = b.last_dx()[0]
            ts = []
            for _ in range(7):
                b.upload(1e-4)
                torch.cuda.synchronize(b.dev)
                a = time.perf_counter()
                b.iterate(1)
                torch.cuda.synchronize(b.dev)
                ts.append(time.perf_counter() - a)
            t = float(np.median(ts))
            b.upload(1e-4)
            k = b.profile_iteration()
            st = b.stats()[0]
            n = pb.n_state()
            rec[prec] = {"value": pb.M / t, "unit": "obs/s per LM iteration", "ms_per_iter": t * 1e3,
                         "plan": tune["options"],
                         "solve_ms": k["chol_solve"], "refinement_solves": st["refine_steps"],
                         "solve_effective_tflops": (n ** 3 / 3.0 + 2.0 * n * n) / (k["chol_solve"] * 1e-3) / 1e12,
                         "kernel_ms": k}
            del b
        rec["dx_rel_err_mixed_vs_fp64"] = float(np.linalg.norm(dxs["mixed"] - dxs["fp64"]) /
                                                np.linalg.norm(dxs["fp64"]))
        if cfg == 4:
            hb = _measured_config4_assembly_bytes()
            if hb is not None:
                ms_asm = rec["mixed"]["kernel_ms"]["lin_schur"] + rec["mixed"]["kernel_ms"]["lin_schur_wide"]
                rec["assembly_hbm"] = {"bytes_per_iteration": hb, "kernel_ms": ms_asm,
                                       "achieved_gbs": hb / (ms_asm * 1e-3) / 1e9,
                                       "frac": hb / (ms_asm * 1e-3) / 1e9 / 8000.0,
                                       "source": os.path.relpath(CFG4_TRAFFIC_SUMMARY, ROOT)}
        log("bench: config %d: mixed %.3e / fp64 %.3e obs/s per LM iteration, mixed-vs-fp64 dx err %.2e" % (
            cfg, rec["mixed"]["value"], rec["fp64"]["value"], rec["dx_rel_err_mixed_vs_fp64"]))
        out["config%d" % cfg] = rec
    return out


def differentiation_config(synth, log, nwin=64, per_window=200):
    """Configuration 3: 64 windows of 15 keyframes x 5,000 landmarks; mixed-precision LM to convergence (10
    iterations), the converged step's implicit-function adjoint w.r.t. pixels and row weights (torch.autograd),
    checked against FP64 central differences (h = 1e-6) at 200 coordinates per window."""
    from xclean_amd import gradcheck
    pbs = [synth.make_config(3, seed=i)[0] for i in range(nwin)]
    r = gradcheck.differentiation_check(pbs, iters=10, per_window=per_window, seed=0)
    M = sum(pb.M for pb in pbs)
    out = {"windows": nwin, "obs_rows": M, "lm_iters": r["lm_iters"], "forward_ms": r["forward_ms"],
           "backward_ms": r["backward_ms"], "forward_obs_per_s_per_lm_iteration": M * r["lm_iters"] /
           (r["forward_ms"] * 1e-3), "coords_per_window": int(np.min(r["coords_per_window"])),
           "max_rel_err_vs_fd": float(np.max(r["rel_err_vs_fd"])),
           "median_rel_err_vs_fd": float(np.median(r["rel_err_vs_fd"])),
           "max_rel_err_fp64_adjoint_vs_fd": float(np.max(r["rel_err_fp64_adjoint_vs_fd"])),
           "max_rel_err_vs_fp64_adjoint": float(np.max(r["rel_err_vs_fp64_adjoint"])),
           "accepted_mean": r["accepted_mean"]}
    log("bench: config 3: forward %.2f ms, backward %.3f ms, adjoint vs FD max rel err %.2e" % (
        out["forward_ms"], out["backward_ms"], out["max_rel_err_vs_fd"]))
    return out


def preintegration_config(torch, synth, log, cpu_budget=3.0, nwin=256, nkf=20, spacing=0.5, reps=20):
    """IMU preintegration, batched: the configuration-1 shape (20 keyframes 0.5 s apart, 19 intervals of ~100
    sub-intervals at 200 Hz) for 256 windows at once (4,864 intervals), EuRoC densities, with the whitening and the
    window-solver records; HIP events over `reps` launches.  CPU baseline: the oracle's preint_core + sqrt_info_core
    on one host thread over the first intervals of the same workload."""
    from oracle import oracle as O
    from xclean_amd import imu, roofline
    streams = synth.imu_streams(nwin, spacing * (nkf - 1), seed=3)
    noise = imu.ImuNoise(synth.IMU_NOISE["sigma_g"], synth.IMU_NOISE["sigma_a"], synth.IMU_NOISE["rw_g"],
                         synth.IMU_NOISE["rw_a"], synth.IMU_NOISE["rate_hz"])
    off = 0.0012  # keyframe stamps fall between IMU samples
    iv = [(w, off + spacing * i, off + spacing * (i + 1) if i + 1 < nkf - 1 else spacing * (i + 1), None,
           np.zeros(3), np.zeros(3)) for w in range(nwin) for i in range(nkf - 1)]
    t, g, a, s0, ns, par = imu.pack(streams, iv, noise)
    dev = torch.device("cuda", torch.cuda.current_device())
    T = lambda x, dt: torch.as_tensor(x, dtype=dt, device=dev).contiguous()
    args = (T(t, torch.float64), T(g, torch.float64), T(a, torch.float64), T(s0, torch.int64), T(ns, torch.int32),
            T(par, torch.float64))
    for _ in range(2):
        out = imu.launch(*args, want_rec=True)  # checks the streams against the sample arrays once
    torch.cuda.synchronize(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out = imu.launch(*args, want_rec=True, check=False)
    e1.record()
    torch.cuda.synchronize(dev)
    ms = e0.elapsed_time(e1) / reps
    info = out["info"].cpu().numpy()
    nsub = info[:, 1]
    nf = len(iv)
    flops = roofline.preint_flops(nsub)
    samples_read = float(np.sum(nsub))  # one new sample per sub-interval (the bracket walk), plus the edge ones
    tfl = flops / (ms * 1e-3) / 1e12
    # spot check against the oracle, and the CPU baseline on the same intervals
    q = noise.squares()
    host = {k: out[k].cpu().numpy() for k in ("dR", "C", "W")}
    err = 0.0
    for f in range(min(64, nf)):
        si, t0, t1, hz, bg, ba = iv[f]
        r = O.preintegrate(*streams[si], t0, t1, hz, bg, ba, *q, par[f, 13])
        W = O.sqrt_info(r["C"], 1.0, 0.0, 0.0, r["J"])
        err = max(err, float(np.abs(host["C"][f] - r["C"]).max() / np.abs(r["C"]).max()),
                  float(np.abs(host["W"][f] - W).max() / np.abs(W).max()), float(np.abs(host["dR"][f] - r["dR"]).max()))
    tc0 = time.perf_counter()
    done = 0
    while done < nf and time.perf_counter() - tc0 < cpu_budget:
        si, t0, t1, hz, bg, ba = iv[done]
        r = O.preintegrate(*streams[si], t0, t1, hz, bg, ba, *q, par[done, 13])
        O.sqrt_info(r["C"], 1.0, 0.0, 0.0, r["J"])
        done += 1
    tc = time.perf_counter() - tc0
    rec = {"workload": "%d windows x %d keyframes %.1f s apart: %d intervals of ~%d sub-intervals at %.0f Hz, "
                       "covariance + whitening + window records" % (nwin, nkf, spacing, nf, int(np.median(nsub)),
                                                                    synth.IMU_NOISE["rate_hz"]),
           "intervals": nf, "sub_intervals": int(np.sum(nsub)), "ms_per_launch": ms,
           "intervals_per_s": nf / (ms * 1e-3), "sub_intervals_per_s": float(np.sum(nsub)) / (ms * 1e-3),
           "all_ok": bool((info[:, 0] == 0).all()),
           "max_rel_err_vs_oracle_first64": err,
           "roofline": {"bound": "fp64 valu (serial sub-interval chain per interval)", "achieved": tfl,
                        "peak": roofline.FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tfl / roofline.FP64_PEAK_TFLOPS,
                        "algorithmic_flops": flops,
                        "algorithmic_bytes": roofline.preint_bytes(samples_read, nf),
                        "kernel": "k_preint"},
           "cpu_baseline": None if done == 0 else {
               "value": done / tc, "unit": "intervals/s", "cores": 1, "kind": "port",
               "sample": "first %d of the %d intervals, oracle preint_core + sqrt_info_core (C), one host thread"
                         % (done, nf)}}
    log("bench: preintegration: %d intervals in %.3f ms (%.3e intervals/s, %.2f TFLOP/s); cpu oracle %s "
        "intervals/s; max rel err vs oracle %.1e" % (nf, ms, rec["intervals_per_s"], tfl,
                                                     "%.3e" % (done / tc) if done else "-", err))
    return rec


def _marg_roofline(ns, ms):
    """The marginalisation launches against the FP64 peak: algorithmic FLOPs (roofline.marg_flops, 6 columns
    marginalised from each window's n + 6) over their HIP-event time.  Bound by the serial Jacobi rounds of the
    largest matrix, not by arithmetic."""
    from xclean_amd import roofline
    fl = float(sum(roofline.marg_flops(n, 6) for n in ns))
    tf = fl / (ms * 1e-3) / 1e12
    return {"bound": "latency (serial Jacobi rounds)", "achieved": tf, "peak": roofline.FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tf / roofline.FP64_PEAK_TFLOPS, "algorithmic_flops": fl,
            "kernels": "k_marg_schur + k_prior_from_info"}


def local_ba_config(torch, synth, log, cpu_budget=3.0, reps=20):
    """The local-BA drop-in (backend/ba.py:713 _solve_np, 5 LM iterations, FP64) at tests/test_ba_solver.py's
    problem (make_problem: 15 keyframes, 600 points, 4 views) and at 30 keyframes / 5,000 points / 6 views: median
    and p95 per call (inputs on the host, as the reference's callers hand them over, so each call includes the
    upload and the download), the allocation-free entry point beside it, and the oracle's C restatement of
    _solve_np on one host thread.  Reference bars: local BA p95 per keyframe 10 ms (tier 1) / 5 ms (tier 2),
    src/xclean/obs/budget.py:84-85."""
    from oracle import oracle as O
    from xclean_amd.ba import _solve_np
    out = {"bars_ms": {"tier1_p95": 10.0, "tier2_p95": 5.0}}
    for name, (nk, nl, op) in (("test_ba_solver_nk15_nl600", (15, 600, 4)), ("nk30_nl5000", (30, 5000, 6))):
        p = synth.make_problem(nk=nk, nl=nl, obs_per=op, seed=5)
        args = (p["poses0"], p["pts0"], p["okf"], p["olm"], p["ouv"], p["rig"], 0, 5)
        rec = {"keyframes": nk, "points": len(p["pts0"]), "obs": int(len(p["okf"])), "lm_iters": 5}
        for mode in (False, True):
            for _ in range(3):
                _solve_np(*args, noalloc=mode)
            ts = []
            for _ in range(reps):
                torch.cuda.synchronize()
                a = time.perf_counter()
                r = _solve_np(*args, noalloc=mode)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - a)
            key = "noalloc" if mode else "gpu"
            rec[key] = {"ms_median": 1e3 * float(np.median(ts)), "ms_p95": 1e3 * float(np.percentile(ts, 95)),
                        "obs_per_s_per_lm_iteration": len(p["okf"]) * 5 / float(np.median(ts))}
        e0 = float(np.sqrt(np.mean(np.sum((p["poses0"][:, :3, 3] - p["poses_true"][:, :3, 3]) ** 2, 1))))
        e1 = float(np.sqrt(np.mean(np.sum((r[0][:, :3, 3] - p["poses_true"][:, :3, 3]) ** 2, 1))))
        rec["pose_error_ratio"] = e1 / e0
        if cpu_budget > 0:
            tc, n = 0.0, 0
            while n < 3 or tc < cpu_budget / 2:
                a = time.perf_counter()
                O.solve_np(*args)
                tc += time.perf_counter() - a
                n += 1
            rec["cpu_baseline"] = {"ms": 1e3 * tc / n, "cores": 1, "kind": "port",
                                   "sample": "%d calls, oracle's C restatement of _solve_np, one host thread" % n}
        log("bench: local BA %s: %.3f ms per call (p95 %.3f; no-alloc %.3f)%s" % (
            name, rec["gpu"]["ms_median"], rec["gpu"]["ms_p95"], rec["noalloc"]["ms_median"],
            "; cpu oracle %.3f ms" % rec["cpu_baseline"]["ms"] if "cpu_baseline" in rec else ""))
        out[name] = rec
    return out


GBA_ROW_RECORD_B = 108 * 8     # per-row products of one assembly (xba_gba.hip RR_*), written once and read once
GBA_ROW_INPUT_B = 4 + 16 + 8   # observing keyframe, pixel, disparity
GBA_LM_RECORD_B = 27 * 8       # per landmark: damped inverse, bl, W_h


def global_ba_config(torch, synth, log, cpu_budget=10.0, nk=6000, nl=450000, iters=8):
    """Run-end global BA (backend/global_ba.py:694 global_ba_id, gba_linsolve.py's deflated PCG) on one trajectory of
    ``nk`` keyframes (6 nk unknowns: the PCG regime above 30,000) and ``nl`` landmarks: LM iterations of pass 1
    timed end to end (assembly, PCG set-up and recursion, trial cost, accept), the default two-pass solve beside
    it, the phase times from HIP events, and the oracle's numpy/scipy restatement on a 2,000-keyframe sample."""
    from xclean_amd.gba import global_ba_id
    from xclean_amd import roofline
    args = ("kf_ids", "poses", "rho", "mbear", "host_idx", "okf", "olm", "ouv", "rig", "odisp", "prior")
    g = synth.make_gba(nk=nk, nl=nl, seed=20261204)
    a = {k: g[k] for k in args}
    global_ba_id(**a, iters=1, passes=1)
    torch.cuda.synchronize()
    h = []
    t0 = time.perf_counter()
    res = global_ba_id(**a, iters=iters, passes=1, history=h, profile=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = res[3]
    gs = st["gba_solver"]
    rows = gs["rows_last_pass"]
    n_it = len(h)
    t1 = time.perf_counter()
    res2 = global_ba_id(**a)
    torch.cuda.synchronize()
    dt2 = time.perf_counter() - t1
    asm_bytes = rows * (2 * GBA_ROW_RECORD_B + GBA_ROW_INPUT_B) + nl * 2 * GBA_LM_RECORD_B + gs["blocks"] * 288
    asm_ms = gs["assembly_ms"]
    gbs = asm_bytes / (asm_ms * 1e-3) / 1e9
    rec = {"workload": "one trajectory: %d keyframes (%d unknowns), %d landmarks, %d observation rows (host rows "
                       "dropped), disparity rows and the landmark prior, keyframe 0 held" % (nk, 6 * nk, nl, rows),
           "unit": "obs rows/s per LM iteration",
           "value": rows * n_it / dt, "ms_per_lm_iteration": 1e3 * dt / n_it, "lm_iterations": n_it,
           "accepted": int(sum(e["accepted"] for e in h)),
           "pcg_iterations_per_solve": gs["pcg_iters"] / max(gs["solves"], 1), "pcg_unconverged": gs["unconverged"],
           "segments": gs["segments"], "seg_len": gs["seg_len"], "blocks": gs["blocks"],
           "phase_ms": {"assembly": asm_ms, "pcg_setup": gs["setup_ms"], "pcg_recursion": gs["pcg_ms"]},
           "default_two_pass": {"seconds": dt2, "rms_px": [res2[3]["rms_px_before"], res2[3]["rms_px_after"]],
                                "rejected": res2[3]["rejected"], "lm_iterations": res2[3]["gba_solver"]["iters_run"]},
           "roofline": {"bound": "hbm", "kernel": "k_gba_rows + k_gba_assemble + k_gba_mirror (the reduced system)",
                        "achieved": gbs, "peak": roofline.HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / roofline.HBM_PEAK_GBS, "traffic": None,
                        "algorithmic_bytes": asm_bytes,
                        "bytes_note": "per row its inputs (%d B) and its product record written and read once "
                                      "(2 x %d B), per landmark its record written and read once (2 x %d B), every "
                                      "6x6 block written once" % (GBA_ROW_INPUT_B, GBA_ROW_RECORD_B, GBA_LM_RECORD_B)}}
    log("bench: global BA %d kf / %d rows: %.2f ms per LM iteration (assembly %.2f, PCG set-up %.2f, recursion "
        "%.2f ms, %.1f PCG iterations per solve); two-pass default %.2f s" % (
            nk, rows, rec["ms_per_lm_iteration"], asm_ms, gs["setup_ms"], gs["pcg_ms"],
            rec["pcg_iterations_per_solve"], dt2))
    if cpu_budget > 0:
        from oracle import gba_oracle as G
        gs_ = synth.make_gba(nk=2000, nl=150000, seed=20261204)
        ac = {k: gs_[k] for k in args}
        tc, n = 0.0, 0
        while n < 1 or tc < cpu_budget:
            t = time.perf_counter()
            G.global_ba_id(**ac, iters=1, passes=1)
            tc += time.perf_counter() - t
            n += 1
        crow = int(np.sum(gs_["host_idx"][gs_["olm"]] != gs_["okf"]))
        rec["cpu_baseline"] = {"value": crow * n / tc, "unit": "obs rows/s per LM iteration", "cores": 1,
                               "kind": "port", "sample": "%d x one LM iteration (system, direct solve, trial) of the "
                                                         "oracle's numpy/scipy restatement on a 2,000-keyframe, "
                                                         "150,000-landmark trajectory (%d rows)" % (n, crow)}
        log("bench: global BA cpu oracle %.3e rows/s per LM iteration" % rec["cpu_baseline"]["value"])
    return rec


def marginalisation_config(torch, pbs, log, cpu_budget=3.0, reps=10):
    """Marginalisation of every window's oldest keyframe at once (vio/engine_live.py:690-739 per window): the
    visual rows of its hosted landmarks and the prior linearised at lambda = 0 (one batched linearisation), its 6
    columns Schur'd out, the square-root prior rebuilt over the remaining blocks.  HIP events around the Schur and
    square-root launches (the linearisation is the solver's own kernel, timed separately).  CPU baseline: the numpy
    oracle of vio/marg.py (schur + from_info) with one BLAS thread on the first windows."""
    from threadpoolctl import threadpool_limits

    from oracle import oracle as O
    from xclean_amd import marg
    from xclean_amd.window import BatchSolver
    qs = []
    for pb in pbs:
        q = pb.copy()
        q.lm_mask = np.asarray(pb.lm_host) == 0
        qs.append(q)
    bs = BatchSolver(qs, precision="fp64", use_graph=False)
    bs.upload(0.0)
    torch.cuda.synchronize(bs.dev)
    t0 = time.perf_counter()
    lin = bs.linearize(0.0, want_chi2=False, want_marginals=False)
    torch.cuda.synchronize(bs.dev)
    lin_ms = (time.perf_counter() - t0) * 1e3  # the batched linearisation and the copy of every H to the host
    Hs, bvs, keeps, margs, ns = [], [], [], [], []
    for q, r in zip(qs, lin):
        H, b = r["H"], r["b"]
        P = q.P
        # the engine keeps the blocks with information (vio/engine_live.py:700-705)
        keep = np.array([c for i in range(1, P) if np.abs(np.diag(H)[6 * i:6 * i + 6]).sum() > 0
                         for c in range(6 * i, 6 * i + 6)], np.int64)
        Hs.append(H)
        bvs.append(b)
        keeps.append(keep)
        margs.append(np.arange(6))
        ns.append(len(keep))
    dev = bs.dev
    ld = max(H.shape[0] for H in Hs)
    Hd = marg._slab(Hs, ld, torch, dev)
    bd = marg._vecs(bvs, ld, torch, dev)
    kp = marg._vecs(keeps, ld, torch, dev, np.int32)
    mp = marg._vecs(margs, ld, torch, dev, np.int32)
    nk = torch.as_tensor(np.array(ns, np.int32), device=dev)
    nm = torch.full((len(Hs),), 6, dtype=torch.int32, device=dev)
    work = torch.empty(len(Hs) * (3 * ld * ld + 2 * ld), dtype=torch.float64, device=dev)
    for _ in range(2):
        H2, b2 = marg.schur_dev(Hd, bd, kp, mp, nk, nm, work)
        S, r, rank, lam, _ = marg.from_info_dev(H2, b2, nk, work=work)
    torch.cuda.synchronize(dev)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ts, tf = 0.0, 0.0
    for _ in range(reps):
        e[0].record()
        H2, b2 = marg.schur_dev(Hd, bd, kp, mp, nk, nm, work)
        e[1].record()
        S, r, rank, lam, _ = marg.from_info_dev(H2, b2, nk, work=work)
        e[2].record()
        torch.cuda.synchronize(dev)
        ts += e[0].elapsed_time(e[1])
        tf += e[1].elapsed_time(e[2])
    ts /= reps
    tf /= reps
    # parity spot check and the CPU baseline
    H2h, b2h, Sh, rh, kh = (x.cpu().numpy() for x in (H2, b2, S, r, rank))
    err = 0.0
    tc0 = time.perf_counter()
    done = 0
    with threadpool_limits(limits=1):
        while done < len(Hs) and (done < 8 or time.perf_counter() - tc0 < cpu_budget):
            W2, w2 = O.marg_schur(Hs[done], bvs[done], keeps[done], margs[done])
            Sr, rr = O.prior_from_info(W2, w2)
            if done < 8:
                n = ns[done]
                Sg, rg = Sh[done, :kh[done], :n], rh[done, :kh[done]]
                err = max(err, float(np.abs(Sg.T @ Sg - Sr.T @ Sr).max() / np.abs(Sr.T @ Sr).max()),
                          float(np.abs(H2h[done, :n, :n] - W2).max() / np.abs(W2).max()))
            done += 1
    tc = time.perf_counter() - tc0
    rec = {"workload": "oldest keyframe of each of the %d batch windows: hosted landmarks' rows + prior at lambda = 0, "
                       "6 columns Schur'd out, square-root prior over the blocks with information (n %d..%d)"
                       % (len(Hs), min(ns), max(ns)),
           "windows": len(Hs), "n_max": int(max(ns)), "linearise_and_copy_ms": lin_ms, "schur_ms": ts, "sqrt_ms": tf,
           "windows_per_s": len(Hs) / ((ts + tf) * 1e-3), "max_rel_err_vs_oracle_first8": err,
           "roofline": _marg_roofline(ns, ts + tf),
           "cpu_baseline": {"value": done / tc, "unit": "windows/s", "cores": 1, "kind": "port",
                            "sample": "first %d windows, numpy oracle of vio/marg.py (schur + from_info), one BLAS "
                                      "thread" % done}}
    log("bench: marginalisation: %d windows, Schur %.3f ms + square root %.3f ms (%.3e windows/s); cpu oracle %.3e "
        "windows/s; max rel err %.1e" % (len(Hs), ts, tf, rec["windows_per_s"], rec["cpu_baseline"]["value"], err))
    return rec


def host_record():
    """The host the CPU baseline ran on: CPU model (lscpu), usable cores, the oracle's compiler and flags."""
    import platform
    import re
    import subprocess
    rec = {"cpu_model": None, "usable_cores": len(os.sched_getaffinity(0)), "compiler": None,
           "cflags": None, "machine": platform.machine()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        rec["cpu_model"] = m.group(1).strip() if m else None
    except (OSError, subprocess.SubprocessError):
        pass
    if rec["cpu_model"] is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    rec["cpu_model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    try:
        rec["compiler"] = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                                         timeout=10).stdout.splitlines()[0]
    except (OSError, subprocess.SubprocessError, IndexError):
        pass
    try:
        m = re.search(r"^CFLAGS \?= (.+)$", open(os.path.join(ROOT, "oracle", "Makefile")).read(), re.M)
        rec["cflags"] = m.group(1).strip() if m else None
    except OSError:
        pass
    return rec


def cpu_oracle_baseline(pbs, budget, threads):
    """The FP64 oracle's window solve (C restatement of vio/window.py:889-982, LM_ITERS iterations) over the batch
    windows in order, one window at a time per worker thread, each thread pinned to its own core (the oracle is one
    C call per window, which releases the GIL).  Workers stop taking windows once `budget` seconds have passed;
    the wall time covers every window taken.  Returns (obs/s per LM iteration, windows done, rows done, cores)."""
    import threading

    from oracle import oracle as O
    cpus = sorted(os.sched_getaffinity(0))[:max(1, threads)]
    nxt = [0]
    lock = threading.Lock()
    done = []
    t0 = time.perf_counter()

    def worker(cpu):
        try:
            os.sched_setaffinity(0, {cpu})  # this thread only
        except OSError:
            pass
        while True:
            with lock:
                i = nxt[0]
                if i >= len(pbs) or time.perf_counter() - t0 > budget:
                    return
                nxt[0] += 1
            O.window_solve(pbs[i].copy(), iters=LM_ITERS)
            with lock:
                done.append(i)

    ws = [threading.Thread(target=worker, args=(c,)) for c in cpus]
    for w in ws:
        w.start()
    for w in ws:
        w.join()
    dt = time.perf_counter() - t0
    rows = sum(pbs[i].M for i in done)
    return rows * LM_ITERS / dt, len(done), rows, len(cpus)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--windows", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of oracle work for each cpu_baseline leg (one thread, then all cores)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the all-cores cpu_baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (three launches) for HBM counter calibration")
    ap.add_argument("--no-tune", action="store_true", help="skip the plan autotuner (solver defaults)")
    ap.add_argument("--no-extra", action="store_true", help="skip configurations 1, 3 and 4")
    ap.add_argument("--marg-only", action="store_true",
                    help="diagnostic: run only the batched marginalisation and print its JSON")
    ap.add_argument("--preint-only", action="store_true",
                    help="diagnostic: run only the batched IMU preintegration and print its JSON")
    ap.add_argument("--local-ba-only", action="store_true",
                    help="diagnostic: run only the local-BA drop-in timings and print their JSON")
    ap.add_argument("--global-ba-only", action="store_true",
                    help="diagnostic: run only the global-BA timings and print their JSON")
    ap.add_argument("--configs-only", action="store_true",
                    help="diagnostic: run only configurations 1 and 4 (for profiling the tiled factorisation) and "
                         "print their JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (independent windows would shard across GPUs "
              "with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from xclean_amd import roofline, synth
    from xclean_amd.window import BatchSolver

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.local_ba_only:
        print(json.dumps(local_ba_config(torch, synth, log, cpu_budget=0.0 if args.no_cpu else 3.0)))
        return
    if args.global_ba_only:
        print(json.dumps(global_ba_config(torch, synth, log, cpu_budget=0.0 if args.no_cpu else 10.0)))
        return
    if args.configs_only:
        print(json.dumps(single_window_configs(torch, synth, BatchSolver, log)))
        return
    if args.preint_only:
        print(json.dumps(preintegration_config(torch, synth, log, cpu_budget=0.0 if args.no_cpu else 3.0)))
        return
    if args.marg_only:
        print(json.dumps(marginalisation_config(torch, synth.make_batch(args.windows, seed=0), log,
                                                cpu_budget=0.0 if args.no_cpu else 3.0)))
        return

    t0 = time.time()
    pbs = synth.make_batch(args.windows, seed=0)
    M = sum(pb.M for pb in pbs)
    L = sum(pb.L for pb in pbs)
    ns = [pb.n_state() for pb in pbs]
    log("bench: %d windows, %d observation rows, %d landmarks (generated in %.1f s)" % (len(pbs), M, L,
                                                                                       time.time() - t0))

    bs = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=True)
    dev = bs.dev
    tune = None
    if not args.no_tune:  # plan options picked on this GPU before the timed region (not timed)
        from xclean_amd import autotune
        tune = autotune.tune(bs, repeats=3, log=log, iters=LM_ITERS)  # the timed step's iterations
        log("bench: autotune -> %s (%.3f ms per LM iteration; %s)" % (
            tune["options"], tune["ms"], "cached" if tune["cached"] else "%.1f s" % tune["seconds"]))
    log("bench: plan", bs.plan_info())

    def step():
        bs.upload(1e-4)
        bs.iterate(LM_ITERS)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize(dev)
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1e3 / max(args.steps, 1)
    # the dominant kernel's launch duration: the same steps replayed at once after the timed region, with HIP events
    # on the solver's stream around the assembly launch of every LM iteration (the iteration then runs as two
    # graphs, the assembly and the rest, which costs ~1 % of the step: not in the timed region above)
    bs.set_assembly_timing(True)
    bs.assembly_times()
    for _ in range(args.steps):
        step()
    asm_ms, asm_n = bs.assembly_times()
    bs.set_assembly_timing(False)
    value = M * LM_ITERS / (ms_step * 1e-3)
    log("bench: %.3f ms per step (%d LM iterations), %.3e obs/s per LM iteration" % (ms_step, LM_ITERS, value))
    st = bs.stats()

    # per-kernel durations of one iteration (events on the solver's stream around each launch, eager), each sample
    # queued right behind three replayed iterations so that the GPU's clocks are those of the steady state
    prof = []
    for _ in range(5):
        bs.upload(1e-4)
        bs.iterate(3)
        prof.append(bs.profile_iteration())
    kern = {k: float(np.mean([p[k] for p in prof])) for k in prof[0]}
    log("bench: kernel ms", kern)
    log("bench: assembly launch over the replayed steps: %.4f ms (mean of %d)" % (asm_ms, asm_n))

    # the same batch on the FP64 path (its own autotuned plan, same step protocol): batched FP64 throughput
    bs64 = BatchSolver([pb.copy() for pb in pbs], precision="fp64", use_graph=True)
    tune64 = None
    if not args.no_tune:
        from xclean_amd import autotune
        tune64 = autotune.tune(bs64, repeats=3, log=None)

    def step64():
        bs64.upload(1e-4)
        bs64.iterate(LM_ITERS)

    for _ in range(args.warmup):
        step64()
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step64()
    torch.cuda.synchronize(dev)
    ms_step64 = (time.perf_counter() - t1) * 1e3 / max(args.steps, 1)
    value64 = M * LM_ITERS / (ms_step64 * 1e-3)
    log("bench: batched FP64 %.3f ms per step, %.3e obs/s per LM iteration" % (ms_step64, value64))

    # error of the timed mixed solver (its tuned plan) against the FP64 solver: first LM iteration of every window
    bs.upload(1e-4)
    bs.iterate(1)
    bs64.upload(1e-4)
    bs64.iterate(1)
    dmx, d64 = bs.last_dx(), bs64.last_dx()
    st_mx = bs.stats()
    errs = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(dmx, d64)]
    fallbacks = int(sum(s_["fallback"] for s_ in st_mx))
    log("bench: mixed vs FP64 first update: max rel err %.2e, median %.2e (%d FP64 fallbacks)"
        % (max(errs), float(np.median(errs)), fallbacks))
    del bs64

    # plain FP64 GPU path: config 0 (one window, one LM iteration, FP64), median of repeated solves
    pb0, _ = synth.make_config(0)
    p0 = BatchSolver([pb0], precision="fp64", use_graph=True)
    ts = []
    for r in range(25):
        p0.upload(1e-4)
        torch.cuda.synchronize(dev)
        a = time.perf_counter()
        p0.iterate(1)
        torch.cuda.synchronize(dev)
        if r >= 5:
            ts.append(time.perf_counter() - a)
    t_plain = float(np.median(ts))
    plain = pb0.M / t_plain
    log("bench: plain FP64 config-0 step %.3f ms -> %.3e obs/s per LM iteration" % (t_plain * 1e3, plain))

    # roofline of the dominant kernel: k_lin_fast<float, 0>, the linearise + Schur launch over the narrow work items
    # (landmarks coupling at most 16 consecutive slots), priced with the algorithmic figures of roofline.py on the
    # observations and landmarks that launch covers
    M_n = L_n = 0
    ks_n = []
    narrow_slots = bs.plan_options["slot_tile"] or NARROW_SLOTS
    for w in bs.packed:
        narrow = roofline.landmark_span(w.lm_host, w.ob_start, w.ob_pose) <= narrow_slots
        cnt = np.diff(w.ob_start)
        M_n += int(cnt[narrow].sum())
        L_n += int(narrow.sum())
        ks_n.append(roofline.touched_slots(w.lm_host, w.ob_start, w.ob_pose)[narrow])
    ks_n = np.concatenate(ks_n)
    # the narrow launch's duration: HIP events on the solver's stream right around the launch (the eager profile,
    # steady clocks, mean of 5); cross-checked by the assembly graph's span over the replayed steps, which also holds
    # the graph boundary
    t_lin = kern["lin_schur"] * 1e-3
    # the launch reads the measurements and whitening in FP64 and computes in FP64: the FP64 figures apply
    lin_bytes = roofline.lin_kernel_bytes(M_n, L_n, ns, fp32=False)
    lin_flops = roofline.lin_kernel_flops(M_n, ks_n)
    gbs = lin_bytes / t_lin / 1e9
    tfl = lin_flops / t_lin / 1e12
    ks = np.concatenate([roofline.touched_slots(w.lm_host, w.ob_start, w.ob_pose) for w in bs.packed])
    it_flops = roofline.iteration_flops(M, ks, ns)
    traffic = _measured_traffic()

    extra = None if args.no_extra else single_window_configs(torch, synth, BatchSolver, log)
    diff = None if args.no_extra else differentiation_config(synth, log)
    preint = None if args.no_extra else preintegration_config(torch, synth, log,
                                                              cpu_budget=0.0 if args.no_cpu else 3.0)
    margin = None if args.no_extra else marginalisation_config(torch, pbs, log,
                                                               cpu_budget=0.0 if args.no_cpu else 3.0)
    local_ba = None if args.no_extra else local_ba_config(torch, synth, log, cpu_budget=0.0 if args.no_cpu else 3.0)
    gba = None if args.no_extra else global_ba_config(torch, synth, log, cpu_budget=0.0 if args.no_cpu else 10.0)

    cpu = None
    if not args.no_cpu:
        host = host_record()
        v1, nw1, rows1, _ = cpu_oracle_baseline(pbs, args.cpu_budget, 1)
        ncores = min(args.cpu_threads, host["usable_cores"])
        vn, nwn, rowsn, ncores = cpu_oracle_baseline(pbs, args.cpu_budget, ncores)
        sample = ("the batch windows in order, %d LM iterations each, FP64 oracle (C restatement of vio/window.py "
                  "solve, gcc %s), one window per thread, threads pinned to distinct cores" % (LM_ITERS, host["cflags"]))
        cpu = {"value": vn, "unit": "obs/s per LM iteration", "cores": ncores, "kind": "port",
               "sample": "first %d of the %d windows (%d observation rows); %s" % (nwn, len(pbs), rowsn, sample),
               "pinning": "one thread per core (sched_setaffinity), cores %s" % sorted(
                   os.sched_getaffinity(0))[:ncores],
               "host": host,
               "single_thread": {"value": v1, "cores": 1, "windows": nw1, "obs_rows": rows1},
               "reference_cost_law_anchor": "vio/budget.py:20-25 (Numba FP64, one Ryzen 9 9950X core): "
                                            "~1.7e6 obs/s per LM iteration at configs 0/1"}
        log("bench: cpu oracle %.3e obs/s per LM iteration on %d cores (%d windows); one core %.3e (%d windows); "
            "%s" % (vn, ncores, nwn, v1, nw1, host["cpu_model"]))

    out = {
        "metric": "BA obs/sec per LM iteration: mixed-prec batched vs plain FP64 at <1e-6 dx err",
        "value": value,
        "unit": "obs/s per LM iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 assembly/Schur (f32 Jacobian record), f32 Cholesky + f64 refinement, f64 state",
        "data": "synthetic (seeded generator, config-0 camera/noise/outlier model)",
        "config": {"workload": "batched: %d VO windows, P 5-30, L 500-30000 log-uniform, views 2..min(10,P), "
                               "%d LM iterations per window per step" % (len(pbs), LM_ITERS),
                   "windows": len(pbs), "obs_rows": M, "landmarks": L, "lm_iters": LM_ITERS,
                   "precision": "mixed"},
        "plain_fp64_gpu": {"value": plain, "unit": "obs/s per LM iteration",
                           "workload": "config 0: one window, P=10, L=2000, %d obs rows, 1 LM iteration, FP64"
                           % pb0.M, "ms_per_iter": t_plain * 1e3},
        "speedup_vs_plain_fp64": value / plain,
        "batched_fp64_gpu": {"value": value64, "unit": "obs/s per LM iteration", "ms_per_step": ms_step64,
                             "workload": "the same %d windows, FP64 assembly and Cholesky, own autotuned plan %s"
                             % (len(pbs), None if tune64 is None else tune64["options"])},
        "speedup_vs_batched_fp64": value / value64,
        "dx_rel_err_vs_fp64": {"max": max(errs), "median": float(np.median(errs)), "limit": DX_ERR_LIMIT,
                               "fp64_fallbacks": fallbacks,
                               "note": "first-iteration pose update of the timed mixed solver (its tuned plan) vs "
                                       "the batched FP64 GPU path (which matches the CPU oracle within 1e-13 in the "
                                       "parity tests)"},
        "accepted_steps_mean": float(np.mean([s["accepted"] for s in st])),
        "kernel_ms": kern,
        "roofline": {"bound": "valu", "achieved": tfl, "peak": roofline.FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tfl / roofline.FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "k_lin_fast<float, 0> (linearise + Schur, narrow work items)",
                     "launch_ms": t_lin * 1e3, "launch_timing": "HIP events on the solver's stream right around the "
                     "narrow launch, mean of 5 eager iterations each queued behind 3 replayed ones (steady clocks)",
                     "span_ms": asm_ms, "span_timing": "mean of %d LM iterations: the timed steps replayed once more "
                     "with HIP events around the assembly graph (the narrow launch, the wide one beside it on the side "
                     "stream ending first, and the graph boundary)" % asm_n,
                     "launch_obs_rows": M_n, "launch_landmarks": L_n,
                     "algorithmic_flops": lin_flops, "algorithmic_bytes": lin_bytes,
                     "hbm_achieved_gbs": gbs, "hbm_frac": gbs / roofline.HBM_PEAK_GBS,
                     "note": "compute-bound on the FP64 VALU (no MFMA in this kernel): intensity %.0f FLOP/B is "
                             "above the FP64 ridge (~10 FLOP/B); the peak is MI355X's FP64 rate, which its vector "
                             "and matrix units share" % (lin_flops / lin_bytes)},
        "iteration_flops": it_flops,
        "autotune": None if tune is None else {"options": tune["options"], "ms_per_lm_iteration": tune["ms"],
                                               "cached": tune["cached"], "seconds": tune["seconds"],
                                               "trials": len(tune["trials"])},
        "single_window_configs": extra,
        "differentiation": diff,
        "preintegration": preint,
        "marginalisation": margin,
        "local_ba": local_ba,
        "global_ba": gba,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    sys.stdout.flush()
    if max(errs) > DX_ERR_LIMIT:
        log("bench: FAILED: mixed-precision pose update error %.2e exceeds %.0e" % (max(errs), DX_ERR_LIMIT))
        sys.exit(3)
    if args.calib_copy:
        import ctypes
        from xclean_amd import _lib
        nb = 1 << 30
        src = torch.ones(nb, dtype=torch.uint8, device=dev)
        dst = torch.empty(nb, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        for _ in range(3):
            _lib.check(_lib.load().xba_copy_kernel(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                                   nb, ctypes.c_void_p(stream)), "xba_copy_kernel")
        torch.cuda.synchronize(dev)
        assert bool(torch.equal(src, dst))
        log("bench: calibration copy of %d bytes x3 done" % nb)


if __name__ == "__main__":
    main()
