"""Finite-difference check of the window-step adjoint on the GPU, many windows at once.

The differentiable quantity is the Gauss-Newton step dx of every window at a fixed state (the LM solution), as a
function of the observed pixel coordinates z and the per-observation weights w (``adjoint.WindowStep``).  For a
loss L = sum_w g_w . dx_w, the check perturbs one coordinate per window at a time (the windows are independent, so
one batched FP64 step evaluates every window's perturbation together) and forms central differences
(L(+h) - L(-h)) / 2h, compared with the adjoint's dL/dz and dL/dw on the same coordinates.  Rows within 1e-3 of the
Huber threshold (whitened norm) are not sampled: the Huber weight's derivative jumps there.
"""
from __future__ import annotations

import numpy as np

from .window import BatchSolver


def sample_coords(solver64, rng, per_window, kink_gap=1e-3):
    """Per window, up to ``per_window`` (row, component) pairs of active observation rows away from the Huber kink;
    component 0/1 are the pixel coordinates (1 only for left-camera rows), 2 the row weight.  Rows are indices in the
    window's original observation order."""
    solver64.upload(0.0)
    lin = solver64.linearize(0.0, want_chi2=True, want_marginals=False)
    out = []
    for pb, r in zip(solver64.problems, lin):
        chi = r["ob_chi2"]
        k_huber = pb.huber_k
        ok = np.isfinite(chi) & (np.abs(np.sqrt(np.where(np.isfinite(chi), chi, 0.0)) - k_huber) > kink_gap)
        ok &= np.asarray(pb.ob_valid, bool)
        rows = rng.permutation(np.nonzero(ok)[0])
        comps = []
        for k in rows:
            nd = 2 if int(pb.ob_cam[k]) == 0 else 1
            for d in list(range(nd)) + [2]:
                comps.append((int(k), d))
            if len(comps) >= per_window:
                break
        out.append(comps[:per_window])
    return out


def fd_gradients(problems, g, coords, h=1e-6):
    """Central differences of L = sum_w g_w . dx_w (FP64 GPU step at the problems' current state) at ``coords``
    (from ``sample_coords``); ``problems`` carry ob_w.  Returns a list (per window) of FD values."""
    import torch
    s = BatchSolver([pb.copy() for pb in problems], precision="fp64", with_weights=True, use_graph=False)
    s.upload(0.0)
    dev = s.dev
    z0 = torch.tensor(np.concatenate([pb.ob_z for pb in problems]), dtype=torch.float64, device=dev)
    w0 = torch.tensor(np.concatenate([pb.ob_w for pb in problems]), dtype=torch.float64, device=dev)
    m0 = np.concatenate([[0], np.cumsum([pb.M for pb in problems])])
    n0 = np.concatenate([[0], np.cumsum([pb.n_state() for pb in problems])])
    gt = torch.as_tensor(np.asarray(g, np.float64), device=dev)
    nmax = max(len(c) for c in coords)
    fd = [np.zeros(len(c)) for c in coords]
    for j in range(nmax):
        idx = [(wi, c[j]) for wi, c in enumerate(coords) if j < len(c)]
        vals = []
        for sgn in (1.0, -1.0):
            z, w = z0.clone(), w0.clone()
            for wi, (k, d) in idx:
                if d < 2:
                    z[m0[wi] + k, d] += sgn * h
                else:
                    w[m0[wi] + k] += sgn * h
            s.set_measurements(z, w)
            dx = s.gn_step()
            prod = (gt * dx).cpu().numpy()
            vals.append(np.array([prod[n0[wi]:n0[wi + 1]].sum() for wi in range(len(problems))]))
        for wi, _kd in idx:
            fd[wi][j] = (vals[0][wi] - vals[1][wi]) / (2 * h)
    return fd


def adjoint_at(coords, gz, gw, problems):
    """The adjoint's values at ``coords`` (gz (M, 2), gw (M,) in the concatenated original observation order)."""
    m0 = np.concatenate([[0], np.cumsum([pb.M for pb in problems])])
    out = []
    for wi, c in enumerate(coords):
        out.append(np.array([gz[m0[wi] + k, d] if d < 2 else gw[m0[wi] + k] for k, d in c]))
    return out


def differentiation_check(problems, iters=10, per_window=200, seed=0, h=1e-6, precision="mixed"):
    """The differentiation workload: LM from the problems' state (``iters`` iterations, ``precision``, one captured
    graph per iteration), the converged step dx as a torch function of (z, w) (``adjoint.WindowStep``), one backward
    of L = g . dx for a random g reusing the step's Cholesky factor, then the adjoint checked against FP64 central
    differences at ``per_window`` coordinates per window and against the FP64 path's own adjoint.  Mutates
    ``problems`` (their state becomes the LM solution; ob_w is set when absent).  Returns timings (ms; the solver's
    planning and graph capture excluded), per-window relative errors, and the sampled coordinates with the adjoint's,
    the FP64 adjoint's and the central differences' values there and the loss weights g."""
    import time

    import torch

    from .adjoint import WindowStep
    rng = np.random.default_rng(seed)
    for pb in problems:
        if getattr(pb, "ob_w", None) is None:
            pb.ob_w = rng.uniform(0.5, 1.5, pb.M)
    solver = BatchSolver(problems, precision=precision, with_weights=True, use_graph=True)
    dev = solver.dev
    z = torch.tensor(np.concatenate([pb.ob_z for pb in problems]), dtype=torch.float64, device=dev,
                     requires_grad=True)
    w = torch.tensor(np.concatenate([pb.ob_w for pb in problems]), dtype=torch.float64, device=dev,
                     requires_grad=True)
    g = torch.as_tensor(rng.standard_normal(solver.n_state), dtype=torch.float64, device=dev)
    # warm-up of the same launches (graph capture, first replays) and of the autograd engine's device thread (started
    # by the first backward of the process), then the timed forward and backward
    solver.set_measurements(z.detach(), w.detach())
    solver.upload(1e-4)
    solver.iterate(1)
    (g * WindowStep.apply(z, w, solver)).sum().backward()
    z.grad = None
    w.grad = None
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    solver.upload(1e-4)
    solver.iterate(iters)
    dx = WindowStep.apply(z, w, solver)
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    (g * dx).sum().backward()
    torch.cuda.synchronize(dev)
    t2 = time.perf_counter()
    gz, gw = z.grad.cpu().numpy(), w.grad.cpu().numpy()
    st = solver.stats()
    solver.download(write_back=True)
    # the FP64 path's adjoint at the same state, and the FD reference
    z64 = z.detach().clone().requires_grad_(True)
    w64 = w.detach().clone().requires_grad_(True)
    s64 = BatchSolver([pb.copy() for pb in problems], precision="fp64", with_weights=True, use_graph=False)
    s64.upload(0.0)
    dx64 = WindowStep.apply(z64, w64, s64)
    (g * dx64).sum().backward()
    gz64, gw64 = z64.grad.cpu().numpy(), w64.grad.cpu().numpy()
    coords = sample_coords(s64, rng, per_window)
    fd = fd_gradients(problems, g.cpu().numpy(), coords, h)
    ad = adjoint_at(coords, gz, gw, problems)
    ad64 = adjoint_at(coords, gz64, gw64, problems)
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    return dict(forward_ms=(t1 - t0) * 1e3, backward_ms=(t2 - t1) * 1e3, lm_iters=iters,
                coords_per_window=[len(c) for c in coords], coords=coords, adjoint=ad, adjoint_fp64=ad64,
                fd=fd, g=g.cpu().numpy(),
                rel_err_vs_fd=[rel(a, f) for a, f in zip(ad, fd)],
                rel_err_fp64_adjoint_vs_fd=[rel(a, f) for a, f in zip(ad64, fd)],
                rel_err_vs_fp64_adjoint=[rel(a, b) for a, b in zip(ad, ad64)],
                accepted_mean=float(np.mean([x["accepted"] for x in st])))
