"""Algorithmic work of one LM iteration of the window solve, and the MI355X peaks the measurements are priced
against.  The per-unit figures are fixed here so every report computes them the same way.

Symbols: M observation rows (entries of ob_pose), L landmarks, n reduced-system dimension of a window, k_l slots
landmark l couples (its host plus every distinct target slot of an observation with t != h).

Bytes (compulsory traffic of one LM iteration):
  FP32-assembly path:  16 M + 40 L + 8 n^2 + 8 n   (observation row: u, v + slot + whitening = 16 B; landmark:
                       (a, b, rho) 12 + host 4 + rho_meas 4 + rho_isig 4 + CSR offset 4 read, update 12 written;
                       reduced system written once and read once)
  FP64 path:           28 M + 72 L + 16 n^2 + 16 n
FLOPs (an FMA counts 2): 810 M + sum_l (60 + 144 k_l + 216 k_l^2) + n^3/3 + 2 n^2
  (per observation: projection + Jacobians ~150, Hll 36, Hlp 144, b 48, pose-pose 432; per landmark: 3x3 inverse 60,
  Hi Hlp 108 + Hlp^T Hi bl 36 per slot, A2^T (Hi A1) 216 per slot pair; dense Cholesky and two triangular solves)

The landmark kernel (linearise + Schur, the dominant one) is priced with the observation and landmark terms only:
  bytes  16 M + 28 L (+ 4 (n^2 + n) per window, the reduced-system partial it writes once)   [FP32 path]
         28 M + 48 L (+ 8 (n^2 + n))                                                       [FP64 path]
  flops  810 M + sum_l (60 + 144 k_l + 216 k_l^2)
"""
from __future__ import annotations

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
FP32_PEAK_TFLOPS = 157.3       # MI355X dense FP32 (matrix cores; equal to the packed-FP32 vector rate)
FP64_PEAK_TFLOPS = 78.6        # MI355X dense FP64 (vector = matrix)


def touched_slots(lm_host, ob_start, ob_pose):
    """k_l per landmark: its host plus every distinct target slot t != host of its observations (0 if none)."""
    L = len(lm_host)
    cnt = np.diff(ob_start)
    owner = np.repeat(np.arange(L), cnt)
    t = np.asarray(ob_pose, np.int64)
    h = np.asarray(lm_host, np.int64)[owner]
    sel = t != h
    pairs = np.unique(np.stack([owner[sel], t[sel]], 1), axis=0) if sel.any() else np.zeros((0, 2), np.int64)
    k = np.bincount(pairs[:, 0], minlength=L) if len(pairs) else np.zeros(L, np.int64)
    return np.where(k > 0, k + 1, 0)


def landmark_span(lm_host, ob_start, ob_pose):
    """Number of consecutive slots each landmark couples: max - min + 1 over its host and its observations' slots."""
    L = len(lm_host)
    cnt = np.diff(ob_start)
    owner = np.repeat(np.arange(L), cnt)
    h = np.asarray(lm_host, np.int64)
    lo, hi = h.copy(), h.copy()
    np.minimum.at(lo, owner, np.asarray(ob_pose, np.int64))
    np.maximum.at(hi, owner, np.asarray(ob_pose, np.int64))
    return hi - lo + 1


def iteration_bytes(M, L, ns, fp32=True):
    ns = np.asarray(ns, np.float64)
    if fp32:
        return 16.0 * M + 40.0 * L + float(np.sum(8 * ns * ns + 8 * ns))
    return 28.0 * M + 72.0 * L + float(np.sum(16 * ns * ns + 16 * ns))


def iteration_flops(M, k, ns):
    k = np.asarray(k, np.float64)
    ns = np.asarray(ns, np.float64)
    lm = k > 0
    return 810.0 * M + float(np.sum(60 + 144 * k[lm] + 216 * k[lm] ** 2)) + float(np.sum(ns ** 3 / 3 + 2 * ns ** 2))


def lin_kernel_bytes(M, L, ns, fp32=True):
    ns = np.asarray(ns, np.float64)
    if fp32:
        return 16.0 * M + 28.0 * L + float(np.sum(4 * (ns * ns + ns)))
    return 28.0 * M + 48.0 * L + float(np.sum(8 * (ns * ns + ns)))


def lin_kernel_flops(M, k):
    k = np.asarray(k, np.float64)
    lm = k > 0
    return 810.0 * M + float(np.sum(60 + 144 * k[lm] + 216 * k[lm] ** 2))


# ---- IMU preintegration (vio/imu_factor.py:392-704), per sub-interval and per interval ------------------------------
# FLOPs per sub-interval (an FMA counts 2): the covariance propagation F P F^T dominates: two passes of 15 columns,
# 153 each (rows phi 18 FMA, v 27 FMA + 3, p 27 FMA + 6) = 4,590; symmetrisation 450; the closed-form noise ~725 and
# its addition 225; bias Jacobians 6 columns x 69 = 414; two Exp / Jr evaluations ~120; the midpoint attitude, M, N,
# T, M M^T and the eight blocks ~360; the mean ~75.  Total 6,960.
PREINT_FLOPS_PER_SUB = 6960.0
# Bytes per interval: its samples (stamp 8 + gyro 24 + accel 24 = 56 B each) read once, the 17 parameters, and the
# outputs dR 72, dv 24, dp 24, J 432, C 1,800, info 72 (+ W 1,800 + w_ok 4 + the 301-entry record 2,408 when asked).
PREINT_BYTES_SAMPLE = 56.0
PREINT_BYTES_OUT = 136.0 + 2424.0
PREINT_BYTES_OUT_W_REC = PREINT_BYTES_OUT + 1800.0 + 4.0 + 2408.0


def preint_flops(n_sub):
    return PREINT_FLOPS_PER_SUB * float(np.sum(n_sub))


def preint_bytes(n_samples_read, n_intervals, whiten_and_records=True):
    return PREINT_BYTES_SAMPLE * float(n_samples_read) + n_intervals * (
        PREINT_BYTES_OUT_W_REC if whiten_and_records else PREINT_BYTES_OUT)


def eig_flops(n):
    """Algorithmic FLOPs of one symmetric eigen-decomposition with eigenvectors of order n: the 9 n^3 of the
    standard count (tridiagonal reduction 4/3 n^3, eigenvectors and back-transformation the rest)."""
    return 9.0 * float(n) ** 3


def marg_flops(n_keep, n_marg):
    """Algorithmic FLOPs of one marginalisation of the square-root prior: the Schur complement (the eigen-
    decomposition of the marginalised block and the rank update of the kept one) and the square root of the result
    (one eigen-decomposition of the kept block)."""
    k, m = float(n_keep), float(n_marg)
    return eig_flops(m) + 2.0 * k * k * m + eig_flops(k)
