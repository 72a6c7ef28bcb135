"""NumPy FP64 restatement of the reference's observation-pair lists and run-end global BA.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module to
check or time the product path; the product package (xclean_amd/) never imports it.

Functions and the reference code they restate (paths relative to the reference's src/xclean/):
  build_pairs        backend/ba.py:902-927
  (the global-BA functions follow below with their own citations; pose_factor_terms restates _rot_blocks and
  _weld_blocks of backend/global_ba.py:1904-2078)

Pins (tests/test_gba_oracle.py): build_pairs against a brute-force loop over every landmark's observation pairs.
"""
from __future__ import annotations

import numpy as np


def build_pairs(olm_sorted, nl):
    """backend/ba.py:902-927: (offs int32 (nl + 1), pa int32, pb int32) - every ordered observation pair (a, b) inside
    each landmark's run of the landmark-sorted observation list, landmark-major, a-major then b."""
    olm = np.asarray(olm_sorted, np.int64).ravel()
    offs = np.searchsorted(olm, np.arange(nl + 1), side="left").astype(np.int64)
    c = np.diff(offs)
    c2 = c * c
    total = int(c2.sum())
    if total == 0:
        return offs.astype(np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32)
    owner = np.repeat(np.arange(nl), c2)
    first = np.concatenate([[0], np.cumsum(c2)[:-1]])
    t = np.arange(total) - first[owner]
    pa = offs[owner] + t // c[owner]
    pb = offs[owner] + t % c[owner]
    return offs.astype(np.int32), pa.astype(np.int32), pb.astype(np.int32)


# ---------------------------------------------------------------------------------------------------------------
# run-end global BA in inverse depth: backend/global_ba.py:694-1785 (global_ba_id, the path with every XTEN_GBA_*
# mode unset, no inertial term, no rotation factors, no welds, no scale or offset nuisance states)
# ---------------------------------------------------------------------------------------------------------------
HUBER_K = 1.345             # core/derive.py:47
DISPARITY_NOISE_PX = 0.5    # core/derive.py:60
D = 6
DIRECT_MAX = 30_000         # global_ba.py:1352: above this many unknowns the step comes from the iterative solver


def se3_inv(T):
    """core/lie.py:328-335."""
    T = np.asarray(T, np.float64)
    out = np.zeros_like(T)
    Rt = np.swapaxes(T[..., :3, :3], -1, -2)
    out[..., :3, :3] = Rt
    out[..., :3, 3] = -np.einsum("...ij,...j->...i", Rt, T[..., :3, 3])
    out[..., 3, 3] = 1.0
    return out


def _hat(v):
    v = np.asarray(v, np.float64)
    H = np.zeros(v.shape[:-1] + (3, 3))
    H[..., 0, 1], H[..., 0, 2] = -v[..., 2], v[..., 1]
    H[..., 1, 0], H[..., 1, 2] = v[..., 2], -v[..., 0]
    H[..., 2, 0], H[..., 2, 1] = -v[..., 1], v[..., 0]
    return H


def se3_exp_many(xi):
    """core/lie.py:301-314 per row, through the C oracle's se3_exp."""
    from oracle import oracle as O
    return np.stack([O.se3_exp(x) for x in np.asarray(xi, np.float64).reshape(-1, 6)])


def so3_project_many(R):
    from oracle import oracle as O
    return np.stack([O.so3_project(r) for r in np.asarray(R, np.float64).reshape(-1, 3, 3)])


def _huber_w(a):
    """IRLS weight of a residual magnitude: 1 inside HUBER_K, HUBER_K / |a| outside."""
    a = np.asarray(a, np.float64)
    return np.where(a <= HUBER_K, 1.0, HUBER_K / np.maximum(a, 1e-9))


def gba_geom(T_cw, rho, mbear, host, okf, olm, ouv, fx, cx, cy):
    """global_ba.py:771-780: (r (N,2) px, Pc (N,3) unclamped, Ph (nl,3), T_wc of every host (nl,4,4))."""
    Th = se3_inv(T_cw)[host]
    Ph = mbear / np.maximum(rho, 1e-12)[:, None]
    Pw = np.einsum("lij,lj->li", Th[:, :3, :3], Ph) + Th[:, :3, 3]
    Pc = np.einsum("nij,nj->ni", T_cw[okf, :3, :3], Pw[olm]) + T_cw[okf, :3, 3]
    z = np.maximum(Pc[:, 2], 1e-6)
    r = np.stack([fx * Pc[:, 0] / z + cx, fx * Pc[:, 1] / z + cy], 1) - ouv
    return r, Pc, Ph, Th


def stereo_chi(zc, d_meas, fB, inv_sig):
    """global_ba.py:338-354: the Huber-weighted chi2 of the per-observation disparity rows."""
    if d_meas is None:
        return 0.0
    ok = d_meas > 1e-6
    if not ok.any():
        return 0.0
    r = (1.0 / np.maximum(zc[ok], 1e-6) - d_meas[ok] / fB) * inv_sig
    return float(np.sum(_huber_w(np.abs(r)) * r * r))


def prior_rows(prior, rho, mbear):
    """global_ba.py:540-566 (beta 0, landed whitening): whitened residual (nl,3), unit Jacobian (nl,3), sqrt IRLS
    weight (nl,3; Huber on the inverse-depth row only) and the valid mask."""
    m0, rho0, wbear, invsig, valid = prior
    rp = np.stack([(mbear[:, 0] - m0[:, 0]) * wbear, (mbear[:, 1] - m0[:, 1]) * wbear, (rho - rho0) * invsig], 1)
    J0 = np.stack([wbear, wbear, invsig], 1)
    swp = np.sqrt(_huber_w(np.abs(rp)))
    swp[:, :2] = 1.0
    return rp, J0, swp, valid > 0.5


def so3_log(R):
    """Rotation vector of R (scipy's quaternion-based conversion)."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_matrix(np.asarray(R, np.float64)).as_rotvec()


def jr_inv(phi):
    """Inverse right Jacobian of SO(3): I + [phi]x / 2 + (1/t^2 - (1 + cos t) / (2 t sin t)) [phi]x^2."""
    phi = np.asarray(phi, np.float64)
    t = float(np.linalg.norm(phi))
    K = _hat(phi)
    c = 1.0 / 12.0 + t * t / 720.0 if t < 1e-4 else 1.0 / (t * t) - (1.0 + np.cos(t)) / (2.0 * t * np.sin(t))
    return np.eye(3) + 0.5 * K + c * (K @ K)


def se3_log(T):
    """[rho, phi] of T: phi = Log(R), rho = V(phi)^-1 t (V^-1 = Jr^-1(-phi))."""
    phi = so3_log(T[:3, :3])
    return np.concatenate([jr_inv(-phi) @ T[:3, 3], phi])


def adjoint(T):
    """Ad_T in the [rho, phi] order: [[R, [t]x R], [0, R]]."""
    R, t = T[:3, :3], T[:3, 3]
    A = np.zeros((6, 6))
    A[:3, :3] = R
    A[:3, 3:] = _hat(t) @ R
    A[3:, 3:] = R
    return A


def pose_factor_terms(T_cw, n, rotf=None, rotf_huber=None, welds=None, beta=0.0):
    """global_ba.py:1904-1967 (_rot_blocks) and 1969-2078 (_weld_blocks, beta held): the dense (n, n) Hessian blocks
    and gradient of the gyro rotation factors (a, b, M, sig) and the loop welds (a, b, Z, w)."""
    H = np.zeros((n, n))
    g = np.zeros(n)
    for (a, b, M, sig) in (rotf or []):
        A = T_cw[a][:3, :3] @ T_cw[b][:3, :3].T
        e = so3_log(np.asarray(M, np.float64).T @ A)
        w = 1.0 / max(sig, 1e-9)
        if rotf_huber is not None:
            en = w * float(np.linalg.norm(e))
            if en > rotf_huber:
                w = w * float(np.sqrt(rotf_huber / en))
        r = w * e
        J = jr_inv(e)
        Ja, Jb = w * (J @ A.T), -w * J
        ia, ib = slice(6 * a + 3, 6 * a + 6), slice(6 * b + 3, 6 * b + 6)
        H[ia, ia] += Ja.T @ Ja
        H[ia, ib] += Ja.T @ Jb
        H[ib, ia] += Jb.T @ Ja
        H[ib, ib] += Jb.T @ Jb
        g[ia] -= Ja.T @ r
        g[ib] -= Jb.T @ r
    for (a, b, Z, w) in (welds or []):
        Zb = np.array(Z, np.float64)
        Zb[:3, 3] /= (1.0 + beta)
        Ti, Tj = se3_inv(T_cw[a]), se3_inv(T_cw[b])
        e = se3_log(se3_inv(Zb) @ (se3_inv(Ti) @ Tj))
        W = np.diag(np.asarray(w, np.float64)) if np.ndim(w) else np.eye(6) * float(w)
        Ja, Jb = adjoint(se3_inv(Tj) @ Ti), -np.eye(6)
        for bi, J1 in ((a, Ja), (b, Jb)):
            g[6 * bi:6 * bi + 6] -= J1.T @ (W @ e)
            for bj, J2 in ((a, Ja), (b, Jb)):
                H[6 * bi:6 * bi + 6, 6 * bj:6 * bj + 6] += J1.T @ W @ J2
    return H, g


def gba_system(T_cw, rho, mbear, host, okf, olm, ouv, hk, cam, d_meas, prior, lam, free, vt_weights=None, rotf=None, rotf_huber=None,
               welds=None, beta=0.0):
    """global_ba.py:1003-1197 and 1326-1336: the damped reduced camera system of one LM iteration.

    cam = (fx, cx, cy, fB, inv_sig_s).  -> dict(H (n,n) scipy CSR with the gauge rows identity and the damping on
    its diagonal, bp (n,), Hlli (nl,3,3), bl (nl,3), U (n, 3 nl) CSR, chi (vision + stereo at this state)).
    """
    import scipy.sparse as sp
    fx, cx, cy, fB, inv_sig_s = cam
    nk, nl, N = len(T_cw), len(rho), len(okf)
    n = D * nk
    r, Pc, Ph, Th = gba_geom(T_cw, rho, mbear, host, okf, olm, ouv, fx, cx, cy)
    chi = float(np.sum(r ** 2)) + stereo_chi(Pc[:, 2], d_meas, fB, inv_sig_s)
    w = _huber_w(np.linalg.norm(r, axis=1))
    if vt_weights is not None:
        w = w * np.asarray(vt_weights, np.float64)[okf]
    sw = np.sqrt(w)
    x, y, z = Pc[:, 0], Pc[:, 1], np.maximum(Pc[:, 2], 1e-6)
    A = np.zeros((N, 2, 3))
    A[:, 0, 0] = A[:, 1, 1] = fx / z
    A[:, 0, 2] = -fx * x / z ** 2
    A[:, 1, 2] = -fx * y / z ** 2
    # d Pc / d(observer twist [rho, phi]) = [I | -[Pc]x], with the clamped depth
    Pcz = np.stack([x, y, z], 1)
    dPo = np.concatenate([np.broadcast_to(np.eye(3), (N, 3, 3)), -_hat(Pcz)], 2)
    Rch = np.einsum("nij,njk->nik", T_cw[okf, :3, :3], Th[olm, :3, :3])
    Pho = Ph[olm]
    # d Ph / d(host twist) = [-I | [Ph]x]
    dPh = np.concatenate([-np.broadcast_to(np.eye(3), (N, 3, 3)), _hat(Pho)], 2)
    ARch = A @ Rch
    irho = 1.0 / np.maximum(rho[olm], 1e-12)
    Jo = (A @ dPo) * sw[:, None, None]
    Jh = (ARch @ dPh) * sw[:, None, None]
    dPdl = np.concatenate([Rch[:, :, :2] * irho[:, None, None],
                           -(Rch @ Pho[:, :, None]) * irho[:, None, None]], 2)     # d Pc / d(a, b, rho)
    Jl = (A @ dPdl) * sw[:, None, None]
    rows_o, rows_h, rows_l, res = [Jo], [Jh], [Jl], [r * sw[:, None]]
    obs_of_row = [np.arange(N)]
    if d_meas is not None:
        ok = d_meas > 1e-6
        if ok.any():
            zs = z[ok]
            rs = (1.0 / zs - d_meas[ok] / fB) * inv_sig_s
            sws = np.sqrt(_huber_w(np.abs(rs)))
            cz = (-1.0 / zs ** 2) * inv_sig_s * sws
            rows_o.append((dPo[ok][:, 2:3, :]) * cz[:, None, None])
            rows_h.append(((Rch @ dPh)[ok][:, 2:3, :]) * cz[:, None, None])
            rows_l.append((dPdl[ok][:, 2:3, :]) * cz[:, None, None])
            res.append((rs * sws)[:, None])
            obs_of_row.append(np.flatnonzero(ok))
    Hll = np.zeros((nl, 3, 3))
    bl = np.zeros((nl, 3))
    trip_r, trip_c, trip_v, lrow, lcol, lval, rvec = [], [], [], [], [], [], []
    row0 = 0
    for Jo_, Jh_, Jl_, r_, k_ in zip(rows_o, rows_h, rows_l, res, obs_of_row):
        m = Jo_.shape[1]
        nr = len(k_) * m
        ridx = row0 + np.arange(nr).reshape(len(k_), m)
        lm = olm[k_]
        np.add.at(Hll, lm, np.einsum("nqa,nqb->nab", Jl_, Jl_))
        np.add.at(bl, lm, -np.einsum("nqa,nq->na", Jl_, r_))
        for J_, kf in ((Jo_, okf[k_]), (Jh_, hk[k_])):
            trip_r.append(np.repeat(ridx.reshape(-1), 6))
            trip_c.append((np.repeat(kf, m)[:, None] * D + np.arange(D)).reshape(-1))
            trip_v.append(J_.reshape(-1))
        lrow.append(np.repeat(ridx.reshape(-1), 3))
        lcol.append((np.repeat(lm, m)[:, None] * 3 + np.arange(3)).reshape(-1))
        lval.append(Jl_.reshape(-1))
        rvec.append(r_.reshape(-1))
        row0 += nr
    Jp = sp.csr_matrix((np.concatenate(trip_v), (np.concatenate(trip_r), np.concatenate(trip_c))), shape=(row0, n))
    Jr = sp.csr_matrix((np.concatenate(lval), (np.concatenate(lrow), np.concatenate(lcol))), shape=(row0, 3 * nl))
    rw = np.concatenate(rvec)
    if prior is not None:
        rp, J0, swp, okp = prior_rows(prior, rho, mbear)
        Jpr = J0 * swp
        for c in range(3):
            Hll[okp, c, c] += Jpr[okp, c] ** 2
        bl[okp] -= (Jpr * rp * swp)[okp]
    tr3 = np.trace(Hll, axis1=1, axis2=2) / 3.0
    Hlli = np.linalg.inv(Hll + (lam * np.maximum(tr3, 1e-9))[:, None, None] * np.eye(3))
    Hl_sp = sp.block_diag(list(Hlli), format="csr") if nl else sp.csr_matrix((0, 0))
    Hpp = (Jp.T @ Jp).tocsr()
    U = (Jp.T @ Jr).tocsr()
    H = (Hpp - U @ Hl_sp @ U.T).tolil()
    bp = -np.asarray(Jp.T @ rw).ravel() - U @ np.einsum("lab,lb->la", Hlli, bl).reshape(-1)
    if rotf or welds:
        Hf, gf = pose_factor_terms(T_cw, n, rotf, rotf_huber, welds, beta)
        H = (H.tocsr() + sp.csr_matrix(Hf)).tolil()
        bp = bp + gf
    for i in np.flatnonzero(~np.asarray(free, bool)):
        s = slice(D * i, D * i + D)
        H[s, :] = 0.0
        H[:, s] = 0.0
        for q in range(D):
            H[D * i + q, D * i + q] = 1.0
        bp[s] = 0.0
    H = H.tocsr()
    H = H + sp.diags(lam * np.maximum(H.diagonal(), 1e-9))
    return {"H": H.tocsr(), "bp": bp, "Hlli": Hlli, "bl": bl, "U": U, "chi": chi, "r": r}


# ---------------------------------------------------------------------------------------------------------------
# two-level deflated PCG: backend/gba_linsolve.py:129-478 (visual-only layout: 6 unknowns per keyframe, no global
# unknowns, no velocity modes)
# ---------------------------------------------------------------------------------------------------------------
RTOL = 1e-8            # gba_linsolve.py:57
MAXITER = 2000         # gba_linsolve.py:62
CHECK_EVERY = 50       # gba_linsolve.py:65
STALL_CHECKS = 4       # gba_linsolve.py:70


def segment_length(nk, b=6, c=7):
    """gba_linsolve.py:147-149: keyframes per segment minimising b^2 m nk + c^2 (nk/m)^2, at least 2."""
    return max(2, int(round((2.0 * c * c * nk / float(b * b)) ** (1.0 / 3.0))))


def segment_bounds(nk, m):
    """gba_linsolve.py:152-154."""
    nseg = max(1, int(round(nk / float(m))))
    return np.unique(np.linspace(0, nk, nseg + 1).round().astype(np.int64))


def mode_vectors(T_cw, free):
    """gba_linsolve.py:249-270 (no velocity): per keyframe (6, 7) - columns translation x3, rotation x3, scale; rows
    [rho, phi] of the left-multiplied twist; rows of fixed keyframes zero."""
    R, t = T_cw[:, :3, :3], T_cw[:, :3, 3]
    M = np.zeros((len(T_cw), 6, 7))
    M[:, :3, :3] = -R
    M[:, :3, 3:6] = -_hat(t) @ R
    M[:, 3:6, 3:6] = -R
    M[:, :3, 6] = t
    M[~np.asarray(free, bool)] = 0.0
    return M


def deflation_basis(T_cw, free, bounds):
    """gba_linsolve.py:273-302: per segment an orthonormal basis (rank-revealing SVD) of its keyframes' mode
    vectors.  -> list of (k0, k1, Z_s (6 (k1 - k0), r_s))."""
    M = mode_vectors(T_cw, free)
    out = []
    eps = np.finfo(float).eps
    for s in range(len(bounds) - 1):
        k0, k1 = int(bounds[s]), int(bounds[s + 1])
        Ms = M[k0:k1].reshape(-1, 7)
        nz = np.abs(Ms).sum(1) > 0
        Z = np.zeros((len(Ms), 0))
        if nz.any():
            U, S, _ = np.linalg.svd(Ms[nz], full_matrices=False)
            r = int(np.sum(S > S[0] * max(Ms[nz].shape) * eps)) if S[0] > 0 else 0
            Z = np.zeros((len(Ms), r))
            Z[nz] = U[:, :r]
        out.append((k0, k1, Z))
    return out


def dpcg(H, bp, T_cw, free, rtol=RTOL, maxiter=MAXITER, seg_len=None):
    """gba_linsolve.py:353-478 for the visual-only system: (dx, info) with info iters / matvecs / restarts /
    converged / reason / relres / seg_len / segments / nc.  H: scipy sparse SPD (n, n), n = 6 nk."""
    import scipy.linalg as sla
    H = H.tocsr()
    n = H.shape[0]
    nk = n // D
    m = int(seg_len) if seg_len else segment_length(nk)
    bounds = segment_bounds(nk, m)
    Hd = H.toarray() if n <= 20000 else None
    # smoother: exact inverses of the segment blocks (gba_linsolve.py:160-235)
    Binv = []
    for s in range(len(bounds) - 1):
        a, b = D * int(bounds[s]), D * int(bounds[s + 1])
        B = H[a:b, a:b].toarray()
        B = 0.5 * (B + B.T)
        L = np.linalg.cholesky(B)
        Li = np.linalg.inv(L)
        Binv.append((a, b, Li.T @ Li))

    def Minv(r):
        z = np.empty_like(r)
        for a, b, Bi in Binv:
            z[a:b] = Bi @ r[a:b]
        return z

    # deflation space and the coarse operator (gba_linsolve.py:273-350)
    basis = deflation_basis(T_cw, free, bounds)
    cols = []
    for k0, k1, Zs in basis:
        for j in range(Zs.shape[1]):
            v = np.zeros(n)
            v[D * k0:D * k1] = Zs[:, j]
            cols.append(v)
    Z = np.stack(cols, 1) if cols else np.zeros((n, 0))
    HZ = H @ Z
    E = Z.T @ HZ
    E = 0.5 * (E + E.T)
    d = np.sqrt(np.maximum(np.diag(E), 1e-300))
    cf = sla.cho_factor(E / d[:, None] / d[None, :], lower=True)

    def csolve(y):
        return sla.cho_solve(cf, y / d) / d

    def coarse_part(r):
        return Z @ csolve(Z.T @ r)

    def project(zv):
        return zv - Z @ csolve(HZ.T @ zv)

    def mv(v):
        return H @ v

    info = {"iters": 0, "matvecs": 0, "restarts": 0, "reason": None, "seg_len": m, "segments": len(bounds) - 1,
            "nc": int(Z.shape[1])}
    b = np.asarray(bp, np.float64)
    nb = float(np.linalg.norm(b))
    if nb == 0.0 or not np.isfinite(nb):
        ok = nb == 0.0
        info.update(converged=ok, relres=0.0 if ok else float("nan"), reason="zero_rhs" if ok else "nonfinite_rhs")
        return np.zeros_like(b), info
    tol = rtol * nb
    x = coarse_part(b)
    r = b - mv(x)
    info["matvecs"] += 1
    best = float(np.linalg.norm(r))
    if best <= tol:
        info.update(converged=True, relres=best / nb, reason="tol")
        return x, info
    x_best = x.copy()
    stall = 0
    it = 0
    z = Minv(r)
    p = project(z)
    rz = float(r @ z)
    while True:
        if it >= maxiter:
            info.update(iters=it, converged=False, relres=best / nb, reason="maxiter")
            return x_best, info
        w = mv(p)
        info["matvecs"] += 1
        it += 1
        pw = float(p @ w)
        if not (pw > 0.0) or not np.isfinite(pw):
            info.update(iters=it, converged=False, relres=best / nb, reason="breakdown")
            return x_best, info
        alpha = rz / pw
        x = x + alpha * p
        r = r - alpha * w
        rn = float(np.linalg.norm(r))
        if rn <= tol or rn <= 0.1 * best or it % CHECK_EVERY == 0:
            rt = b - mv(x)
            info["matvecs"] += 1
            true = float(np.linalg.norm(rt))
            if true <= tol:
                info.update(iters=it, converged=True, relres=true / nb, reason="tol")
                return x, info
            if true < best:
                best, stall = true, 0
                x_best = x.copy()
            else:
                stall += 1
            if true > 10.0 * best:
                info.update(iters=it, converged=False, relres=best / nb, reason="divergence")
                return x_best, info
            if stall >= STALL_CHECKS:
                info.update(iters=it, converged=False, relres=best / nb, reason="stagnation")
                return x_best, info
            if rn <= tol or true > 10.0 * rn:
                info["restarts"] += 1
                x = x + coarse_part(rt)
                r = b - mv(x)
                info["matvecs"] += 1
                z = Minv(r)
                p = project(z)
                rz = float(r @ z)
                continue
        z = Minv(r)
        rz_new = float(r @ z)
        beta = rz_new / rz
        rz = rz_new
        p = project(z) + beta * p


# ---------------------------------------------------------------------------------------------------------------
# the LM driver: global_ba.py:694-1785
# ---------------------------------------------------------------------------------------------------------------
def marginal_cov(H, nk, want, free):
    """global_ba.py:2080-2126: the 6x6 diagonal blocks of H^-1 (H the last damped reduced system) for want free
    keyframes sampled evenly, symmetrised; {} above DIRECT_MAX unknowns."""
    import scipy.sparse.linalg as spl
    ids = [i for i in range(nk) if free[i]]
    if not ids or H.shape[0] > DIRECT_MAX:
        return {}
    step = max(1, len(ids) // max(want, 1))
    ids = ids[::step][:want]
    lu = spl.splu(H.tocsc())
    out = {}
    for i in ids:
        C = np.zeros((D, D))
        for c in range(D):
            e = np.zeros(H.shape[0])
            e[D * i + c] = 1.0
            C[:, c] = lu.solve(e)[D * i:D * i + D]
        if np.isfinite(C).all():
            out[i] = 0.5 * (C + C.T)
    return out


def global_ba_id(kf_ids, poses, rho, mbear, host_idx, okf, olm, ouv, rig, odisp=None, prior=None, rotf=None,
                 welds=None, iters=8, fixed=(0,), lam0=1e-4, passes=2, reject_sigma=3.0, want_cov=0, rho_floor=1e-4,
                 rotf_huber=None, vt_weights=None, solver="direct", history=None):
    """global_ba.py:694-1785, the default path.  poses: (nk,4,4) T_wc; -> (poses T_wc, rho, mbear, stats).

    solver "direct" solves every damped reduced system exactly (scipy spsolve, the reference's path up to 30,000
    unknowns); "dpcg" uses the deflated PCG above at rtol RTOL.  history, a list, receives per LM iteration
    dict(lam, chi, chi2, accepted, dx)."""
    import scipy.sparse.linalg as spl
    nk, nl = len(poses), len(rho)
    T_cw = np.stack([se3_inv(T) for T in poses])
    rho = np.array(rho, np.float64)
    mbear = np.array(mbear, np.float64)
    host_idx = np.asarray(host_idx, np.int64)
    free = np.ones(nk, bool)
    for f in fixed:
        if 0 <= f < nk:
            free[f] = False
    okf = np.asarray(okf, np.int64)
    olm = np.asarray(olm, np.int64)
    ouv = np.asarray(ouv, np.float64)
    hk = host_idx[olm]
    keep0 = hk != okf                      # a landmark's rows at its own host carry no pose information
    okf, olm, ouv, hk = okf[keep0], olm[keep0], ouv[keep0], hk[keep0]
    N = len(okf)
    fx, cx, cy = float(rig.fx_rect), float(rig.cx), float(rig.cy)
    d_meas = None if odisp is None else np.asarray(odisp, np.float64)[keep0]
    fB = float(getattr(rig, "fB", 0.0) or 0.0) if d_meas is not None else 0.0
    inv_sig_s = (1.0 / max(float(getattr(rig, "rho_sigma", 0.0) or 1e-9), 1e-12) if d_meas is not None else 0.0)
    if d_meas is not None and fB <= 0.0:
        d_meas = None
    cam = (fx, cx, cy, fB, inv_sig_s)
    r0, Pc0, _, _ = gba_geom(T_cw, rho, mbear, host_idx, okf, olm, ouv, fx, cx, cy)
    rms0 = float(np.sqrt(np.mean(np.sum(r0 ** 2, 1)))) if N else float("nan")
    chi0_vision = float(np.sum(r0 ** 2))
    chi0_stereo = stereo_chi(Pc0[:, 2], d_meas, fB, inv_sig_s)
    lam = lam0
    last_H = None
    best = (T_cw.copy(), rho.copy(), mbear.copy(), chi0_vision)
    n_acc = n_best = 0
    n_stereo = int((d_meas > 1e-6).sum()) if d_meas is not None else 0
    for it in range(iters):
        sysm = gba_system(T_cw, rho, mbear, host_idx, okf, olm, ouv, hk, cam, d_meas, prior, lam, free, vt_weights,
                          rotf=rotf, rotf_huber=rotf_huber, welds=welds)
        chi = sysm["chi"]
        H, bp = sysm["H"], sysm["bp"]
        last_H = H
        if solver == "dpcg" or H.shape[0] > DIRECT_MAX:
            dx, info = dpcg(H, bp, T_cw, free)
            if not info["converged"]:
                lam = min(lam * 10, 1e3)
                if history is not None:
                    history.append({"lam": lam, "chi": chi, "chi2": None, "accepted": False, "dx": None})
                continue
        else:
            dx = spl.spsolve(H.tocsc(), bp)
        if not np.isfinite(dx).all():
            lam = min(lam * 10, 1e3)
            continue
        dxl = np.einsum("lab,lb->la", sysm["Hlli"], sysm["bl"] - np.asarray(sysm["U"].T @ dx).reshape(nl, 3))
        T_new = T_cw.copy()
        if free.any():
            T_new[free] = se3_exp_many(dx.reshape(nk, D)[free]) @ T_cw[free]
        m_new = mbear.copy()
        m_new[:, :2] += dxl[:, :2]
        rho_new = np.maximum(rho + dxl[:, 2], rho_floor)
        r2, Pc2, _, _ = gba_geom(T_new, rho_new, m_new, host_idx, okf, olm, ouv, fx, cx, cy)
        chi2 = float(np.sum(r2 ** 2)) + stereo_chi(Pc2[:, 2], d_meas, fB, inv_sig_s)
        acc = chi2 <= chi
        if history is not None:
            history.append({"lam": lam, "chi": chi, "chi2": chi2, "accepted": bool(acc), "dx": dx.copy()})
        if acc:
            n_acc += 1
            T_cw, rho, mbear = T_new, rho_new, m_new
            lam = max(lam * 0.4, 1e-7)
            if chi2 < best[3]:
                n_best += 1
                best = (T_cw.copy(), rho.copy(), m_new.copy(), chi2)
        else:
            lam = min(lam * 5.0, 1e3)
    T_cw, rho, mbear = best[0], best[1], best[2]
    lm = {"chi0_vision": chi0_vision, "chi0_stereo": chi0_stereo, "chi_best": float(best[3]), "accepted": n_acc,
          "best_updates": n_best, "iters_allowed": int(iters)}
    n_rej, thr = 0, None
    if passes > 1:
        rr, _, _, _ = gba_geom(T_cw, rho, mbear, host_idx, okf, olm, ouv, fx, cx, cy)
        e = np.linalg.norm(rr, axis=1)
        med = float(np.median(e))
        mad = float(np.median(np.abs(e - med))) * 1.4826
        thr = med + reject_sigma * max(mad, 1e-6)
        keep = e <= thr
        n_rej = int((~keep).sum())
        first_sigma = None
        prior_rw = prior
        if prior is not None:
            m0P, rho0P, wbP, invsigP, validP = prior
            okp = validP > 0.5
            if okp.sum() > 50:
                rp = ((rho - rho0P) * invsigP)[okp]
                spread = float(np.median(np.abs(rp)) * 1.4826)
                first_sigma = spread * DISPARITY_NOISE_PX
                if 0.05 < spread < 20.0:
                    prior_rw = (m0P, rho0P, wbP, invsigP / spread, validP)
        if n_rej and keep.sum() > 0.5 * N:
            # pass 2 re-solves the kept rows from pass 1's state: the vision rows only (no disparity rows)
            p2, rho2, m2, st = global_ba_id(kf_ids, se3_inv(T_cw), rho, mbear, host_idx, okf[keep], olm[keep],
                                            ouv[keep], rig, prior=prior_rw, rotf=rotf, welds=welds, iters=iters,
                                            fixed=fixed, lam0=lam0, passes=1, want_cov=want_cov, rho_floor=rho_floor,
                                            rotf_huber=rotf_huber, vt_weights=vt_weights, solver=solver,
                                            history=history)
            st.update(stereo_rows=n_stereo, rms_px_before=rms0, observations=N, rejected=n_rej,
                      rejected_pct=100.0 * n_rej / max(N, 1), reject_threshold_px=thr, lm_pass1=lm)
            if first_sigma is not None:
                st["measured_disparity_sigma_px"] = first_sigma
                st["measured_sigma_is_first_pass"] = True
            return p2, rho2, m2, st
    T_cw = T_cw.copy()
    T_cw[:, :3, :3] = so3_project_many(T_cw[:, :3, :3])
    rn, _, _, _ = gba_geom(T_cw, rho, mbear, host_idx, okf, olm, ouv, fx, cx, cy)
    sig = bias = None
    if prior is not None:
        m0P, rho0P, wbP, invsigP, validP = prior
        okp = validP > 0.5
        if okp.any():
            rp = ((rho - rho0P) * invsigP)[okp]
            sig = float(np.median(np.abs(rp)) * 1.4826 * DISPARITY_NOISE_PX)
            bias = float(np.median(rp) * DISPARITY_NOISE_PX)
    cov = marginal_cov(last_H, nk, want_cov, free) if (want_cov and last_H is not None) else {}
    st = {"rms_px_before": rms0, "rms_px_after": float(np.sqrt(np.mean(np.sum(rn ** 2, 1)))) if N else float("nan"),
          "cov": cov,
          "keyframes": nk, "landmarks": nl, "observations": N, "rejected": n_rej,
          "rejected_pct": 100.0 * n_rej / max(N, 1), "stereo_rows": n_stereo, "lm": lm,
          "measured_disparity_sigma_px": sig, "measured_disparity_bias_px": bias}
    return se3_inv(T_cw), rho, mbear, st
