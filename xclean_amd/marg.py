"""Marginalisation of the square-root prior on the GPU (vio/marg.py's `schur`, `Prior.from_info`, `combine`,
`drop_slot`, `shrink_slot`, `add_block_info`; vio/window.py:1011-1028 `marg_system`), batched over matrices.

    from xclean_amd import marg
    H, b, off = marg.marg_system(pb, lm_mask)                          # factors touching what is marginalised
    H2, b2 = marg.schur(H, b, keep, drop)                              # the kept block's Schur complement
    prior = marg.combine(prior_old, sids, dims, states, H2, b2)       # FEJ square-root prior

The eigen-decompositions (H_mm's floored inverse, the square root S = diag(sqrt l) V^T) and the re-expression of b
at the first linearisation points run in `k_marg_schur` / `k_prior_from_info` (csrc/xba_marg.hip, FP64 cyclic
Jacobi, one workgroup per matrix); `schur_batch` / `from_info_batch` take many matrices per launch.  Inputs and
outputs are host arrays, as the reference's are; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .problem import Prior

EIG_FLOOR = 1e-12  # vio/marg.py:32


class XbaSchurBatch(C.Structure):
    vp = C.c_void_p
    _fields_ = [("nb", C.c_int32), ("ld", C.c_int32), ("nkeep", vp), ("nmarg", vp), ("keep", vp), ("marg", vp),
                ("H", vp), ("b", vp), ("H2", vp), ("b2", vp), ("work", vp)]


class XbaPriorBatch(C.Structure):
    vp = C.c_void_p
    _fields_ = [("nb", C.c_int32), ("ld", C.c_int32), ("n", vp), ("H", vp), ("b", vp), ("blk0", vp), ("boff", vp),
                ("bdim", vp), ("bcur", vp), ("bx0", vp), ("S", vp), ("r", vp), ("rank", vp), ("lam", vp),
                ("bx", vp), ("work", vp)]


def _torch():
    import torch
    _lib.require_gpu()
    return torch, torch.device("cuda", torch.cuda.current_device())


def _ld(sizes):
    ld = max([1] + [int(s) for s in sizes])
    if ld > 1024:
        raise _lib.XbaError("marginalisation: matrices larger than XBA_MAX_STATE (1024)")
    return ld


def _slab(mats, ld, torch, dev):
    out = np.zeros((len(mats), ld, ld))
    for w, A in enumerate(mats):
        A = np.asarray(A, np.float64)
        out[w, :A.shape[0], :A.shape[1]] = A
    return torch.as_tensor(out, device=dev).contiguous()


def _vecs(vs, ld, torch, dev, dtype=np.float64):
    out = np.zeros((len(vs), ld), dtype)
    for w, v in enumerate(vs):
        v = np.asarray(v).reshape(-1)
        out[w, :len(v)] = v
    return torch.as_tensor(out, device=dev).contiguous()


def schur_batch(Hs, bs, keeps, margs):
    """vio/marg.py:165-177 for many (H, b, keep, marg) at once: [(H2, b2)]."""
    torch, dev = _torch()
    nb = len(Hs)
    if nb == 0:
        return []
    ns = [np.asarray(H).shape[0] for H in Hs]
    ld = _ld(ns)
    keeps = [np.asarray(k, np.int64).reshape(-1) for k in keeps]
    margs = [np.asarray(m, np.int64).reshape(-1) for m in margs]
    for n, k, m in zip(ns, keeps, margs):
        if len(k) + len(m) > n or (len(k) and (k.min() < 0 or k.max() >= n)) or (len(m) and (m.min() < 0 or
                                                                                         m.max() >= n)):
            raise _lib.XbaError("schur: keep / marg indices out of range")
    H = _slab(Hs, ld, torch, dev)
    b = _vecs(bs, ld, torch, dev)
    kp = _vecs(keeps, ld, torch, dev, np.int32)
    mp = _vecs(margs, ld, torch, dev, np.int32)
    nk = torch.as_tensor(np.array([len(k) for k in keeps], np.int32), device=dev)
    nm = torch.as_tensor(np.array([len(m) for m in margs], np.int32), device=dev)
    H2, b2 = schur_dev(H, b, kp, mp, nk, nm)
    H2h, b2h = H2.cpu().numpy(), b2.cpu().numpy()
    return [(H2h[w, :len(k), :len(k)].copy(), b2h[w, :len(k)].copy()) for w, k in enumerate(keeps)]


def schur_dev(H, b, kp, mp, nk, nm, work=None):
    """One `xba_marg_schur` launch on device tensors: H (nb, ld, ld), b (nb, ld) FP64, kp / mp (nb, ld) int32 index
    lists with nk / nm (nb,) int32 entries.  Returns (H2, b2) device tensors; asynchronous on torch's current stream."""
    import torch
    nb, ld = int(H.shape[0]), int(H.shape[1])
    H2 = torch.zeros_like(H)
    b2 = torch.zeros_like(b)
    if work is None:
        work = torch.empty(nb * (3 * ld * ld + 2 * ld), dtype=torch.float64, device=H.device)
    q = XbaSchurBatch(nb, ld, nk.data_ptr(), nm.data_ptr(), kp.data_ptr(), mp.data_ptr(), H.data_ptr(), b.data_ptr(),
                      H2.data_ptr(), b2.data_ptr(), work.data_ptr())
    st = torch.cuda.current_stream(H.device).cuda_stream
    _lib.check(_lib.load().xba_marg_schur(C.cast(C.byref(q), C.c_void_p), C.c_void_p(st)), "xba_marg_schur")
    return H2, b2


def from_info_dev(H, b, n, blocks=None, want_bx=False, work=None):
    """One `xba_prior_from_info` launch on device tensors: H (nb, ld, ld), b (nb, ld), n (nb,) int32; blocks None or
    (blk0, boff, bdim, bcur, bx0) device tensors.  Returns (S, r, rank, lam, bx or None); asynchronous."""
    import torch
    nb, ld = int(H.shape[0]), int(H.shape[1])
    S = torch.zeros_like(H)
    r = torch.zeros_like(b)
    rank = torch.zeros(nb, dtype=torch.int32, device=H.device)
    lam = torch.zeros_like(b)
    bx = torch.zeros_like(b) if want_bx else None
    if work is None:
        work = torch.empty(nb * (3 * ld * ld + 2 * ld), dtype=torch.float64, device=H.device)
    ptr = lambda i: blocks[i].data_ptr() if blocks is not None else None
    q = XbaPriorBatch(nb, ld, n.data_ptr(), H.data_ptr(), b.data_ptr(), ptr(0), ptr(1), ptr(2), ptr(3), ptr(4),
                      S.data_ptr(), r.data_ptr(), rank.data_ptr(), lam.data_ptr(),
                      bx.data_ptr() if want_bx else None, work.data_ptr())
    st = torch.cuda.current_stream(H.device).cuda_stream
    _lib.check(_lib.load().xba_prior_from_info(C.cast(C.byref(q), C.c_void_p), C.c_void_p(st)),
               "xba_prior_from_info")
    return S, r, rank, lam, bx


def schur(H, b, keep, marg):
    """vio/marg.py:165-177: (H2, b2) of the kept indices (in keep's order)."""
    return schur_batch([H], [b], [keep], [marg])[0]


def _state21(st):
    R, p, v, bg, ba = (tuple(st) + (None,) * 5)[:5]
    z = np.zeros(3)
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(p, float).reshape(3),
                           *[np.asarray(z if x is None else x, float).reshape(3) for x in (v, bg, ba)]])


def from_info_batch(items, want_bx=False):
    """Many square roots at once.  items: [(H, b) or (H, b, blocks)] with blocks = [(offset, dim, state, x0)]
    (states as (R, p, v, bg, ba)) to re-express b at x0 first (combine).  Returns [(S (rank, n), r (rank,),
    eigenvalues ascending (n,))], with b re-expressed at x0 appended to each tuple when `want_bx`."""
    torch, dev = _torch()
    nb = len(items)
    if nb == 0:
        return []
    ns = [np.asarray(it[0]).shape[0] for it in items]
    ld = _ld(ns)
    H = _slab([it[0] for it in items], ld, torch, dev)
    b = _vecs([it[1] for it in items], ld, torch, dev)
    blk0 = [0]
    boff, bdim, bcur, bx0 = [], [], [], []
    for it in items:
        for (o, d, st, x0) in (it[2] if len(it) > 2 and it[2] is not None else []):
            boff.append(int(o))
            bdim.append(int(d))
            bcur.append(_state21(st))
            bx0.append(_state21(x0))
        blk0.append(len(boff))
    with_blocks = len(boff) > 0
    T = lambda x, dt: torch.as_tensor(np.asarray(x, dt), device=dev).contiguous()
    n_t = T(ns, np.int32)
    blocks = None
    if with_blocks:
        blocks = (T(blk0, np.int32), T(boff, np.int32), T(bdim, np.int32), T(np.stack(bcur), np.float64),
                  T(np.stack(bx0), np.float64))
    S, r, rank, lam, bx = from_info_dev(H, b, n_t, blocks, want_bx)
    Sh, rh, kh, lh = S.cpu().numpy(), r.cpu().numpy(), rank.cpu().numpy(), lam.cpu().numpy()
    out = [(Sh[w, :int(kh[w]), :n].copy(), rh[w, :int(kh[w])].copy(), lh[w, :n].copy()) for w, n in enumerate(ns)]
    if want_bx:
        bxh = bx.cpu().numpy()
        out = [o + (bxh[w, :n].copy(),) for w, (o, n) in enumerate(zip(out, ns))]
    return out


def prior_from_info(sids, dims, x0, H, b):
    """vio/marg.py:73-84 `Prior.from_info`."""
    S, r, _lam = from_info_batch([(H, b)])[0]
    return Prior(sids, dims, x0, S, r)


def info(prior):
    """vio/marg.py:70-71: (S^T S, S^T r) at x0."""
    torch, dev = _torch()
    S = torch.as_tensor(np.asarray(prior.S, np.float64), device=dev)
    r = torch.as_tensor(np.asarray(prior.r, np.float64), device=dev)
    return (S.T @ S).cpu().numpy(), (S.T @ r).cpu().numpy()


def _x0_of(st):
    R, p, v, bg, ba = (tuple(st) + (None,) * 5)[:5]
    z = np.zeros(3)
    return (np.array(R, float), np.array(p, float), np.array(z if v is None else v, float),
            np.array(z if bg is None else bg, float), np.array(z if ba is None else ba, float))


def combine(prior_old, sids_new, dims_new, states_cur, H_cur, b_cur):
    """vio/marg.py:212-242: the prior over (sids_new, dims_new) from (H_cur, b_cur) linearised at the current states.
    Blocks already in the old prior keep their x0 (FEJ; a pose block that grew to 15 dims keeps its pose x0 and takes
    the current v, bg, ba) and b is re-expressed there (b - H d, on the GPU); new blocks start at the current state."""
    x0, blocks, o = [], [], 0
    for s, dim, st in zip(sids_new, dims_new, states_cur):
        cur = _x0_of(st)
        if prior_old is not None and s in prior_old.sids:
            j = prior_old.sids.index(s)
            old = prior_old.x0[j]
            xb = old if (prior_old.dims[j] == 15 or dim == 6) else (old[0], old[1], cur[2], cur[3], cur[4])
            x0.append(xb)
            blocks.append((o, dim, cur, xb))
        else:
            x0.append(cur)
        o += int(dim)
    S, r, _lam = from_info_batch([(H_cur, b_cur, blocks)])[0]
    return Prior(sids_new, dims_new, x0, S, r)


def _without(prior, k, dims_marg):
    H, b = info(prior)
    bo = np.concatenate([[0], np.cumsum(prior.dims)]).astype(np.int64)
    marg = np.array([bo[k] + j for j in dims_marg], np.int64)
    ms = set(marg.tolist())
    keep = np.array([i for i in range(prior.n) if i not in ms], np.int64)
    return schur(H, b, keep, marg)


def drop_slot(prior, sid):
    """vio/marg.py:136-147: marginalise one slot's whole block out of the prior (None when nothing is left)."""
    if sid not in prior.sids:
        return prior
    k = prior.sids.index(sid)
    H2, b2 = _without(prior, k, range(prior.dims[k]))
    sids = [s for i, s in enumerate(prior.sids) if i != k]
    if not sids:
        return None
    return prior_from_info(sids, [d for i, d in enumerate(prior.dims) if i != k],
                           [x for i, x in enumerate(prior.x0) if i != k], H2, b2)


def shrink_slot(prior, sid):
    """vio/marg.py:149-159: marginalise the (v, bg, ba) dims of a full block; the block becomes a pose block."""
    if sid not in prior.sids:
        return prior
    k = prior.sids.index(sid)
    if prior.dims[k] == 6:
        return prior
    H2, b2 = _without(prior, k, range(6, 15))
    dims = list(prior.dims)
    dims[k] = 6
    return prior_from_info(prior.sids, dims, prior.x0, H2, b2)


def add_block_info(prior, sid, dim, st, Hb, bb):
    """vio/marg.py:212-245 `add_block_info`: merge independent information (Hb, bb at the current state st, over
    block sid's first `dim` dims) into the prior."""
    if prior is None:
        return prior_from_info([sid], [dim], [st], Hb, bb)
    H0, b0 = info(prior)
    sids, dims, x0 = list(prior.sids), list(prior.dims), list(prior.x0)
    if sid in sids:
        k = sids.index(sid)
        if dims[k] < dim:
            bo = np.concatenate([[0], np.cumsum(dims)]).astype(np.int64)
            ins, pad = bo[k] + dims[k], dim - dims[k]
            H0 = np.insert(np.insert(H0, [ins] * pad, 0.0, axis=0), [ins] * pad, 0.0, axis=1)
            b0 = np.insert(b0, [ins] * pad, 0.0)
            dims[k] = dim
            cur = _x0_of(st)
            x0[k] = (x0[k][0], x0[k][1], cur[2], cur[3], cur[4])
        bo = np.concatenate([[0], np.cumsum(dims)]).astype(np.int64)
        cols = np.arange(bo[k], bo[k] + dim)
        # (Hb, bb) at the current state, expressed at the block's x0: bb - Hb d (d = st [-] x0 of the block)
        Hloc = np.asarray(Hb, float)
        bloc = from_info_batch([(Hloc, np.asarray(bb, float), [(0, dim, _x0_of(st), x0[k])])], want_bx=True)[0][3]
        H0[np.ix_(cols, cols)] += Hloc
        b0[cols] += bloc
        return prior_from_info(sids, dims, x0, H0, b0)
    n0 = H0.shape[0]
    H = np.zeros((n0 + dim, n0 + dim))
    H[:n0, :n0] = H0
    H[n0:, n0:] = Hb
    return prior_from_info(sids + [sid], dims + [dim], x0 + [_x0_of(st)], H, np.concatenate([b0, bb]))


def marg_system(pb, lm_mask, precision="fp64"):
    """vio/window.py:1011-1028: (H, b, off) at the current estimate of the factors touching what is marginalised: the
    visual rows of the masked landmarks (Schur'd out), the prior and the IMU factors."""
    from .window import final_linearisation
    q = pb.copy()
    q.lm_mask = np.ascontiguousarray(lm_mask, np.bool_)
    H, b, _chi2, _ok, _Hi, _imu, off = final_linearisation(q, precision=precision)
    return H, b, off
