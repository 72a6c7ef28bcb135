"""Window packing from engine state (vio/engine_live.py:407-553 _problem / _writeback): round trips and the
packer's rules on hand-made engine states."""
import types

import numpy as np
import pytest

from xclean_amd import pack, synth
from xclean_amd.problem import gauge_prior


def _engine_from_window(pb, seed=0, n_dead=7, kappa_d=1.5, sig_px=1.0, shuffle=False):
    """Scatter a window's CSR rows back into per-keyframe observation lists, with dead landmarks in between and
    keyframe ids that are not slot indices; landmark ids ascending with the window's landmark order unless
    ``shuffle``."""
    rng = np.random.default_rng(seed)
    P, L = pb.P, pb.L
    n_lm = L + n_dead
    lid_of = rng.permutation(n_lm)[:L]            # landmark position -> engine id
    if not shuffle:
        lid_of = np.sort(lid_of)
    alive = np.zeros(n_lm + 3, bool)
    alive[lid_of] = True
    sids = list(100 + 3 * np.arange(P))
    lm_of_row = np.repeat(np.arange(L), np.diff(pb.ob_start))
    slots = []
    for i in range(P):
        rows = np.flatnonzero(pb.ob_pose == i)
        # a right-camera row's sigma s is recovered from its whitening 1 / max(kappa_d s, 1e-9)
        sig = np.where(pb.ob_cam[rows] == 0, 0.0, 1.0 / (kappa_d * pb.ob_isig[rows]))
        s = {"sid": sids[i], "R": pb.Rwb[i].copy(), "p": pb.pwb[i].copy(), "full": pb.dims is not None and
             pb.dims[i] == 15, "lid": lid_of[lm_of_row[rows]], "uv": pb.ob_z[rows].copy(),
             "cam": pb.ob_cam[rows].astype(np.int64), "valid": pb.ob_valid[rows].copy(), "sig": sig}
        if pb.vel is not None:
            s.update(v=pb.vel[i].copy(), bg=pb.bgs[i].copy(), ba=pb.bas[i].copy())
        slots.append(s)
    lm_p = rng.normal(size=(n_lm + 3, 3))
    lm_p[lid_of] = pb.lmp
    lm_host = np.zeros(n_lm + 3, np.int64)
    lm_host[lid_of] = np.array(sids)[pb.lm_host]
    fB = pb.fx * pb.B
    rho_meas = np.zeros(n_lm + 3)
    rho_meas[lid_of] = pb.rho_meas
    # sigma of rho = max(kappa_d sig_lk, span) / fB: put it all in sig_lk
    sig_lk = np.full(n_lm + 3, 1.0)
    ok = pb.rho_isig > 0
    sig_lk[lid_of[ok]] = 1.0 / (pb.rho_isig[ok] * kappa_d) * fB
    span = np.zeros(n_lm + 3)
    eng = types.SimpleNamespace(
        slots=slots, lm_alive=alive, n_lm=n_lm, lm_host=lm_host, lm_p=lm_p,
        stereo_mode="prior" if not np.any(pb.ob_cam == 1) else "ur", lm_rho_meas=rho_meas, lm_sig_lk=sig_lk,
        lm_span=span, kappa_d=kappa_d, fB=fB, sig_px=sig_px, vi=pb.dims is not None, prior=None,
        imu_factors=[{"i": sids[i], "j": sids[j], "pre": pre, "W": W} for (i, j, pre, W) in (pb.imu or [])],
        Rbc=pb.Rbc, pbc=pb.pbc, fx=pb.fx, fy=pb.fy, cx=pb.cx, cy=pb.cy, B=pb.B, g_w=pb.g_w)
    if pb.prior is not None:
        pr = pb.prior.copy()
        pr.sids = [sids[i] for i in pb.prior.slots()]
        eng.prior = pr
    return eng, lid_of


def _same_rows(a, b):
    for k in ("ob_start", "ob_pose", "ob_cam", "ob_valid", "lm_host"):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k
    for k in ("ob_z", "ob_isig", "lmp", "rho_meas", "rho_isig", "Rwb", "pwb"):
        assert np.allclose(getattr(a, k), getattr(b, k), rtol=1e-14, atol=0), k


@pytest.mark.parametrize("vi,stereo", [(False, "prior"), (True, "prior"), (False, "ur")])
def test_pack_round_trip_reproduces_the_window(vi, stereo):
    pb, _ = synth.make_window(P=7, L=300, views=(2, 7), seed=[5, int(vi)], vi=vi, stereo=stereo)
    pb.ob_isig[pb.ob_cam == 0] = 1.0   # the engine whitens every left row with one sigma_px
    if stereo == "ur":                 # the engine's right-camera mode carries no inverse-depth prior
        pb.rho_meas[:] = 0.0
        pb.rho_isig[:] = 0.0
    eng, lid_of = _engine_from_window(pb, sig_px=1.0)
    out, meta = pack.pack_problem(eng)
    _same_rows(out, pb)
    assert np.array_equal(meta["alive"], lid_of)
    # landmark order follows ascending engine id: map back through lid_of
    assert np.array_equal(out.lmp, eng.lm_p[meta["alive"]])
    # every row's origin points at the keyframe observation it came from
    for k in range(out.M):
        i, j = meta["src"][k]
        assert eng.slots[i]["lid"][j] == meta["alive"][meta["local"][k]]
        assert np.array_equal(eng.slots[i]["uv"][j], out.ob_z[k])
    if vi:
        assert out.dims is not None and len(out.imu) == len(pb.imu)
        assert np.array_equal(meta["off"], np.concatenate([[0], np.cumsum(out.dims)]))


def test_pack_landmark_order_is_ascending_engine_id():
    pb, _ = synth.make_window(P=5, L=120, views=(2, 5), seed=[6, 0])
    pb.ob_isig[:] = 1.0
    eng, lid_of = _engine_from_window(pb, seed=3, shuffle=True)
    out, meta = pack.pack_problem(eng)
    order = np.argsort(lid_of)                      # window landmark positions in engine-id order
    assert np.array_equal(meta["alive"], lid_of[order])
    assert np.allclose(out.lmp, pb.lmp[order])
    cnt = np.diff(pb.ob_start)[order]
    assert np.array_equal(np.diff(out.ob_start), cnt)


def _hand_engine():
    """Three keyframes (sids 10, 11, 12) and six landmarks with hand-picked observation sets."""
    def slot(sid, obs):
        lid = np.array([o[0] for o in obs], np.int64)
        return {"sid": sid, "R": np.eye(3), "p": np.zeros(3) + sid, "full": False, "lid": lid,
                "uv": np.array([[o[0] + 0.5, sid + 0.25] for o in obs]).reshape(-1, 2),
                "cam": np.array([o[1] for o in obs], np.int64), "valid": np.array([o[2] for o in obs], bool),
                "sig": np.array([o[3] for o in obs], np.float64)}
    slots = [
        slot(10, [(0, 0, True, 0.0), (1, 0, True, 0.0), (2, 1, True, 0.4), (3, 0, True, 0.0), (5, 0, True, 0.0)]),
        slot(11, [(1, 0, False, 0.0), (3, 0, True, 0.0), (4, 0, True, 0.0), (5, 1, True, 1e-12)]),
        slot(12, [(4, 0, False, 0.0)]),
    ]
    # hosts: lm 0,1,2,3 at 10; lm 4 at 11; lm 5 at 10
    eng = types.SimpleNamespace(
        slots=slots, lm_alive=np.array([1, 1, 1, 1, 1, 1, 0], bool), n_lm=6,
        lm_host=np.array([10, 10, 10, 10, 11, 10, 10]), lm_p=np.arange(21, dtype=float).reshape(7, 3),
        stereo_mode="prior", lm_rho_meas=np.array([0.5, 0.2, 0.3, -0.1, 0.4, 0.6, 0.0]),
        lm_sig_lk=np.array([0.2, 0.2, np.inf, 0.2, 0.1, 0.0, 0.0]), lm_span=np.array([0.0, 0.9, 0.0, 0.0, 0.0, 0.0, 0]),
        kappa_d=2.0, fB=40.0, sig_px=0.5, vi=False, prior=None, imu_factors=[], Rbc=np.eye(3), pbc=np.zeros(3),
        fx=400.0, fy=400.0, cx=320.0, cy=240.0, B=0.1, g_w=np.array([0, 0, -9.81]))
    return eng


def test_pack_whitening_and_inverse_depth_prior_rules():
    eng = _hand_engine()
    pb, meta = pack.pack_problem(eng)
    assert pb.L == 6 and pb.M == 10
    # rho sigma = max(kappa_d sig_lk, span) / fB; zero information for a non-finite sigma, a zero sigma or rho <= 0
    exp = np.array([40.0 / 0.4, 40.0 / 0.9, 0.0, 0.0, 40.0 / 0.2, 0.0])
    assert np.allclose(pb.rho_isig, exp, rtol=1e-15)
    # left rows 1 / sigma_px; right rows 1 / max(kappa_d sig, 1e-9)
    right = pb.ob_cam == 1
    assert np.allclose(pb.ob_isig[~right], 2.0)
    sig_r = np.array([0.4, 1e-12])
    assert np.allclose(np.sort(pb.ob_isig[right]), np.sort(1.0 / np.maximum(2.0 * sig_r, 1e-9)))
    # landmark 1: observations at keyframe 10 then 11 (keyframe order inside a landmark)
    r = slice(pb.ob_start[1], pb.ob_start[2])
    assert pb.ob_pose[r].tolist() == [0, 1] and pb.ob_valid[r].tolist() == [True, False]


def test_pack_posefree_selection():
    eng = _hand_engine()
    pb, meta = pack.pack_problem(eng, posefree=True)
    # pose-free: lm 0 (host-left only), lm 1 (host-left valid + an invalid row elsewhere);
    # not: lm 2 (host-right only), lm 3 (a second valid keyframe), lm 4 (host-left valid, invalid row elsewhere) is,
    # lm 5 (host-left + a valid right row at another keyframe) is not
    assert sorted(meta["pf_lids"].tolist()) == [0, 1, 4]
    assert meta["alive"].tolist() == [2, 3, 5]
    assert pb.pf.host.tolist() == [0, 0, 1]
    assert np.allclose(pb.pf.z, [[0.5, 10.25], [1.5, 10.25], [4.5, 11.25]])
    assert pb.pf.isg == 2.0
    assert np.allclose(pb.pf.lmp, eng.lm_p[[0, 1, 4]])
    # the kept landmarks keep their order and their rows only
    assert np.array_equal(np.diff(pb.ob_start), [1, 2, 2])
    assert pb.M == 5
    # "ur" mode never moves landmarks out
    eng.stereo_mode = "ur"
    pb2, meta2 = pack.pack_problem(eng, posefree=True)
    assert pb2.pf is None and len(meta2["pf_lids"]) == 0 and np.all(pb2.rho_isig == 0)


def test_pack_landmark_selection_and_prior_binding():
    eng = _hand_engine()
    pb, meta = pack.pack_problem(eng, lm_select=lambda a: a % 2 == 0)
    assert meta["alive"].tolist() == [0, 2, 4]
    assert pb.M == 4
    eng.prior = gauge_prior(11, np.eye(3), np.zeros(3))
    pb, _ = pack.pack_problem(eng)
    assert pb.prior.slots().tolist() == [1]
    pb, _ = pack.pack_problem(eng, with_prior=False)
    assert pb.prior is None
    eng.prior = gauge_prior(99, np.eye(3), np.zeros(3))
    with pytest.raises(RuntimeError):
        pack.pack_problem(eng)
    # a 15-dim prior block on a 6-dim keyframe
    from xclean_amd.problem import Prior
    eng.prior = Prior([10], [15], [(np.eye(3), np.zeros(3))], np.eye(15), np.zeros(15))
    with pytest.raises(RuntimeError):
        pack.pack_problem(eng)


def test_pack_imu_selection_and_writeback():
    pb, _ = synth.make_window(P=5, L=150, views=(2, 5), seed=[8, 1], vi=True)
    pb.ob_isig[:] = 1.0
    eng, lid_of = _engine_from_window(pb)
    out, meta = pack.pack_problem(eng, imu_sel=[0, 2])
    assert [(i, j) for i, j, _, _ in out.imu] == [(pb.imu[0][0], pb.imu[0][1]), (pb.imu[2][0], pb.imu[2][1])]
    # writeback of a moved state, then a second pack reads it
    out.pwb = out.pwb + 1.0
    out.lmp = out.lmp * 2.0
    out.vel = out.vel - 3.0
    pack.writeback(eng, out, meta)
    again, _ = pack.pack_problem(eng)
    assert np.allclose(again.pwb, pb.pwb + 1.0) and np.allclose(again.lmp, out.lmp)
    full = np.asarray(pb.dims) == 15
    assert np.allclose(again.vel[full], pb.vel[full] - 3.0)


@pytest.mark.gpu
@pytest.mark.parametrize("vi", [False, True])
def test_packed_window_solves_on_the_gpu_like_the_oracle(vi):
    """Engine state -> pack -> GPU LM (3 iterations, FP64) -> writeback, against the oracle on the same packed
    window; pose-free landmarks included."""
    from oracle import oracle as O
    from xclean_amd.window import BatchSolver
    pb, _ = synth.make_window(P=8, L=500, views=(1, 8), seed=[9, int(vi)], vi=vi)
    pb.ob_isig[pb.ob_cam == 0] = 1.0
    eng, lid_of = _engine_from_window(pb)
    eng.lm_span = np.zeros_like(eng.lm_span)
    out, meta = pack.pack_problem(eng, posefree=True)
    print("packed L", out.L, "pose-free", len(meta["pf_lids"]), "M", out.M, flush=True)
    assert len(meta["pf_lids"]) > 0
    ref = out.copy()
    O.window_solve(ref, iters=3)
    bs = BatchSolver([out], precision="fp64", use_graph=False)
    bs.upload(1e-4)
    bs.iterate(3)
    bs.download()
    for k in ("Rwb", "pwb", "lmp"):
        a, b = getattr(out, k), getattr(ref, k)
        assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b), k
    assert np.linalg.norm(out.pf.lmp - ref.pf.lmp) <= 1e-9 * np.linalg.norm(ref.pf.lmp)
    pack.writeback(eng, out, meta)
    assert np.array_equal(eng.lm_p[meta["alive"]], out.lmp)
    assert np.array_equal(eng.lm_p[meta["pf_lids"]], out.pf.lmp)
