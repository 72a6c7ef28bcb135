"""The reference engine's window dump format (vio/engine_live.py:804-821, win_%05d.npz): field layout, round trip,
refusals, and that a reloaded window solves exactly like the original (through the CPU oracle here; through the
GPU solver in the gpu-marked test)."""
import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import formats, synth


def _win(seed=0):
    return synth.make_window(P=8, L=400, views=(2, 6), seed=[7, seed])[0]


def _same(a, b):
    for k in ("Rwb", "pwb", "Rbc", "pbc", "lmp", "lm_host", "lm_mask", "rho_meas", "rho_isig", "ob_start", "ob_pose",
              "ob_cam", "ob_z", "ob_isig", "ob_valid"):
        np.testing.assert_array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k)), err_msg=k)
    for k in ("fx", "fy", "cx", "cy", "B", "huber_k"):
        assert getattr(a, k) == getattr(b, k), k
    if a.prior is None:
        assert b.prior is None
        return
    np.testing.assert_array_equal(a.prior.S, b.prior.S)
    np.testing.assert_array_equal(a.prior.r, b.prior.r)
    assert list(a.prior._idx) == list(b.prior._idx) and list(a.prior.dims) == list(b.prior.dims)
    for x, y in zip(a.prior.x0, b.prior.x0):
        np.testing.assert_array_equal(x[0], y[0])
        np.testing.assert_array_equal(x[1], y[1])


def test_field_layout_and_round_trip(tmp_path):
    pb = _win()
    assert pb.prior is not None
    p = tmp_path / "win_00000.npz"
    formats.save_window_npz(pb, p)
    with np.load(p, allow_pickle=False) as z:
        assert set(z.files) == set(formats.FIELDS)
        np.testing.assert_array_equal(z["cam"], [pb.fx, pb.fy, pb.cx, pb.cy, pb.B])
        assert z["prior_R0"].shape == (len(pb.prior.dims), 3, 3)
    _same(pb, formats.load_window_npz(p))
    pb.prior = None
    formats.save_window_npz(pb, tmp_path / "win_00001.npz")
    got = formats.load_window_dir(tmp_path)
    assert len(got) == 2 and got[1].prior is None


def test_refusals(tmp_path):
    pb = _win()
    p = tmp_path / "bad.npz"
    d = {k: np.zeros(1) for k in formats.FIELDS[:-1]}
    np.savez(p, **d)
    with pytest.raises(ValueError, match="not a window dump"):
        formats.load_window_npz(p)
    formats.save_window_npz(pb, p)
    with np.load(p) as z:
        d = {k: z[k] for k in z.files}
    d["prior_dims"] = np.array([15])
    np.savez(p, **d)
    with pytest.raises(ValueError, match="15-dim"):
        formats.load_window_npz(p)
    d["prior_dims"] = np.array([6])
    d["ob_start"] = d["ob_start"][:-1]
    np.savez(p, **d)
    with pytest.raises(ValueError, match="CSR"):
        formats.load_window_npz(p)


def test_reloaded_window_solves_like_the_original(tmp_path):
    pb = _win(1)
    formats.save_window_npz(pb, tmp_path / "win_00000.npz")
    back = formats.load_window_npz(tmp_path / "win_00000.npz")
    a, b = pb.copy(), back.copy()
    sa, dxa, _ = O.window_solve(a, iters=3, history=True)
    sb, dxb, _ = O.window_solve(b, iters=3, history=True)
    assert sa["accepted"] == sb["accepted"] and sa["cost"] == sb["cost"]
    np.testing.assert_array_equal(a.pwb, b.pwb)
    np.testing.assert_array_equal(a.lmp, b.lmp)


@pytest.mark.gpu
def test_reloaded_windows_replay_on_the_gpu(tmp_path):
    from xclean_amd.window import BatchSolver
    pbs = [_win(s) for s in range(4)]
    for i, pb in enumerate(pbs):
        formats.save_window_npz(pb, tmp_path / ("win_%05d.npz" % i))
    back = formats.load_window_dir(tmp_path)
    refs = [pb.copy() for pb in back]
    outs = [O.window_solve(r, iters=3, history=True)[0] for r in refs]
    st = BatchSolver(back, precision="fp64").solve(iters=3)
    for s, o, pb, ref in zip(st, outs, back, refs):
        assert s["accepted"] == o["accepted"]
        assert abs(s["cost"] - o["cost"]) <= 1e-9 * abs(o["cost"])
        assert np.linalg.norm(pb.pwb - ref.pwb) <= 1e-9 * np.linalg.norm(ref.pwb)
