"""Parity of the GPU inverse-depth local-BA drop-in (xclean_amd.ba._solve_np_id -> libxba xba_np_id_solve) with the
reference's backend/ba.py:1212 `_solve_np_id`.

Bars (FP64 throughout): poses, bearings and inverse depths within 1e-9 relative of the golden outputs that the
reference's own `_solve_np_id` wrote (tests/golden/solve_np_id_*.npz: out-of-window hosts; a held-keyframe mask
with the bearing / inverse-range prior, rotation factors and an active inverse-depth floor; a hard problem with
rejected steps) and of the CPU oracle on larger seeded problems; the first pose update within 1e-9 relative of the
oracle's; chi2/rms statistics within 1e-9 relative.  Plus the interface's edge cases.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xclean_amd import synth

from test_oracle_id import load_id_case  # noqa: E402  (tests/ is on sys.path under pytest)

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", ["default", "mask_prior_rotf", "rejects"])
def test_solve_np_id_matches_reference_golden(case):
    from xclean_amd.ba import _solve_np_id
    d, rig, free, prior, rotf = load_id_case(case)
    print("solve_np_id golden", case, flush=True)
    p, rho, mb, st = _solve_np_id(d["poses0"], d["rho0"], d["mbear0"], d["hidx"], d["hpose_ext"], d["okf"], d["olm"],
                                  d["ouv"], rig, free, int(d["iters"]), float(d["lam0"]), prior=prior, rotf=rotf,
                                  rho_floor=float(d["rho_floor"]))
    print("  poses %.2e rho %.2e mbear %.2e" % (_rel(p, d["out_poses"]), _rel(rho, d["out_rho"]),
                                                 _rel(mb, d["out_mbear"])), flush=True)
    assert _rel(p, d["out_poses"]) < 1e-9
    assert _rel(rho, d["out_rho"]) < 1e-9
    assert _rel(mb, d["out_mbear"]) < 1e-9
    np.testing.assert_allclose([st["chi2_0"], st["chi2_n"], st["rms_0"], st["rms_n"]], d["stats"], rtol=1e-9)


@pytest.mark.parametrize("seed,nk,nl,ext", [(21, 10, 300, 6), (22, 14, 500, 0), (23, 6, 60, 2)])
def test_solve_np_id_matches_oracle_iteration_by_iteration(seed, nk, nl, ext):
    from xclean_amd.ba import _solve_np_id
    q = synth.to_inverse_depth(synth.make_problem(nk=nk, nl=nl, obs_per=4, seed=seed), ext_every=ext,
                               ext_shift=[0.02, 0.0, 0.01, 0.0, 0.003, 0.0])
    args = (q["poses0"], q["rho0"], q["mbear0"], q["hidx"], q["hpose_ext"], q["okf"], q["olm"], q["ouv"], q["rig"],
            0, 5)
    po, ro, mo, so, dxo, acco = O.solve_np_id(*args, history=True)
    p, rho, mb, st, dxh = _solve_np_id(*args, history=True)
    print("seed", seed, "first update rel err %.2e" % _rel(dxh[0], dxo[0]), "accepts", acco, flush=True)
    assert _rel(dxh[0], dxo[0]) < 1e-9
    for i in range(1, len(dxo)):
        assert np.linalg.norm(dxh[i] - dxo[i]) <= 1e-9 * np.linalg.norm(dxo[0])
    assert _rel(p, po) < 1e-9 and _rel(rho, ro) < 1e-9 and _rel(mb, mo) < 1e-9
    np.testing.assert_allclose([st[k] for k in ("chi2_0", "chi2_n", "rms_0", "rms_n")],
                               [so[k] for k in ("chi2_0", "chi2_n", "rms_0", "rms_n")], rtol=1e-9)


def test_solve_np_id_edge_cases():
    from xclean_amd import _lib
    from xclean_amd.ba import _solve_np_id
    q = synth.to_inverse_depth(synth.make_problem(nk=5, nl=40, obs_per=3, seed=4), ext_every=4)
    base = (q["poses0"], q["rho0"], q["mbear0"], q["hidx"], q["hpose_ext"], q["okf"], q["olm"], q["ouv"], q["rig"])
    # no iterations: the state comes back unchanged, chi2_n == chi2_0
    p, rho, mb, st = _solve_np_id(*base, 0, 0)
    np.testing.assert_allclose(p, q["poses0"], atol=1e-15)
    np.testing.assert_array_equal(rho, q["rho0"])
    assert st["chi2_n"] == st["chi2_0"]
    # every keyframe held: poses unchanged, landmarks still move, as the oracle
    held = np.zeros(5, bool)
    p, rho, mb, st = _solve_np_id(*base, held, 3)
    po, ro, mo, so = O.solve_np_id(*base, held, 3)
    np.testing.assert_allclose(p, q["poses0"], atol=1e-15)
    assert _rel(rho, ro) < 1e-9 and _rel(mb, mo) < 1e-9
    # out-of-range host / observation indices are rejected before any compute
    bad = q["hidx"].copy()
    bad[0] = 7
    with pytest.raises(_lib.XbaError):
        _solve_np_id(q["poses0"], q["rho0"], q["mbear0"], bad, q["hpose_ext"], q["okf"], q["olm"], q["ouv"],
                     q["rig"], 0, 1)
    okf = q["okf"].copy()
    okf[3] = -1
    with pytest.raises(_lib.XbaError):
        _solve_np_id(q["poses0"], q["rho0"], q["mbear0"], q["hidx"], q["hpose_ext"], okf, q["olm"], q["ouv"],
                     q["rig"], 0, 1)
