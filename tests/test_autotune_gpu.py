"""The autotuner on the GPU: it times real candidate plans of a small uneven batch, leaves the solver planned with
the fastest, the tuned plan solves to the same result as the default plan (FP64 within 1e-9 relative), and a second
call with the same workload comes from the cache without timing."""
import numpy as np
import pytest

from xclean_amd import synth

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_plan_that_solves_identically_and_caches_it(tmp_path):
    from xclean_amd import autotune
    from xclean_amd.window import BatchSolver
    pbs = [synth.make_window(P=P, L=L, views=(2, min(10, P)), seed=[11, i])[0]
           for i, (P, L) in enumerate([(5, 200), (12, 900), (30, 1500), (20, 400)])]
    pbs.append(synth.make_window(P=40, L=1200, views=(2, 10), seed=[11, 9])[0])   # n = 240: tiled by default
    ref = [pb.copy() for pb in pbs]
    BatchSolver(ref, precision="fp64").solve(iters=2)
    bs = BatchSolver([pb.copy() for pb in pbs], precision="fp64")
    cache = str(tmp_path / "tune.json")
    r = autotune.tune(bs, repeats=2, cache_path=cache)
    print("autotune", r["options"], "%.3f ms" % r["ms"], len(r["trials"]), "trials", flush=True)
    assert not r["cached"] and len(r["trials"]) >= 5
    assert bs.plan_options == r["options"]
    bs.solve(iters=2)
    for a, b in zip(bs.problems, ref):
        assert np.linalg.norm(a.pwb - b.pwb) <= 1e-9 * np.linalg.norm(b.pwb)
        assert np.linalg.norm(a.lmp - b.lmp) <= 1e-9 * np.linalg.norm(b.lmp)
    bs2 = BatchSolver([pb.copy() for pb in pbs], precision="fp64")
    r2 = autotune.tune(bs2, cache_path=cache)
    assert r2["cached"] and r2["options"] == r["options"] and bs2.plan_options == r["options"]
