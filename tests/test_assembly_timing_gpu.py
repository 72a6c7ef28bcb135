"""The assembly-timing record of the window solver (xba_set_assembly_timing / xba_assembly_times): splitting the
captured LM iteration into the assembly graph and the rest leaves the solve unchanged (same state, same costs, same
accept decisions as one graph), and every replayed iteration records one positive assembly duration."""
import numpy as np
import pytest

from xclean_amd import synth

pytestmark = pytest.mark.gpu


def test_split_graph_solves_like_one_graph_and_records_every_iteration():
    from xclean_amd.window import BatchSolver
    pbs = [synth.make_window(P=p, L=l, views=(2, min(6, p)), seed=[31, i])[0]
           for i, (p, l) in enumerate([(6, 300), (9, 800), (12, 1500), (20, 400)])]
    ref = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=True)
    ref.upload(1e-4)
    ref.iterate(4)
    st_ref = ref.stats()
    dx_ref = ref.last_dx()
    bs = BatchSolver([pb.copy() for pb in pbs], precision="mixed", use_graph=True)
    bs.set_assembly_timing(True)
    bs.upload(1e-4)
    bs.iterate(4)
    ms, n = bs.assembly_times()
    assert n == 4 and ms > 0.0
    st = bs.stats()
    for a, b in zip(st, st_ref):
        assert a["accepted"] == b["accepted"]
        assert abs(a["cost"] - b["cost"]) <= 1e-9 * abs(b["cost"])
    for a, b in zip(bs.last_dx(), dx_ref):
        assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)
    # a second record starts empty; switching off captures one graph again and records nothing
    assert bs.assembly_times()[1] == 0
    bs.set_assembly_timing(False)
    bs.upload(1e-4)
    bs.iterate(2)
    assert bs.assembly_times()[1] == 0
