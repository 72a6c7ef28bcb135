"""The autotuner's search and candidate space (host logic, no GPU): coordinate descent keeps the fastest value of
each option in turn (two rounds), skips plans the solver rejects, never re-times a plan, and the candidate space
always holds the solver's defaults."""
import numpy as np
import pytest

from xclean_amd import autotune as AT


def test_coordinate_descent_finds_the_separable_optimum_and_skips_rejected_plans():
    best = dict(slot_tile=12, lm_per_chunk=64, big_n=128, big_fp64=1, grid=0)
    space = dict(slot_tile=[0, 8, 12, 20], lm_per_chunk=[0, 64, 256], big_n=[0, 128, -1], big_fp64=[0, 1],
                 grid=[0, 512])
    seen = []

    def measure(o):
        seen.append(tuple(sorted(o.items())))
        if o["slot_tile"] == 20:
            raise RuntimeError("tile does not fit")
        return 1.0 + sum(0.1 * (o[k] != best[k]) for k in best)

    opts, t, trials = AT.search(measure, space)
    assert opts == best
    assert t == pytest.approx(1.0)
    assert len(seen) == len(set(seen))                       # no plan timed twice
    assert any(tt is None for _o, tt in trials)              # the rejected plan is recorded as such


def test_second_round_revisits_an_option_after_a_later_one_moved():
    # every window in FP64 pays only with the lower tiling threshold, which alone is slower: one round stops at
    # (0, 2), the second finds (64, 2)
    table = {(0, 0): 1.0, (64, 0): 1.05, (0, 2): 0.99, (64, 2): 0.9}
    space = dict(big_n=[0, 64], big_fp64=[0, 2])

    def measure(o):
        return table[(o["big_n"], o["big_fp64"])]

    one, _, _ = AT.search(measure, space, start=dict(big_n=0, big_fp64=0), order=("big_n", "big_fp64"), rounds=1)
    assert one == dict(big_n=0, big_fp64=2)
    two, t, _ = AT.search(measure, space, start=dict(big_n=0, big_fp64=0), order=("big_n", "big_fp64"))
    assert two == dict(big_n=64, big_fp64=2) and t == pytest.approx(0.9)


def test_search_raises_when_the_default_plan_is_rejected():
    with pytest.raises(RuntimeError):
        AT.search(lambda o: None, dict(slot_tile=[8]))


def test_candidate_space_holds_the_defaults_and_only_moving_thresholds():
    sp = AT.candidate_space(nlm=2_000_000, ncu=256, ns=np.array([30, 90, 180]))
    for k in AT.ORDER:
        assert sp[k][0] == 0
    assert sp["big_n"] == [0, 64, 96, 112, 128, 144, 160]    # 256 / 384 / 512 / never change nothing here
    assert sp["big_fp64"] == [0, 1, 2]
    sp2 = AT.candidate_space(nlm=200_000, ncu=256, ns=np.array([900]))
    assert sp2["big_n"] == [0, -1]                           # the tiled solve or one workgroup
    assert AT.candidate_space(nlm=200_000, ncu=256, ns=np.array([900]), mixed=False)["big_fp64"] == [0]
    assert AT.candidate_space(nlm=2_000, ncu=256, ns=np.array([30, 60]))["big_fp64"] == [0, 2]  # nothing tiled
    assert all(8 <= v <= 4096 for v in sp2["lm_per_chunk"][1:])
