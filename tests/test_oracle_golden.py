"""The CPU oracle against golden vectors written by the compiled reference CPU path.

tests/golden/golden.json holds, per case, what CPU_version1 printed for the stored series:
1-based first/second index (CPU_version1/new_dtw_motifGUI.cpp:575), final distance and seed
distance (:578, printed with %f / %5.3f), the prune fraction of every stage -- diag_fast, the
global mask stage, the alpha-tuned local masks, LB_KK, LB_Petitjean (:569-574, printed with 6
significant digits) --, the number of diagonals each stage ran on ("fast / global / local_up /
local_down times", :561-564) and the number of pairs that reached the DTW ("DTW times", :565).
"""
import math

import pytest

from conftest import golden_cases, golden_series

CASES = golden_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference(name, oracle):
    c = CASES[name]
    a = golden_series(name)
    r = oracle.motif_discovery(a, c["m"], c["w"])
    ref = c["ref"]
    # indices exactly as the reference reports them (1-based if found by the sweep,
    # 0-based seed fallback otherwise)
    assert r.first_min == int(ref["first_min"]) and r.sec_min == int(ref["sec_min"])
    assert abs(r.bsf - ref["final_distance"]) <= 5e-7
    assert abs(r.seed_bsf - ref["input_bsf"]) <= 5e-4
    check_stage_counts(r, ref, name)
    assert r.pe_cnt >= r.dtw_cnt


def check_stage_counts(r, ref, what):
    """Every stage's prune fraction (6 significant digits, as printed), the number of diagonals
    each stage ran on and the DTW count (exact)."""
    all_cnt = r.all_cnt
    for key, val in (("fast_prunes", r.fast_cnt), ("global_prunes", r.global_cnt),
                     ("local_prunes", r.local_cnt), ("kk_prunes", r.kk_cnt - r.pe_cnt),
                     ("pe_prunes", r.pe_cnt - r.dtw_cnt)):
        got = val / all_cnt
        want = ref[key]
        assert math.isclose(got, want, rel_tol=6e-6, abs_tol=1e-12), (what, key, got, want)
    for key, val in (("fast_times", r.fast_stage_rounds), ("global_times", r.global_rounds),
                     ("local_up_times", r.local_rounds), ("local_down_times", r.local_down_rounds),
                     ("dtw_times", r.dtw_cnt)):
        assert val == ref[key], (what, key, val, ref[key])
    assert r.bad_bin == 0


def test_local_masks_prune_on_the_fixtures():
    """The fixtures exercise the alpha-tuned local masks: some cases prune with them."""
    pruning = [n for n, c in CASES.items() if c["ref"]["local_prunes"] > 0]
    assert len(pruning) >= 3, pruning


def test_golden_pairs_are_admissible():
    for name, c in CASES.items():
        ref = c["ref"]
        i, j = int(ref["first_min"]), int(ref["sec_min"])
        assert j > i and j - i >= c["m"]


def grid_fixture():
    import json
    import os
    from conftest import GOLDEN_DIR
    with open(os.path.join(GOLDEN_DIR, "grid.json")) as f:
        return json.load(f)


def test_grid_fixture_follows_the_reference_grid():
    # CPU_version1/AAAA.cpp:30-38: ratios outer, lengths inner, w = MAX(round(m r), 1)
    g = grid_fixture()
    cfg = [(c["m"], c["r"], c["w"]) for c in g["configs"]]
    want = [(m, r, max(int(math.floor(m * r + 0.5)), 1)) for r in (0.01, 0.02, 0.05, 0.1)
            for m in (32, 64, 128, 256, 512, 1024)]
    assert cfg == want
    for c in g["configs"]:
        i, j = int(c["ref"]["first_min"]), int(c["ref"]["sec_min"])
        assert j > i and j - i >= c["m"]


def test_oracle_matches_reference_grid(oracle):
    """The reference CPU main's 24-configuration grid on one series (tests/golden/grid.json)."""
    g = grid_fixture()
    a = golden_series(g["name"])
    assert len(a) == g["n"]
    for c in g["configs"]:
        r = oracle.motif_discovery(a, c["m"], c["w"])
        ref = c["ref"]
        assert (r.first_min, r.sec_min) == (int(ref["first_min"]), int(ref["sec_min"])), c
        assert abs(r.bsf - ref["final_distance"]) <= 5e-7, c
        assert abs(r.seed_bsf - ref["input_bsf"]) <= 5e-4, c
        check_stage_counts(r, ref, (c["m"], c["w"]))
