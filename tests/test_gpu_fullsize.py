"""GPU: the full-size workloads (n = 1,000,000) against the answers the compiled reference CPU
path and the oracle gave for them (tests/golden/fullsize.json, written in the build container
by tests/golden/make_fullsize.py; the reference took about 2.2 hours per case on one core).

The series are rebuilt with the package's deterministic generator and must hash to the SHA-256
the fixture recorded.  Bar: the pair exactly (the reference prints it 1-based), the distance
within the reference's printed 6 decimals and within 1e-9 relative of the oracle's FP64 value,
and the seed distance the reference printed ("input of bsf", 3 decimals) from the matrix-profile
pipeline.  The oracle entries' dtw_cnt counts the pairs that passed both LB_KK directions: those
runs predate the oracle's LB_Petitjean, which the reference's DTW count includes.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def cases():
    path = os.path.join(GOLDEN_DIR, "fullsize.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


CASES = cases()


def rebuild(gpu, c):
    a, src, dst = gpu.generate(c["n"], c["seed"], **c["gen"])
    assert hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest() == c["sha256"]
    assert list(src) == c["plants_src"] and list(dst) == c["plants_dst"]
    return a


@pytest.mark.parametrize("name", sorted(CASES))
def test_full_size_answer_matches_the_reference(gpu, name):
    c = CASES[name]
    a = rebuild(gpu, c)
    r = gpu.top1(a, c["m"], c["w"])
    ref, orc = c["ref"], c["oracle"]
    assert (r.i + 1, r.j + 1) == (int(ref["first_min"]), int(ref["sec_min"]))
    assert (r.i, r.j) == (orc["i"], orc["j"])
    assert abs(r.dist - ref["final_distance"]) <= 5e-7
    assert abs(r.dist - orc["dist"]) <= 1e-9 * orc["dist"]


@pytest.mark.parametrize("name", sorted(CASES))
def test_full_size_reference_pipeline_seed(gpu, name, tmp_path):
    """dtwj_motif_discovery (GPU matrix profile, the reference's seed, seeded join) prints the
    reference's seed distance and answer for the full-size series."""
    c = CASES[name]
    a = rebuild(gpu, c)
    out = tmp_path / "time.txt"
    r = gpu.DTWMotifDiscoveryGUI(a, c["m"], c["w"], str(out))
    ref = c["ref"]
    assert abs(r.seed_dist - ref["input_bsf"]) <= 5e-4
    assert (r.i + 1, r.j + 1) == (int(ref["first_min"]), int(ref["sec_min"]))
    assert abs(r.dist - ref["final_distance"]) <= 5e-7
