"""Writes golden vectors for the top-1 join by running the reference CPU path.

The reference (CPU_version1, compiled from its own sources by oracle/build_ref.sh into
oracle/_ref/ref_dtw) is run once per configuration on a series produced by the package's
deterministic generator; its printed result lines are parsed and stored with the series:

    tests/golden/<name>.npy   the FP64 series (exactly what the reference read: the text file
                              is written with 17 significant digits)
    tests/golden/golden.json  per case: generator parameters, m, w and the reference outputs
                              (1-based first/second index, final distance, seed distance,
                              the prune fractions of every stage -- diag_fast, global, local
                              masks, LB_KK, LB_Petitjean -- the DTW count, the number of
                              diagonals each stage ran on, its own timer).
    tests/golden/grid.json    the reference CPU main's grid (CPU_version1/AAAA.cpp:30-38: band
                              ratios 0.01, 0.02, 0.05, 0.1 outer, m = 32 .. 1024 inner,
                              w = MAX(round(m r), 1)) on one series, one result per configuration.

Run in the build container only (needs /root/reference): python tests/golden/make_golden.py
"""
import importlib.util
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "scaling-subsequence-similarity-join-based-on-dynamic-time-warping-cpu-gpu_amd")
REF = os.path.join(ROOT, "oracle", "_ref", "ref_dtw")

CASES = [
    # name, n, seed, gen kwargs, m, w
    ("cfg1_n16384_m128_w6", 16384, 0, dict(n_plants=1, plant_len=128, plant_src=[2000], plant_dst=[11000]), 128, 6),
    ("n8192_m64_w1", 8192, 10, dict(n_plants=1, plant_len=64), 64, 1),
    ("n8192_m64_w3", 8192, 11, dict(n_plants=1, plant_len=64), 64, 3),
    ("n8192_m64_w6", 8192, 12, dict(n_plants=1, plant_len=64), 64, 6),
    ("n8192_m32_w1", 8192, 13, dict(n_plants=1, plant_len=32), 32, 1),
    ("n16384_m256_w3", 16384, 14, dict(n_plants=3, plant_len=256), 256, 3),
    ("n16384_m256_w13", 16384, 15, dict(n_plants=2, plant_len=256), 256, 13),
    ("n16384_m256_w26", 16384, 16, dict(n_plants=1, plant_len=256), 256, 26),
    ("n16384_m512_w51", 16384, 17, dict(n_plants=1, plant_len=512), 512, 51),
    ("n16384_m1024_w102", 16384, 18, dict(n_plants=1, plant_len=1024), 1024, 102),
    ("n16384_m1024_w10", 16384, 19, dict(n_plants=1, plant_len=1024), 1024, 10),
    ("n8192_regime_m128_w13", 8192, 20, dict(n_plants=1, plant_len=128, step_std=0.5, step_std_alt=2.0,
                                              regime_len=1000), 128, 13),
    ("n8192_noplant_m128_w6", 8192, 21, dict(), 128, 6),
    ("n4096_m64_w2", 4096, 22, dict(n_plants=1, plant_len=64), 64, 2),
    # the wide-band workload's regime shifts (step std 0.5 / 2.0) at its window and band
    ("n24576_regime_m1024_w102", 24576, 24, dict(n_plants=1, plant_len=1024, step_std=0.5, step_std_alt=2.0,
                                                 regime_len=3000), 1024, 102),
    # the medium workload's window and band as the configuration states it (w = 2)
    ("n16384_m256_w2", 16384, 25, dict(n_plants=3, plant_len=256), 256, 2),
]

# the reference CPU main's configuration grid (CPU_version1/AAAA.cpp:30-38)
GRID_M = [32, 64, 128, 256, 512, 1024]
GRID_R = [0.01, 0.02, 0.05, 0.1]
GRID_CASE = ("grid_n4096", 4096, 23, dict(n_plants=1, plant_len=128))


def grid_configs():
    """(m, r, w) in the reference's order, w = MAX(round(m r), 1) (round half away from zero)."""
    import math
    out = []
    for r in GRID_R:
        for m in GRID_M:
            x = m * r
            w = max(int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5)), 1)
            out.append((m, r, w))
    return out


def load_pkg():
    spec = importlib.util.spec_from_file_location("dtwj_pkg", os.path.join(PKG, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def parse(out):
    r = {}
    m = re.search(r"first min is (\S+) sec min is (\S+)", out)
    r["first_min"], r["sec_min"] = float(m.group(1)), float(m.group(2))
    m = re.search(r"total time is (\S+)s\s+input of bsf:\s*(\S+)\s*,final distance :(\S+)", out)
    r["ref_seconds"], r["input_bsf"], r["final_distance"] = float(m.group(1)), float(m.group(2)), float(m.group(3))
    for key, pat in (("fast_prunes", r"fast prunes (\S+)"), ("global_prunes", r"global prunes (\S+)"),
                     ("dtw_times", r"DTW times (\S+)"), ("local_prunes", r"local prunes (\S+)"),
                     ("kk_prunes", r"KK prunes (\S+)"), ("pe_prunes", r"Pe prunes (\S+)"),
                     ("fast_times", r"fast times :(\S+)"), ("global_times", r"global times :(\S+)"),
                     ("local_up_times", r"local_up times :(\S+)"), ("local_down_times", r"local_down times :(\S+)")):
        m = re.search(pat, out)
        r[key] = float(m.group(1)) if m else None
    return r


def main(names=None):
    if not os.path.exists(REF):
        sys.exit("reference binary missing: run bash oracle/build_ref.sh first")
    pkg = load_pkg()
    path = os.path.join(HERE, "golden.json")
    db = json.load(open(path)) if os.path.exists(path) else {}
    for name, n, seed, kw, m, w in CASES:
        if names and name not in names:
            continue
        series, src, dst = pkg.generate(n, seed, **kw)
        with tempfile.TemporaryDirectory() as td:
            txt = os.path.join(td, "ts.txt")
            np.savetxt(txt, series, fmt="%.17g")
            back = np.loadtxt(txt)
            assert np.array_equal(back, series)
            t0 = time.time()
            out = subprocess.run([REF, txt, os.path.join(td, "time.txt"), str(m), str(w)], capture_output=True,
                                 text=True, check=True).stdout
            wall = time.time() - t0
        res = parse(out)
        np.save(os.path.join(HERE, name + ".npy"), series)
        gen = {k: (list(map(int, v)) if isinstance(v, (list, tuple)) else v) for k, v in kw.items()}
        db[name] = dict(n=n, seed=seed, gen=gen, m=m, w=w, plants_src=list(map(int, src)),
                        plants_dst=list(map(int, dst)), ref=res, ref_wall_seconds=round(wall, 3))
        print(name, res, "%.1fs" % wall, flush=True)
        json.dump(db, open(path, "w"), indent=1, sort_keys=True)


def main_grid():
    """Runs the reference once per grid configuration on one series; writes grid.json."""
    if not os.path.exists(REF):
        sys.exit("reference binary missing: run bash oracle/build_ref.sh first")
    pkg = load_pkg()
    name, n, seed, kw = GRID_CASE
    series, src, dst = pkg.generate(n, seed, **kw)
    results = []
    with tempfile.TemporaryDirectory() as td:
        txt = os.path.join(td, "ts.txt")
        np.savetxt(txt, series, fmt="%.17g")
        assert np.array_equal(np.loadtxt(txt), series)
        for m, r, w in grid_configs():
            out = subprocess.run([REF, txt, os.path.join(td, "time.txt"), str(m), str(w)], capture_output=True,
                                 text=True, check=True).stdout
            res = parse(out)
            results.append(dict(m=m, r=r, w=w, ref=res))
            print(m, r, w, res, flush=True)
    np.save(os.path.join(HERE, name + ".npy"), series)
    gen = {k: (list(map(int, v)) if isinstance(v, (list, tuple)) else v) for k, v in kw.items()}
    json.dump(dict(name=name, n=n, seed=seed, gen=gen, configs=results), open(os.path.join(HERE, "grid.json"), "w"),
              indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["--grid"]:
        main_grid()
    else:
        main(sys.argv[1:] or None)
