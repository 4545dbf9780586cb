"""GPU parity: every test calls the HIP library through the C ABI and checks it against the
CPU oracle or the reference's golden vectors.

Tolerances: integer outputs (indices) bit-exact; FP64 stages (statistics, verified distance)
relative 1e-9; FP32 DTW cells relative 1e-4 of the cost (the FP32-vs-FP64 error is reported);
required bar for the final distance: relative 1e-5.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_DIR, golden_cases, golden_series

pytestmark = pytest.mark.gpu
CASES = golden_cases()


def rw(n, seed):
    return np.cumsum(np.random.default_rng(seed).standard_normal(n))


def test_sliding_stats_match_oracle(gpu, oracle):
    for n, m, off in ((100000, 256, 0.0), (20000, 1024, 5e4), (5000, 8, -1e3)):
        a = rw(n, 1) + off
        mu, sg = gpu.sliding_stats(a, m)
        W = np.lib.stride_tricks.sliding_window_view(a, m)
        mu2 = W.mean(1)
        sg2 = np.sqrt(((W - mu2[:, None]) ** 2).mean(1))     # two-pass, FP64
        assert np.max(np.abs(mu - mu2) / np.maximum(np.abs(mu2), 1)) < 1e-13
        assert np.max(np.abs(sg - sg2) / sg2) < 1e-10
        mu0, sg0 = oracle.mvmean(a, m)                        # the reference's running sums
        assert np.max(np.abs(sg - sg0) / sg0) < 1e-6


def test_envelope_is_exact(gpu, oracle):
    a = rw(50000, 2)
    for w in (0, 1, 2, 3, 13, 102, 511, 1024, 1536):
        lo, up = gpu.envelope(a, w)
        lo0, up0 = oracle.lemire(a, w)
        assert np.array_equal(lo, lo0) and np.array_equal(up, up0), w


@pytest.mark.parametrize("n", [2, 7, 255, 256, 257, 1023, 1024, 1025, 3000])
def test_envelope_tile_edges(gpu, oracle, n):
    """Series shorter than, equal to and just past one tile (256 outputs for w <= 16, 1024 above,
    so both tile sizes meet each of these lengths); radii up to and past n - 1
    (clipped windows at both ends, and the whole-series case); repeated values (ties)."""
    a = np.round(rw(n, n), 1)
    for w in sorted({0, 1, 5, 16, 17, 100, n // 2, n - 1, n + 7, 4000}):
        if min(w, n - 1) > 1536:
            with pytest.raises(gpu.DTWJError):
                gpu.envelope(a, w)
            continue
        lo, up = gpu.envelope(a, w)
        lo0, up0 = oracle.lemire(a, min(w, n - 1))
        assert np.array_equal(lo, lo0) and np.array_equal(up, up0), (n, w)


@pytest.mark.parametrize("m,w", [(64, 0), (64, 3), (128, 6), (256, 13), (512, 51), (1024, 102), (64, 63), (96, 200)])
def test_dtw_pairs_match_oracle(gpu, oracle, m, w):
    a = rw(20000, 3)
    rng = np.random.default_rng(m + w)
    pi = rng.integers(0, a.size - m, size=24)
    pj = rng.integers(0, a.size - m, size=24)
    c64 = gpu.dtw_pairs(a, m, w, pi, pj, 64)
    c32 = gpu.dtw_pairs(a, m, w, pi, pj, 32)
    ref = np.array([oracle.dtw_sq(oracle.znorm(a, int(i), m), oracle.znorm(a, int(j), m), min(w, m - 1))
                    for i, j in zip(pi, pj)])
    assert np.max(np.abs(c64 - ref) / np.maximum(ref, 1e-12)) < 1e-9
    assert np.max(np.abs(c32 - ref) / np.maximum(ref, 1e-3)) < 1e-4


@pytest.mark.parametrize("name", sorted(CASES))
def test_top1_matches_golden_and_oracle(gpu, oracle, name):
    c = CASES[name]
    a = golden_series(name)
    r = gpu.top1(a, c["m"], c["w"])
    o = oracle.motif_discovery(a, c["m"], c["w"])
    # identical pair (0-based; the oracle's pair is the reference's answer, with the seed pair
    # when the reference reports its seed fallback) and distance
    assert (r.i, r.j) == (o.i, o.j), (r.i, r.j, o.i, o.j)
    assert abs(r.dist - o.bsf) <= 1e-9 * o.bsf
    assert abs(r.dist - c["ref"]["final_distance"]) <= 5e-7 + 1e-5 * r.dist
    if o.found_in_sweep:
        assert (r.i + 1, r.j + 1) == (int(c["ref"]["first_min"]), int(c["ref"]["sec_min"]))
    # FP32-vs-FP64 error of the winner (reported; bounded by the join's error model)
    assert abs(r.dist_fp32 - r.dist) <= 2e-3 * r.dist + 1e-4
    assert r.pairs_swept == r.pairs_total


@pytest.mark.parametrize("n,m,w,seed", [(900, 32, 1, 11), (1000, 32, 4, 12), (800, 48, 2, 13), (700, 64, 9, 14)])
def test_top1_matches_brute_force(gpu, oracle, n, m, w, seed):
    a = rw(n, seed)
    bi, bj, bd = oracle.bruteforce_top1(a, m, w)
    r = gpu.top1(a, m, w)
    assert (r.i, r.j) == (bi, bj)
    assert abs(r.dist - bd) <= 1e-9 * bd


@pytest.mark.parametrize("n,m,w,seed", [(3000, 300, 150, 11), (2500, 200, 180, 11), (2000, 300, 260, 12),
                                        (1600, 300, 280, 13)])
def test_top1_wide_bands_match_oracle(gpu, oracle, n, m, w, seed):
    """Bands of 2w + 1 > 256 offsets: the join's DTW runs with 4 (w = 150, 180) or 8 (w = 260,
    280) offset pairs per lane, in every sweep mode and with small tiles (several items and
    drains per wave)."""
    a = rw(n, seed)
    o = oracle.motif_discovery(a, m, w)
    for opts in ({}, {"sweep": 1, "rows_per_item": 512, "waves_per_cu": 1}, {"sweep": 2}):
        r = gpu.top1(a, m, w, **opts)
        assert (r.i, r.j) == (o.i, o.j), (opts, r.i, r.j, o.i, o.j)
        assert abs(r.dist - o.bsf) <= 1e-9 * o.bsf
        assert r.pairs_swept == r.pairs_total


@pytest.mark.parametrize("n,m,w,seed", [(3000, 256, 70, 14), (4000, 800, 80, 15), (3500, 1024, 120, 16)])
def test_top1_two_offset_pair_bands_match_oracle(gpu, oracle, n, m, w, seed):
    """Bands of 129..255 offsets run two offset pairs per lane.  Where that fits more waves per
    CU (m = 800, 1024 here) the kernel reads the column windows' mean and 1/sigma from global
    memory and runs shorter cascade queues; at m = 256 it stages them in LDS.  Both forms, in
    every sweep mode and with small tiles (several items and drains per wave)."""
    a = rw(n, seed)
    o = oracle.motif_discovery(a, m, w)
    for opts in ({}, {"sweep": 1, "rows_per_item": 512, "waves_per_cu": 1}, {"sweep": 2}):
        r = gpu.top1(a, m, w, **opts)
        assert (r.i, r.j) == (o.i, o.j), (opts, r.i, r.j, o.i, o.j)
        assert abs(r.dist - o.bsf) <= 1e-9 * o.bsf
        assert r.pairs_swept == r.pairs_total


def test_single_pair_and_extreme_bands(gpu, oracle):
    a = rw(256, 21)
    r = gpu.top1(a, 128, 5)                     # n = 2m: exactly one admissible pair
    assert (r.i, r.j) == (0, 128)
    d = np.sqrt(oracle.dtw_sq(oracle.znorm(a, 0, 128), oracle.znorm(a, 128, 128), 5))
    assert abs(r.dist - d) <= 1e-9 * d
    a = rw(1500, 22)
    for w in (0, 31, 200):                      # w = 0 is Euclidean; w >= m-1 is unconstrained
        bi, bj, bd = oracle.bruteforce_top1(a, 32, min(w, 31))
        r = gpu.top1(a, 32, w)
        assert (r.i, r.j) == (bi, bj) and abs(r.dist - bd) <= 1e-9 * bd


def test_errors_are_reported(gpu):
    a = rw(3000, 23)
    a[1000:1300] = 4.0                          # constant window of length >= m
    with pytest.raises(gpu.DTWJError) as e:
        gpu.top1(a, 128, 6)
    assert e.value.code == gpu.DTWJ_EINVAL and "constant" in e.value.message
    with pytest.raises(gpu.DTWJError):
        gpu.top1(rw(3000, 24), 600, 513)        # 2*min(w,m-1)+1 > 1024


def test_affine_invariance_and_planted_recovery(gpu):
    a, src, dst = gpu.generate(200000, 31, n_plants=2, plant_len=256)
    r1 = gpu.top1(a, 256, 13)
    r2 = gpu.top1(3.5 * a - 1234.0, 256, 13)
    assert (r1.i, r1.j) == (r2.i, r2.j) and abs(r1.dist - r2.dist) <= 1e-7 * r1.dist
    planted = gpu.dtw_pairs(a, 256, 13, src, dst, 64)
    assert r1.dist <= np.sqrt(planted.min()) * (1 + 1e-12)
    assert r1.j - r1.i >= 256


def test_seeded_entry_point_matches(gpu, oracle):
    name = "cfg1_n16384_m128_w6"
    c = CASES[name]
    a = golden_series(name)
    mp, mpi = oracle.mpx(a, c["m"])
    r = gpu.top1_seeded(a, c["m"], c["w"], mp, mpi)
    o = oracle.join(a, c["m"], c["w"], mp, mpi)
    assert (r.i, r.j) == (o.i, o.j) and abs(r.dist - o.bsf) <= 1e-9 * o.bsf
    assert abs(r.seed_dist - o.seed_bsf) <= 1e-4 * o.seed_bsf


def test_context_reuse_is_stable(gpu):
    a, _, _ = gpu.generate(100000, 41, n_plants=1, plant_len=128)
    ctx = gpu.Context()
    ctx.upload(a)
    r1 = ctx.top1(128, 6)
    r2 = ctx.top1(128, 6)
    r3 = ctx.top1(128, 6, rows_per_item=512, waves_per_cu=4)
    ctx.close()
    assert (r1.i, r1.j, r1.dist) == (r2.i, r2.j, r2.dist) == (r3.i, r3.j, r3.dist)


def test_cli_drop_in(gpu, tmp_path):
    name = "cfg1_n16384_m128_w6"
    a = golden_series(name)
    data = tmp_path / "ts.txt"
    np.savetxt(data, a, fmt="%.17g")
    tfile = tmp_path / "time.txt"
    for _ in range(2):
        out = subprocess.run([gpu.CLI_PATH, str(data), str(tfile), "128", "0.05"], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stderr
    ref = CASES[name]["ref"]
    assert "first min is %d sec min is %d" % (ref["first_min"], ref["sec_min"]) in out.stdout
    assert "final distance :%f" % ref["final_distance"] in out.stdout
    lines = tfile.read_text().split()
    assert len(lines) == 2 and all(float(x) > 0 for x in lines)


def test_cli_grid_matches_reference_grid(gpu, oracle, tmp_path):
    """gpu_dtw --grid against the compiled reference's 24-configuration grid
    (CPU_version1/AAAA.cpp:30-38) on one series (tests/golden/grid.json).  Where the reference's
    sweep found the pair, the printed pair is its pair; where it kept its seed, the reference
    prints the last examined seed's windows (CPU_version1/new_dtw_motifGUI.cpp:75-76), and
    gpu_dtw prints the true pair, which the oracle's pair pins."""
    with open(os.path.join(GOLDEN_DIR, "grid.json")) as f:
        g = json.load(f)
    a = golden_series(g["name"])
    data = tmp_path / "ts.txt"
    np.savetxt(data, a, fmt="%.17g")
    tfile = tmp_path / "time.txt"
    out = subprocess.run([gpu.CLI_PATH, str(data), str(tfile), "--grid"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    blocks = re.split(r"subsequence length m=", out.stdout)[1:]
    assert len(blocks) == len(g["configs"]) == 24
    for blk, c in zip(blocks, g["configs"]):
        ref = c["ref"]
        assert blk.startswith("%d, band radius w=%d, n=%d" % (c["m"], c["w"], g["n"])), blk[:60]
        o = oracle.motif_discovery(a, c["m"], c["w"])
        assert (o.first_min, o.sec_min) == (int(ref["first_min"]), int(ref["sec_min"])), c
        assert "first min is %d sec min is %d" % (o.i + 1, o.j + 1) in blk, (c, blk)
        if o.found_in_sweep:
            assert (o.i + 1, o.j + 1) == (int(ref["first_min"]), int(ref["sec_min"])), c
        assert "final distance :%f" % ref["final_distance"] in blk, (c, blk)
        # the seed from the ED matrix profile, as the reference prints it
        assert "input of bsf:%5.3f" % ref["input_bsf"] in blk, (c, blk)
    lines = tfile.read_text().split()
    assert len(lines) == 24 and all(float(x) > 0 for x in lines)


def test_python_mirror_prints_reference_lines(gpu, capfd, tmp_path):
    name = "n8192_m64_w3"
    a = golden_series(name)
    r = gpu.DTWMotifDiscoveryGUI(a, 64, 3, str(tmp_path / "t.txt"))
    out = capfd.readouterr().out
    ref = CASES[name]["ref"]
    assert "first min is %d sec min is %d" % (ref["first_min"], ref["sec_min"]) in out
    assert (r.first_min, r.sec_min) == (int(ref["first_min"]), int(ref["sec_min"]))
