"""GPU parity of the batched form (dtwj_top1_batch): many series of uneven length with mixed
(m, w) in one persistent join launch.  Every series' answer must equal the single-series
path's answer and, at sizes the oracle finishes quickly, the CPU oracle's brute-force answer.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (n, m, w, seed): uneven lengths, all four window lengths of the batched configuration,
# band ratios 0.01 / 0.05 / 0.1, one planted pair each
MIX = [(8192, 64, 1, 100), (12000, 128, 6, 101), (20000, 256, 26, 102), (9000, 512, 5, 103),
       (15000, 64, 6, 104), (30000, 128, 13, 105), (10000, 256, 3, 106), (18000, 512, 51, 107),
       (8192, 128, 1, 108), (25000, 64, 3, 109), (11000, 256, 13, 110), (16000, 512, 26, 111)]


def make_batch(gpu, mix):
    series = [gpu.generate(n, seed, n_plants=1, plant_len=m)[0] for n, m, w, seed in mix]
    return series, [m for _, m, _, _ in mix], [w for _, _, w, _ in mix]


def test_batch_matches_single_series_path(gpu):
    series, ms, ws = make_batch(gpu, MIX)
    res = gpu.top1_batch(series, ms, ws)
    for t, (a, m, w) in enumerate(zip(series, ms, ws)):
        r1 = gpu.top1(a, m, w)
        rb = res[t]
        assert not isinstance(rb, Exception), (t, rb)
        assert (rb.i, rb.j) == (r1.i, r1.j), (t, m, w)
        assert abs(rb.dist - r1.dist) <= 1e-9 * r1.dist
        assert rb.pairs_swept == r1.pairs_swept == r1.pairs_total


def test_batch_matches_oracle_brute_force(gpu, oracle):
    mix = [(1500, 64, 3, 120), (2200, 32, 1, 121), (1800, 96, 9, 122), (2600, 48, 4, 123)]
    series, ms, ws = make_batch(gpu, mix)
    res = gpu.top1_batch(series, ms, ws)
    for t, (a, m, w) in enumerate(zip(series, ms, ws)):
        bi, bj, bd = oracle.bruteforce_top1(a, m, w)
        assert (res[t].i, res[t].j) == (bi, bj)
        assert abs(res[t].dist - bd) <= 1e-9 * bd


def test_batch_with_several_band_classes_matches_oracle(gpu, oracle):
    """Series whose bands need 1, 2, 4 and 8 offset pairs per lane in one batch: the seed DTWs
    run as one launch per class, the join with the widest class's kernel."""
    mix = [(3000, 300, 150, 11), (1800, 128, 6, 130), (2500, 200, 80, 131), (1600, 300, 280, 13)]
    series = [np.cumsum(np.random.default_rng(seed).standard_normal(n)) for n, _, _, seed in mix]
    ms, ws = [m for _, m, _, _ in mix], [w for _, _, w, _ in mix]
    refs = [oracle.motif_discovery(a, m, w) for a, m, w in zip(series, ms, ws)]
    for opts in ({}, {"graph": -1}):
        res = gpu.top1_batch(series, ms, ws, **opts)
        for t, (o, m, w) in enumerate(zip(refs, ms, ws)):
            assert not isinstance(res[t], Exception), (t, res[t])
            assert (res[t].i, res[t].j) == (o.i, o.j), (opts, t, m, w)
            assert abs(res[t].dist - o.bsf) <= 1e-9 * o.bsf


def test_batch_of_two_offset_pair_bands_with_long_windows_matches_oracle(gpu, oracle):
    """Widest band 205 offsets (two offset pairs per lane) and longest window 1024: the batch
    runs the form that reads column-window statistics from global memory (each series' own) and
    sizes LDS by m = 1024, with shorter members and narrower bands beside it."""
    mix = [(3500, 1024, 102, 140), (3000, 256, 13, 141), (4000, 800, 80, 142), (2500, 512, 70, 143)]
    series = [np.cumsum(np.random.default_rng(seed).standard_normal(n)) for n, _, _, seed in mix]
    ms, ws = [m for _, m, _, _ in mix], [w for _, _, w, _ in mix]
    refs = [oracle.motif_discovery(a, m, w) for a, m, w in zip(series, ms, ws)]
    for opts in ({}, {"graph": -1}):
        res = gpu.top1_batch(series, ms, ws, **opts)
        for t, (o, m, w) in enumerate(zip(refs, ms, ws)):
            assert not isinstance(res[t], Exception), (t, res[t])
            assert (res[t].i, res[t].j) == (o.i, o.j), (opts, t, m, w)
            assert abs(res[t].dist - o.bsf) <= 1e-9 * o.bsf


def test_graph_and_direct_launch_agree_and_repeat(gpu):
    series, ms, ws = make_batch(gpu, MIX[:6])
    ctx = gpu.Context()
    runs = [ctx.top1_batch(series, ms, ws), ctx.top1_batch(series, ms, ws, graph=-1), ctx.top1_batch(series, ms, ws)]
    ctx.close()
    for t in range(len(series)):
        assert len({(r[t].i, r[t].j, r[t].dist) for r in runs}) == 1


def test_autotuned_batch_is_exact_and_reports_its_choice(gpu):
    series, ms, ws = make_batch(gpu, MIX)
    ctx = gpu.Context()
    tuned = ctx.top1_batch(series, ms, ws, autotune=1)
    again = ctx.top1_batch(series, ms, ws, autotune=1)  # cached choice
    ctx.close()
    plain = gpu.top1_batch(series, ms, ws)
    for t in range(len(series)):
        assert (tuned[t].i, tuned[t].j, tuned[t].dist) == (plain[t].i, plain[t].j, plain[t].dist)
        assert tuned[t].rows_per_item in (1024, 2048, 4096)
        assert tuned[t].waves_per_cu in (6, 8)
        assert (again[t].rows_per_item, again[t].waves_per_cu) == (tuned[t].rows_per_item, tuned[t].waves_per_cu)


def test_bad_series_is_rejected_alone(gpu):
    series, ms, ws = make_batch(gpu, MIX[:4])
    const = series[1].copy()
    const[3000:3000 + 200] = 7.0  # a constant window of length 128
    bad_short = series[2][:100]
    res = gpu.top1_batch([series[0], const, bad_short, series[3]], ms[:4], ws[:4])
    assert isinstance(res[1], gpu.DTWJError) and res[1].code == gpu.DTWJ_EINVAL
    assert isinstance(res[2], gpu.DTWJError) and res[2].code == gpu.DTWJ_EINVAL
    for t in (0, 3):
        r1 = gpu.top1(series[t], ms[t], ws[t])
        assert (res[t].i, res[t].j) == (r1.i, r1.j)


def test_subset_of_the_512_series_workload_matches_single_series(gpu):
    """64 series of the 512-series batched workload (seeds 100..155 and the first eight more of
    its m = 512, band ratio 0.1 members, w = 51, the batch's heaviest class), in one batch:
    every series equal to its own single-series join (pair, FP64 distance, pairs swept)."""
    heavy = [s for s in range(156, 612) if gpu.batch_spec(s)[1:] == (512, 0.1)][:8]
    seeds = list(range(100, 156)) + heavy
    series, ms, ws = gpu.batch_series(seeds)
    assert sum(1 for m, w in zip(ms, ws) if (m, w) == (512, 51)) >= 10
    res = gpu.top1_batch(series, ms, ws)
    for t, (a, m, w) in enumerate(zip(series, ms, ws)):
        rb = res[t]
        assert not isinstance(rb, Exception), (seeds[t], rb)
        r1 = gpu.top1(a, m, w)
        assert (rb.i, rb.j) == (r1.i, r1.j), (seeds[t], m, w)
        assert abs(rb.dist - r1.dist) <= 1e-9 * r1.dist
        assert rb.pairs_swept == r1.pairs_swept == r1.pairs_total
