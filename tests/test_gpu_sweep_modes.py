"""GPU parity of the sweep's two ball tests.  The sweep chooses per work item between the
two-sided ball bound (64-row chunks) and the one-sided one (256-row chunks, drained in 64-row
parts); both are exact.  With the choice forced (dtwj_options.sweep = 1 or 2) every golden case
must give the pair and distance of the adaptive default, which test_gpu_parity checks against
the oracle, and the golden distance.  Small tiles and few waves make each wave sweep several
items and chunks.  A deliberately poor seed (a distant pair) makes the first chunks overfill the
survivor queue, which exercises the drain in the middle of a chunk and the restaging after it.

Tolerances: pair indices bit-exact; distance equal to the default mode's (both are the FP64
re-verification of the same pair) and within 1e-5 relative of the golden distance.
"""
import numpy as np
import pytest

from conftest import golden_cases, golden_series

pytestmark = pytest.mark.gpu
CASES = golden_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_forced_sweep_modes_match_the_default(gpu, name):
    c = CASES[name]
    a = golden_series(name)
    m, w = c["m"], c["w"]
    r0 = gpu.top1(a, m, w)
    for sweep in (1, 2):
        for opts in ({}, {"rows_per_item": 512, "waves_per_cu": 1}):
            r = gpu.top1(a, m, w, sweep=sweep, **opts)
            assert (r.i, r.j, r.dist) == (r0.i, r0.j, r0.dist), (sweep, opts, (r.i, r.j, r.dist), (r0.i, r0.j, r0.dist))
            assert abs(r.dist - c["ref"]["final_distance"]) <= 5e-7 + 1e-5 * r.dist
            assert r.pairs_swept == r.pairs_total
            if sweep == 1:
                assert r.pairs_swept_one_sided == 0
            else:
                assert r.pairs_swept_one_sided == r.pairs_swept


@pytest.mark.parametrize("name", ["cfg1_n16384_m128_w6", "n16384_m256_w13", "n8192_m64_w3"])
def test_poor_seed_overfills_the_queue_and_stays_exact(gpu, name):
    c = CASES[name]
    a = golden_series(name)
    m, w = c["m"], c["w"]
    nw = a.size - m + 1
    r0 = gpu.top1(a, m, w)
    # one finite profile entry: the only seed pair besides (0, m) is (0, nw - 1)
    mp = np.full(nw, np.inf)
    mpi = np.ones(nw, dtype=np.int32)
    mp[0], mpi[0] = 0.0, nw
    for sweep in (0, 1, 2):
        r = gpu.top1_seeded(a, m, w, mp, mpi, sweep=sweep, rows_per_item=1024, waves_per_cu=1)
        assert r.seed_dist > 1.5 * r0.dist
        assert (r.i, r.j, r.dist) == (r0.i, r0.j, r0.dist), (sweep, (r.i, r.j, r.dist), (r0.i, r0.j, r0.dist))


def test_forced_sweep_in_a_batch(gpu):
    specs = [(9000, 64, 3, 501), (12000, 128, 6, 502), (15000, 256, 13, 503), (8192, 32, 1, 504)]
    series, ms, ws = [], [], []
    for n, m, w, seed in specs:
        s, _, _ = gpu.generate(n, seed, n_plants=1, plant_len=m)
        series.append(s)
        ms.append(m)
        ws.append(w)
    ref = gpu.top1_batch(series, ms, ws)
    for sweep in (1, 2):
        got = gpu.top1_batch(series, ms, ws, sweep=sweep, waves_per_cu=1)
        for k in range(len(specs)):
            assert (got[k].i, got[k].j, got[k].dist) == (ref[k].i, ref[k].j, ref[k].dist), (sweep, k)


def test_sweep_option_is_validated(gpu):
    a = np.cumsum(np.random.default_rng(5).standard_normal(2000))
    with pytest.raises(gpu.DTWJError) as e:
        gpu.top1(a, 64, 3, sweep=3)
    assert e.value.code == gpu.DTWJ_EINVAL
