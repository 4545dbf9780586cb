"""GPU: the segment-mean (PAA) bound the join applies between LB_Kim and LB_Keogh.

It only prunes, so it must (1) never exceed LB_Keogh or the banded DTW of a pair, (2) be the
segment-mean quantity it claims to be, and (3) leave every answer unchanged.  (1) is checked
against the oracle's LB_Keogh (per-window envelope, at least the raw-series one the bound uses)
and textbook DTW; (2) against an FP64 numpy evaluation of the definition, from which the FP32
kernel may only fall below by its rounding margin; (3) on the golden series with the stage
forced on, forced off and automatic.
"""
import numpy as np
import pytest

from conftest import golden_cases, golden_series

pytestmark = pytest.mark.gpu
CASES = golden_cases()


def rw(n, seed, regime=False):
    r = np.random.default_rng(seed)
    steps = r.standard_normal(n)
    if regime:
        steps *= np.where((np.arange(n) // 3000) % 2 == 0, 0.5, 2.0)
    return np.cumsum(steps)


def paa_fp64(a, m, w, pi, pj, lemire):
    """len * sum_s max(Z_j - U_i, L_i - Z_j, 0)^2 over S segments of len = m // S samples
    (S = 32 for m >= 512, else 16), the larger of the two directions, with the raw-series
    envelope of radius min(w, m - 1)."""
    lo, up = lemire(a, min(w, m - 1))
    S = 32 if m >= 512 else 16
    seg = m // S
    def rows(i):
        x = a[i:i + m]
        mu, sd = x.mean(), x.std()
        z = ((x - mu) / sd)[:S * seg].reshape(S, seg).mean(1)
        u = ((up[i:i + m] - mu) / sd)[:S * seg].reshape(S, seg).mean(1)
        l = ((lo[i:i + m] - mu) / sd)[:S * seg].reshape(S, seg).mean(1)
        return z, u, l
    out = []
    for i, j in zip(pi, pj):
        zi, ui, li = rows(int(i))
        zj, uj, lj = rows(int(j))
        ta = np.maximum(np.maximum(zj - ui, li - zj), 0.0)
        tb = np.maximum(np.maximum(zi - uj, lj - zi), 0.0)
        out.append(seg * max(float(ta @ ta), float(tb @ tb)))
    return np.array(out)


@pytest.mark.parametrize("m,w,regime", [(16, 2, False), (128, 6, False), (256, 13, True), (511, 40, False),
                                        (512, 51, False), (1000, 102, True), (1024, 300, False)])
def test_paa_bound_is_a_lower_bound(gpu, oracle, m, w, regime):
    a = rw(30000, m + w, regime) + 1e3  # an offset: the table is built against the window mean
    rng = np.random.default_rng(m)
    nw = a.size - m + 1
    pi = rng.integers(0, nw, 300)
    pj = rng.integers(0, nw, 300)
    # near pairs (overlapping windows) as well: the bound is tight there
    pj[:60] = np.minimum(pi[:60] + rng.integers(0, 4, 60), nw - 1)
    lb = gpu.paa_pairs(a, m, w, pi, pj)
    ref = paa_fp64(a, m, w, pi, pj, oracle.lemire)
    # the FP32 table is widened outwards, so the kernel's value never exceeds the FP64 one by
    # more than its rounding (2^-19 relative, scaled away in the join) ...
    assert np.all(lb <= ref * (1 + 2.0 ** -17) + 1e-9)
    # ... and stays close to it
    assert np.all(lb >= ref - 1e-3 * np.maximum(ref, 1.0))
    assert np.count_nonzero(ref > 0) > 100
    we = min(w, m - 1)
    for k in range(0, 300, 7):
        zi = oracle.znorm(a, int(pi[k]), m)
        zj = oracle.znorm(a, int(pj[k]), m)
        keogh = max(oracle.lb_keogh_sq(zi, zj, we), oracle.lb_keogh_sq(zj, zi, we))
        assert lb[k] <= keogh * (1 + 1e-6) + 1e-9, (k, lb[k], keogh)
        assert lb[k] <= oracle.dtw_sq(zi, zj, we) * (1 + 1e-6) + 1e-9


@pytest.mark.parametrize("name", sorted(n for n in CASES if CASES[n]["w"] >= 6))
def test_answers_do_not_depend_on_the_stage(gpu, name):
    c = CASES[name]
    a = golden_series(name)
    res = {mode: gpu.top1(a, c["m"], c["w"], paa=mode) for mode in (0, 1, 2)}
    for mode in (0, 1):
        assert (res[mode].i, res[mode].j) == (res[2].i, res[2].j), (name, mode)
        assert res[mode].dist == res[2].dist
    # the reference's distance (its printed pair can be a stale seed pair; the parity tests
    # check the pair itself)
    assert abs(res[2].dist - c["ref"]["final_distance"]) <= 5e-7
    on = res[1]
    assert on.pass_paa <= on.pass_kim
    assert res[2].pass_paa == res[2].pass_kim
    # the stage only removes pairs that LB_Keogh would remove as well
    assert on.pass_keogh <= on.pass_paa


def test_stage_prunes_on_a_wide_band(gpu):
    c = CASES["n24576_regime_m1024_w102"]
    a = golden_series("n24576_regime_m1024_w102")
    on = gpu.top1(a, c["m"], c["w"], paa=1)
    off = gpu.top1(a, c["m"], c["w"], paa=2)
    assert (on.i, on.j, on.dist) == (off.i, off.j, off.dist)
    assert on.pass_paa < 0.6 * on.pass_kim
    assert on.keogh_elems < off.keogh_elems


def test_batch_answers_do_not_depend_on_the_stage(gpu):
    series, ms, ws = gpu.batch_series(range(200, 216))
    on = gpu.top1_batch(series, ms, ws, paa=1)
    off = gpu.top1_batch(series, ms, ws, paa=2)
    for x, y in zip(on, off):
        assert (x.i, x.j, x.dist) == (y.i, y.j, y.dist)
