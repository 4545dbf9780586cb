"""Pins the CPU oracle where the reference has no test: closed forms, invariants and
brute force on tiny inputs (the reference repository ships no tests or fixtures)."""
import numpy as np
import pytest


def rw(n, seed):
    return np.cumsum(np.random.default_rng(seed).standard_normal(n))


def np_dtw_full(x, y, w):
    """Independent full-table DTW (band |r-c| <= w), squared cost."""
    m = len(x)
    D = np.full((m + 1, m + 1), np.inf)
    D[0, 0] = 0.0
    for r in range(1, m + 1):
        for c in range(max(1, r - w), min(m, r + w) + 1):
            D[r, c] = (x[r - 1] - y[c - 1]) ** 2 + min(D[r - 1, c], D[r, c - 1], D[r - 1, c - 1])
    return D[m, m]


def test_mvmean_matches_two_pass(oracle):
    a = rw(5000, 1) + 300.0
    mu, sg = oracle.mvmean(a, 128)
    W = np.lib.stride_tricks.sliding_window_view(a, 128)
    assert np.allclose(mu, W.mean(1), rtol=1e-12, atol=0)
    assert np.allclose(sg, W.std(1), rtol=1e-7, atol=0)  # running sums lose a few digits


def test_lemire_matches_naive(oracle):
    a = rw(3000, 2)
    for w in (0, 1, 6, 51):
        lo, up = oracle.lemire(a, w)
        p_hi = np.concatenate([np.full(w, -np.inf), a, np.full(w, -np.inf)])
        p_lo = np.concatenate([np.full(w, np.inf), a, np.full(w, np.inf)])
        win = np.lib.stride_tricks.sliding_window_view
        assert np.array_equal(up, win(p_hi, 2 * w + 1).max(1))
        assert np.array_equal(lo, win(p_lo, 2 * w + 1).min(1))
        assert np.all(lo <= a) and np.all(a <= up)


def test_dtw_known_answers(oracle):
    x = np.array([0.0, 1.0, 1.0, 2.0])
    y = np.array([0.0, 0.0, 1.0, 2.0])
    assert oracle.dtw_sq(x, y, 0) == 1.0          # w = 0: squared Euclidean distance
    assert oracle.dtw_sq(x, y, 1) == 0.0          # one step of warping aligns them exactly
    x = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    y = np.array([2.0, 3.0, 4.0, 5.0, 6.0])
    assert oracle.dtw_sq(x, y, 0) == 5.0
    assert oracle.dtw_sq(x, y, 1) == 2.0          # (1-2)^2 at the start, (5-6)^2 at the end


def test_dtw_invariants(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        m = int(rng.integers(8, 40))
        x, y = rng.standard_normal(m), rng.standard_normal(m)
        ed = float(np.sum((x - y) ** 2))
        prev = np.inf
        for w in range(0, m):
            d = oracle.dtw_sq(x, y, w)
            assert d <= prev + 1e-12            # monotone non-increasing in w
            assert d <= ed + 1e-12              # DTW_w <= ED
            assert abs(d - oracle.dtw_sq(y, x, w)) <= 1e-12 * max(1, d)  # symmetric
            prev = d
        assert abs(oracle.dtw_sq(x, y, 0) - ed) <= 1e-12 * max(1, ed)
        w = int(rng.integers(0, m))
        assert abs(oracle.dtw_sq(x, y, w) - np_dtw_full(x, y, w)) <= 1e-9 * max(1, ed)


def test_lb_keogh_and_mon_dtw(oracle):
    rng = np.random.default_rng(4)
    a = rw(4000, 5)
    m = 64
    for _ in range(40):
        i, j = sorted(rng.integers(0, a.size - m, size=2))
        zi, zj = oracle.znorm(a, int(i), m), oracle.znorm(a, int(j), m)
        for w in (1, 4, 12):
            d = oracle.dtw_sq(zi, zj, w)
            assert oracle.lb_keogh_sq(zi, zj, w) <= d + 1e-9
            cb = np.zeros(m + w + 1)
            assert abs(oracle.mon_dtw(zi, zj, cb, w, 1e300) - d) <= 1e-9 * max(1.0, d)
            assert oracle.mon_dtw(zi, zj, cb, w, d * 0.999) == 1e20   # exact or "infinite"


def test_mpx_against_brute_force(oracle):
    a = rw(700, 6)
    m = 32
    mp, mpi = oracle.mpx(a, m)
    nw = a.size - m + 1
    Z = np.stack([oracle.znorm(a, i, m) for i in range(nw)])
    for k in range(0, nw, 37):
        d = np.sqrt(np.maximum(((Z - Z[k]) ** 2).sum(1), 0))
        d[max(0, k - m + 1):k + m] = np.inf          # exclusion |j - k| < m
        assert abs(mp[k] - d.min()) <= 1e-6
        assert abs(d[mpi[k] - 1] - d.min()) <= 1e-6  # 1-based neighbour index


@pytest.mark.parametrize("n,m,w,seed", [(500, 32, 1, 7), (600, 32, 3, 8), (450, 40, 6, 9)])
def test_oracle_join_equals_brute_force(oracle, n, m, w, seed):
    a = rw(n, seed)
    bi, bj, bd = oracle.bruteforce_top1(a, m, w)
    r = oracle.motif_discovery(a, m, w)
    assert (r.i, r.j) == (bi, bj)
    assert abs(r.bsf - bd) <= 1e-9 * max(1.0, bd)


def test_oracle_finds_planted_motif(oracle, dtwj):
    a, src, dst = dtwj.generate(6000, 42, n_plants=1, plant_len=96, plant_src=[700], plant_dst=[4100])
    r = oracle.motif_discovery(a, 96, 5)
    assert abs(r.i - 700) <= 5 and abs(r.j - 4100) <= 5


def test_corner_bound_used_by_the_join_is_valid(oracle):
    """The LB_Kim corner bound of the GPU join (join.hip, stage 2): any warping path first
    leaves {0,1}^2 through (2,0),(2,1),(2,2),(1,2),(0,2), so
    D00 + min(D10+D20, min(D10,D11)+D21, D11+D22, min(D01,D11)+D12, D01+D02) plus the mirrored
    far-corner term is at most the DTW cost.  Checked against the oracle DTW."""
    rng = np.random.default_rng(10)
    a = rw(5000, 11)
    sq = lambda u: u * u
    for _ in range(300):
        m = int(rng.integers(8, 80))
        i, j = sorted(rng.integers(0, a.size - m, size=2))
        x, y = oracle.znorm(a, int(i), m), oracle.znorm(a, int(j), m)
        for w in (0, 1, 2, 5):
            c = lambda r, cc: sq(y[r] - x[cc])
            bl = min(c(1, 0) + c(2, 0), min(c(1, 0), c(1, 1)) + c(2, 1), c(1, 1) + c(2, 2),
                     min(c(0, 1), c(1, 1)) + c(1, 2), c(0, 1) + c(0, 2))
            e = lambda r, cc: sq(y[m - 1 - r] - x[m - 1 - cc])
            br = min(e(1, 0) + e(2, 0), min(e(1, 0), e(1, 1)) + e(2, 1), e(1, 1) + e(2, 2),
                     min(e(0, 1), e(1, 1)) + e(1, 2), e(0, 1) + e(0, 2))
            kim = c(0, 0) + bl + e(0, 0) + br
            assert kim <= oracle.dtw_sq(y, x, w) + 1e-9


def test_ball_bound_used_by_the_join_is_valid(oracle):
    """Stage 1 of the GPU join: (||z_j - M_i|| - R_i)_+^2 <= LB_Keogh(z_j, raw envelope of i)
    <= DTW, with M_i, R_i the centre / radius of the normalised raw-series envelope box."""
    rng = np.random.default_rng(12)
    a = rw(6000, 13)
    for _ in range(200):
        m = int(rng.integers(16, 128))
        w = int(rng.integers(1, m // 4))
        lo, up = oracle.lemire(a, w)
        i, j = sorted(rng.integers(0, a.size - m, size=2))
        mu_i, sd_i = a[i:i + m].mean(), a[i:i + m].std()
        U, L = (up[i:i + m] - mu_i) / sd_i, (lo[i:i + m] - mu_i) / sd_i
        zi, zj = oracle.znorm(a, int(i), m), oracle.znorm(a, int(j), m)
        M, R = (U + L) / 2, np.linalg.norm(U - L) / 2
        ball = max(np.linalg.norm(zj - M) - R, 0.0) ** 2
        keogh = np.sum(np.maximum(zj - U, 0) ** 2 + np.maximum(L - zj, 0) ** 2)
        assert ball <= keogh + 1e-9
        assert keogh <= oracle.dtw_sq(zi, zj, w) + 1e-9


def test_oracle_profile_is_the_brute_force_profile(oracle):
    """orc_mpx (the reference's mpx_v2 restated) against a brute-force NumPy ED matrix profile
    with exclusion m: every distance and nearest-neighbour index."""
    a = np.cumsum(np.random.default_rng(5).standard_normal(1500))
    m = 64
    W = np.lib.stride_tricks.sliding_window_view(a, m)
    Z = (W - W.mean(1, keepdims=True)) / W.std(1, keepdims=True)
    D2 = np.maximum(2 * m - 2 * (Z @ Z.T), 0.0)
    idx = np.arange(Z.shape[0])
    D2[np.abs(idx[:, None] - idx[None, :]) < m] = np.inf
    mp0, mpi0 = oracle.mpx(a, m)
    assert np.max(np.abs(mp0 - np.sqrt(D2.min(1))) / np.sqrt(D2.min(1))) < 1e-9
    assert np.array_equal(mpi0, D2.argmin(1) + 1)
