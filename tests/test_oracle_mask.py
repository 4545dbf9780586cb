"""CPU: the oracle's global occupancy mask and masked ball bound, pinned independently.

The oracle restates compute_shared_data_global + compute_MASK_global
(CPU_version1/compute_shared_data_global.cpp:12-46, CPU_version1/compute_MASK_global.cpp:11-93) and
diag_mask_global (CPU_version1/diag_mask_global.cpp:11-226); its global prune count already equals
the compiled reference's on every golden case (test_oracle_golden.py), which pins both through the
join.  Here each is also checked on its own:
  * the occupancy statistic against a brute-force evaluation of its definition on a tiny series
    (every window's clipped envelope from scratch, counts by direct comparison with the bin
    edges, no cumulative tables), bit for bit, and the mask against its threshold;
  * the running-sum masked ball bound along a diagonal against a direct NumPy sum of
    (z - Uhat)^2 + (z - Lhat)^2 over the masked offsets, and against the banded DTW it bounds.
"""
import math

import numpy as np
import pytest

from conftest import golden_series


def rw(n, seed):
    return np.cumsum(np.random.default_rng(seed).standard_normal(n))


def find_pos(v, ln):
    p = math.floor(v * 100 + ln * 0.5)
    return min(max(p, 1), ln)


def occupancy_bruteforce(a, m, w, mu, sig):
    """The statistic of compute_MASK_global by definition, in the reference's operation order."""
    n = a.size
    nw = n - m + 1
    z = [[(a[i + k] - mu[i]) / sig[i] for k in range(m)] for i in range(nw)]
    zmax = max(max(abs(v) for v in row) for row in z)
    ln = int(max(600.0, math.floor(zmax) * 2 * 100))
    # bins of every offset's values, kept as lists: the cumulative count cdf[k][p - 1] is the
    # number of windows whose offset-k value falls in a bin <= p
    bins = [[find_pos(z[i][k], ln) for i in range(nw)] for k in range(m)]
    nd = [0.0] * n
    for i in range(nw):
        for k in range(m):
            lo, hi = max(0, k - w), min(m - 1, k + w)
            U = max(z[i][lo:hi + 1])
            L = min(z[i][lo:hi + 1])
            pu, pl = find_pos(U + 0.0, ln), find_pos(L - 0.0, ln)
            cu = sum(1 for b in bins[k] if b <= pu)
            cl = sum(1 for b in bins[k] if b <= pl)
            nd[i + k] += float(cu - cl) / nw
    out = np.zeros(n)
    for t in range(n):
        mod = t + 1.0 if t < m else (float(m) if t < n - m else float(m) - (t - (n - m)))
        out[t] = nd[t] / mod
    s = 0.0
    for v in out:
        s += v
    return out, s / n, ln


@pytest.mark.parametrize("n,m,w,seed", [(90, 16, 2, 1), (120, 20, 4, 2), (70, 16, 0, 3), (64, 16, 9, 4)])
def test_occupancy_statistic_matches_its_definition(oracle, n, m, w, seed):
    a = rw(n, seed)
    mu, sig = oracle.mvmean(a, m)
    mask, diff, ths, ln = oracle.global_mask(a, m, w)
    ref, ref_ths, ref_ln = occupancy_bruteforce(a, m, w, mu, sig)
    assert ln == ref_ln
    assert np.array_equal(diff, ref)
    assert ths == ref_ths
    assert np.array_equal(mask, (ref <= ref_ths).astype(np.uint8))
    if w > 0:
        assert 0 < mask.sum() < n
    else:  # envelope = value: every count is 0 and every position at the (zero) mean
        assert mask.sum() == n


def masked_ball_direct(a, m, w, mask, mu, sig, lemire, e, d):
    """(sqrt(2 dist - DUL2) - DUL) / 2 with dist = sum_K (z_d - Uhat_e)^2 + (z_d - Lhat_e)^2 over
    K = {k in [3, m-4]: mask[e + k]}, DUL2 over the masked offsets of the whole window."""
    L, U = lemire
    k = np.arange(m)
    msk = mask[e:e + m].astype(bool)
    uh = (U[e:e + m] - mu[e]) / sig[e]
    lh = (L[e:e + m] - mu[e]) / sig[e]
    zd = (a[d:d + m] - mu[d]) / sig[d]
    K = msk & (k >= 3) & (k <= m - 4)
    dist = np.sum((zd[K] - uh[K]) ** 2 + (zd[K] - lh[K]) ** 2)
    dul2 = np.sum(((U[e:e + m] - L[e:e + m]) / sig[e])[msk] ** 2)
    return 0.0 if dist < dul2 else 0.5 * (math.sqrt(2 * dist - dul2) - math.sqrt(dul2))


@pytest.mark.parametrize("name,m,w,diag", [("cfg1_n16384_m128_w6", 128, 6, 300), ("n8192_m64_w3", 64, 3, 2000),
                                           ("n16384_m256_w13", 256, 13, 5000)])
def test_masked_ball_running_sums_match_direct_sums(oracle, name, m, w, diag):
    a = golden_series(name)[:6000]
    mu, sig = oracle.mvmean(a, m)
    mask, _, _, _ = oracle.global_mask(a, m, w)
    fir, sec = oracle.masked_ball_diag(a, m, w, diag)
    env = oracle.lemire(a, w)
    rows = list(range(1, len(fir), max(1, len(fir) // 40)))
    for lo in rows:
        hi = lo + diag - 1
        f = masked_ball_direct(a, m, w, mask, mu, sig, env, lo, hi)
        s = masked_ball_direct(a, m, w, mask, mu, sig, env, hi, lo)
        assert abs(fir[lo] - f) <= 1e-7 * (1 + abs(f)), (lo, fir[lo], f)
        assert abs(sec[lo] - s) <= 1e-7 * (1 + abs(s)), (lo, sec[lo], s)
        # a lower bound of the banded DTW distance of the pair
        zi = (a[lo:lo + m] - mu[lo]) / sig[lo]
        zj = (a[hi:hi + m] - mu[hi]) / sig[hi]
        dtw = math.sqrt(oracle.dtw_sq(zi, zj, w))
        assert max(fir[lo], sec[lo]) <= dtw * (1 + 1e-12)
