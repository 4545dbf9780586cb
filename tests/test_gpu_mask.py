"""GPU: the reference's global occupancy mask and masked ball bound (dtwj_global_mask,
dtwj_masked_ball_pairs) against the oracle.

The mask is integer output of FP64 arithmetic that the kernels perform in the reference's order
(z-values, bins, each position's sum over its windows, a sequential mean), so given the same
window statistics it must equal the oracle's bit for bit: the oracle is run with the GPU's own
statistics (dtwj_sliding_stats), and the occupancy statistic, its threshold, the table length
and the mask are compared exactly.  With the reference's own (uncompensated) statistics the
masks may only differ at positions whose statistic lies within rounding of the threshold.

The masked ball bound is compared with the oracle's running sums along diagonals (same
statistics and mask) to 1e-7 relative (1e-5 absolute: the oracle's running-sum rounding), and checked to stay below the banded DTW of each pair.
"""
import numpy as np
import pytest

from conftest import golden_series

pytestmark = pytest.mark.gpu

CASES = [("cfg1_n16384_m128_w6", 128, 6), ("n8192_m64_w3", 64, 3), ("n16384_m256_w13", 256, 13),
         ("n16384_m512_w51", 512, 51), ("n16384_m1024_w102", 1024, 102), ("n24576_regime_m1024_w102", 1024, 102),
         ("n8192_m32_w1", 32, 1), ("n16384_m256_w2", 256, 2)]


@pytest.mark.parametrize("name,m,w", CASES)
def test_global_mask_equals_the_oracle_bit_for_bit(gpu, oracle, name, m, w):
    a = golden_series(name)
    mu, sig = gpu.sliding_stats(a, m)
    mask, occ, ths, ln = gpu.global_mask(a, m, w)
    omask, oocc, oths, oln = oracle.global_mask(a, m, w, mu, sig)
    assert ln == oln
    assert np.array_equal(occ, oocc), int(np.sum(occ != oocc))
    assert ths == oths
    assert np.array_equal(mask, omask)
    assert 0 < int(mask.sum()) < a.size


@pytest.mark.parametrize("name,m,w", CASES[:3])
def test_global_mask_near_the_references_own(gpu, oracle, name, m, w):
    a = golden_series(name)
    mask, occ, ths, _ = gpu.global_mask(a, m, w)
    rmask, rocc, rths, _ = oracle.global_mask(a, m, w)  # the reference's running-sum statistics
    diff = np.nonzero(mask != rmask)[0]
    # differences only where the statistic is within rounding of the threshold
    assert np.all(np.abs(rocc[diff] - rths) <= 1e-9 * abs(rths)), diff[:10]
    assert np.allclose(occ, rocc, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n,m,w,seed", [(3000, 16, 9, 5), (5000, 40, 39, 6), (4000, 100, 0, 7)])
def test_global_mask_edge_bands(gpu, oracle, n, m, w, seed):
    """Bands as wide as the window (no interior offsets) and w = 0 (the envelope is the value)."""
    a = np.cumsum(np.random.default_rng(seed).standard_normal(n)) + 50.0
    mu, sig = gpu.sliding_stats(a, m)
    mask, occ, ths, ln = gpu.global_mask(a, m, w)
    omask, oocc, oths, oln = oracle.global_mask(a, m, w, mu, sig)
    assert (ln, ths) == (oln, oths)
    assert np.array_equal(occ, oocc)
    assert np.array_equal(mask, omask)


@pytest.mark.parametrize("name,m,w,diags", [("cfg1_n16384_m128_w6", 128, 6, (129, 1500, 9000)),
                                            ("n16384_m512_w51", 512, 51, (513, 4000)),
                                            ("n24576_regime_m1024_w102", 1024, 102, (1025, 9000))])
def test_masked_ball_matches_the_oracles_running_sums(gpu, oracle, name, m, w, diags):
    a = golden_series(name)
    mu, sig = gpu.sliding_stats(a, m)
    nw = a.size - m + 1
    for diag in diags:
        fir, sec = oracle.masked_ball_diag(a, m, w, diag, mu, sig)
        lo = np.arange(1, nw - diag + 1, max(1, (nw - diag) // 300), dtype=np.int64)
        lb = gpu.masked_ball_pairs(a, m, w, lo, lo + diag - 1)
        # the oracle's running sums carry FP64 rounding of raw squared values (~1e9 on the
        # regime series, cancelling to ~1e2) along the diagonal: ~1e-4 in dist, ~1e-5 in the bound
        assert np.allclose(lb[:, 0], fir[lo], rtol=1e-7, atol=1e-5), np.max(np.abs(lb[:, 0] - fir[lo]))
        assert np.allclose(lb[:, 1], sec[lo], rtol=1e-7, atol=1e-5), np.max(np.abs(lb[:, 1] - sec[lo]))
        assert np.count_nonzero(lb) > 0


@pytest.mark.parametrize("name,m,w", [("cfg1_n16384_m128_w6", 128, 6), ("n16384_m256_w13", 256, 13),
                                      ("n16384_m1024_w102", 1024, 102)])
def test_masked_ball_is_a_lower_bound_of_dtw(gpu, oracle, name, m, w):
    a = golden_series(name)
    nw = a.size - m + 1
    rng = np.random.default_rng(m + w)
    pi = rng.integers(0, nw - m, 200)
    pj = pi + m + (rng.random(200) * (nw - m - pi)).astype(np.int64)
    pj = np.minimum(pj, nw - 1)
    lb = gpu.masked_ball_pairs(a, m, w, pi, pj)
    d = gpu.dtw_pairs(a, m, w, pi, pj, prec=64)
    dist = np.sqrt(d)
    assert np.all(lb[:, 0] <= dist * (1 + 1e-12) + 1e-12)
    assert np.all(lb[:, 1] <= dist * (1 + 1e-12) + 1e-12)


def test_mask_limits_and_errors(gpu):
    a = np.cumsum(np.random.default_rng(1).standard_normal(4000))
    with pytest.raises(gpu.DTWJError):
        gpu.global_mask(a, 600, 257)  # radius beyond the sparse tables
    b = a.copy()
    b[1000:1100] = 3.0  # a constant window of length 64
    with pytest.raises(gpu.DTWJError):
        gpu.global_mask(b, 64, 3)
    with pytest.raises(gpu.DTWJError):
        gpu.masked_ball_pairs(a, 64, 3, np.array([0]), np.array([a.size]))
