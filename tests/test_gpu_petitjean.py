"""GPU: LB_Keogh plus the LB_Petitjean increments of explicit window pairs (FP64) against the
oracle's restatement of the reference's lbPetitjean_new (CPU_version1/LB_PETITJEAN.cpp:9-123,
whose DTW counts the oracle reproduces on every golden case), relative 1e-9; and as a lower
bound: at least LB_Keogh of the same direction, at most the exact squared DTW."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,w,seed", [(3000, 64, 6, 1), (4000, 256, 13, 2), (5000, 512, 51, 3), (6000, 1024, 102, 4),
                                        (900, 8, 2, 5)])
def test_lb_petitjean_matches_oracle_and_bounds_dtw(gpu, oracle, n, m, w, seed):
    rng = np.random.default_rng(seed)
    a = np.cumsum(rng.standard_normal(n))
    nw = n - m + 1
    pi = rng.integers(0, nw, 40)
    pj = rng.integers(0, nw, 40)
    got = gpu.lb_petitjean_pairs(a, m, w, pi, pj)
    dtw = gpu.dtw_pairs(a, m, w, pi, pj, 64)
    for k in range(pi.size):
        ref = oracle.lb_petitjean_pair(a, m, w, pi[k], pj[k])
        assert abs(got[k] - ref) <= 1e-9 * max(ref, 1.0), (k, got[k], ref)
        zi, zj = oracle.znorm(a, int(pi[k]), m), oracle.znorm(a, int(pj[k]), m)
        assert got[k] >= oracle.lb_keogh_sq(zi, zj, w) * (1 - 1e-12) - 1e-12
        assert got[k] <= dtw[k] * (1 + 1e-9) + 1e-9


def test_lb_petitjean_rejects_bad_arguments(gpu):
    a = np.cumsum(np.random.default_rng(0).standard_normal(500))
    with pytest.raises(Exception):
        gpu.lb_petitjean_pairs(a, 64, 6, [0], [500])
    with pytest.raises(Exception):
        gpu.lb_petitjean_pairs(a, 6, 2, [0], [10])
