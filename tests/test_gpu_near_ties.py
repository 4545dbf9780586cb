"""GPU parity on input with very many near-ties: an exactly periodic integer series, where every
pair of windows whose lag is a multiple of the period has banded DTW cost exactly 0 (identical
windows give bitwise identical statistics and z-values, in FP32 and in FP64).

Every such pair is within the FP32 error margin of the best, so the join pushes all of them as
candidates for the FP64 re-verification.  When they outnumber the candidate buffer the join is
re-run from its own best-so-far with a buffer sized to the count (result field `reruns`), never
failing and never writing past the buffer.  The exact answer under the stated tie-break (smallest
i, then smallest j) is pinned by the oracle's brute force (textbook DP over every pair).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tiled(n, period, seed):
    """Integer random walk of one period, repeated: exact ties at lags k * period."""
    base = np.cumsum(np.random.default_rng(seed).integers(-3, 4, size=period)).astype(np.float64)
    return np.tile(base, n // period + 1)[:n]


def near_tie_pairs(n, m, period):
    nw = n - m + 1
    k0 = -(-m // period)  # smallest multiple of the period that is >= m
    return sum(max(0, (nw - 1 - i) // period - k0 + 1) for i in range(nw))


@pytest.mark.parametrize("n,m,w,period", [(3000, 32, 1, 50), (2500, 40, 4, 64)])
def test_regrown_candidate_buffer_matches_brute_force(gpu, oracle, n, m, w, period):
    a = tiled(n, period, 7)
    ties = near_tie_pairs(n, m, period)
    assert ties > 20000  # more than the small buffer below and than a batch series' 2^14 slots
    bi, bj, bd = oracle.bruteforce_top1(a, m, w)
    assert bd == 0.0 and bi == 0 and bj == -(-m // period) * period
    small = gpu.top1(a, m, w, cand_cap=1024)   # overflows: re-run with a regrown buffer
    assert small.reruns >= 1
    assert (small.i, small.j) == (bi, bj) and small.dist == 0.0
    assert small.candidates >= ties           # every exact tie was re-scored in FP64
    full = gpu.top1(a, m, w)                  # 2^20 slots: no re-run
    assert full.reruns == 0
    assert (full.i, full.j, full.dist) == (small.i, small.j, small.dist)


def test_batch_series_with_more_near_ties_than_its_slots(gpu, oracle):
    """A batch keeps 2^14 candidate slots per series; a series that pushes more is joined again
    on its own (the other series keep their batch results)."""
    a = tiled(3000, 50, 7)
    b = np.cumsum(np.random.default_rng(3).standard_normal(4000))
    res = gpu.top1_batch([a, b], [32, 64], [1, 3])
    ra, rb = res
    assert not isinstance(ra, Exception) and not isinstance(rb, Exception)
    bi, bj, bd = oracle.bruteforce_top1(a, 32, 1)
    assert ra.reruns >= 1 and (ra.i, ra.j) == (bi, bj) and ra.dist == 0.0
    o = oracle.motif_discovery(b, 64, 3)
    assert rb.reruns == 0 and (rb.i, rb.j) == (o.i, o.j)
    assert abs(rb.dist - o.bsf) <= 1e-9 * o.bsf
