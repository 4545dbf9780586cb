"""GPU parity of the ED matrix profile (dtwj_matrix_profile, the reference's mpx_v2 with
minlag = m) and of the reference pipeline built on it (dtwj_motif_discovery / gpu_dtw: the
profile's 100 smallest entries seed the join and are printed as "input of bsf").

Oracles: orc_mpx (the reference's mpx_v2 restated, FP64), itself pinned here by a brute-force
NumPy profile on a small series; the compiled reference's printed seed and result lines in the
golden fixtures.  Tolerances: profile values relative 1e-9 where the partner is the oracle's;
partner indices equal except where two partners' squared distances agree to 2^-21 relative
(both distances then within 5e-7 relative); printed lines identical.
"""
import re
import subprocess

import numpy as np
import pytest

from conftest import golden_cases, golden_series

pytestmark = pytest.mark.gpu
CASES = golden_cases()


def brute_profile(a, m):
    W = np.lib.stride_tricks.sliding_window_view(a, m)
    Z = (W - W.mean(1, keepdims=True)) / W.std(1, keepdims=True)
    D2 = np.maximum(2 * m - 2 * (Z @ Z.T), 0.0)
    nw = Z.shape[0]
    idx = np.arange(nw)
    D2[np.abs(idx[:, None] - idx[None, :]) < m] = np.inf
    return np.sqrt(D2.min(1)), D2.argmin(1) + 1


def check_profile(a, m, mp, mpi, mp0, mpi0):
    assert mp.shape == mp0.shape and mpi.shape == mpi0.shape
    same = mpi == mpi0
    rel = np.abs(mp - mp0) / np.maximum(mp0, 1e-6)
    assert np.all(rel[same] <= 1e-9), rel[same].max()
    # a different partner only where the two are tied at the sweep's key precision (2^-21 in ED^2)
    assert np.all(rel[~same] <= 5e-7), rel[~same].max()
    assert (~same).mean() <= 1e-3
    nw = a.size - m + 1
    none = mpi0 == 0  # no window m or more away (n < 3m - 1): sqrt(4m), index 0
    assert np.array_equal(mpi[none], mpi0[none]) and np.all(mp[none] == np.sqrt(4.0 * m))
    has = ~none
    assert np.all((mpi[has] >= 1) & (mpi[has] <= nw))
    assert np.all(np.abs(mpi[has] - 1 - np.arange(nw)[has]) >= m)


@pytest.mark.parametrize("n,m,seed", [(1500, 64, 5), (5000, 32, 6), (12000, 256, 7), (3000, 1024, 8),
                                      (2050, 1024, 9)])
def test_matrix_profile_matches_oracle(gpu, oracle, n, m, seed):
    a = np.cumsum(np.random.default_rng(seed).standard_normal(n)) + 1e3 * seed
    mp, mpi = gpu.matrix_profile(a, m)
    mp0, mpi0 = oracle.mpx(a, m)
    check_profile(a, m, mp, mpi, mp0, mpi0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_matrix_profile_on_goldens(gpu, oracle, name):
    c = CASES[name]
    a = golden_series(name)
    mp, mpi = gpu.matrix_profile(a, c["m"])
    mp0, mpi0 = oracle.mpx(a, c["m"])
    check_profile(a, c["m"], mp, mpi, mp0, mpi0)


def test_matrix_profile_rejects_bad_input(gpu):
    a = np.cumsum(np.random.default_rng(1).standard_normal(3000))
    b = a.copy()
    b[100:400] = 2.0
    with pytest.raises(gpu.DTWJError) as e:
        gpu.matrix_profile(b, 128)
    assert e.value.code == gpu.DTWJ_EINVAL
    b = a.copy()
    b[7] = np.inf
    with pytest.raises(gpu.DTWJError):
        gpu.matrix_profile(b, 128)


@pytest.mark.parametrize("name", sorted(CASES))
def test_pipeline_prints_the_reference_seed_and_result(gpu, oracle, name, tmp_path):
    """gpu_dtw on the golden series prints the compiled reference's "input of bsf" (the seed
    from the profile) and final distance; the pair line is the reference's where its sweep
    found the pair, else the true pair (the oracle's)."""
    c = CASES[name]
    a = golden_series(name)
    data = tmp_path / "ts.txt"
    np.savetxt(data, a, fmt="%.17g")
    out = subprocess.run([gpu.CLI_PATH, str(data), str(tmp_path / "t.txt"), str(c["m"]), "w=%d" % c["w"]],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    ref = c["ref"]
    assert "input of bsf:%5.3f" % ref["input_bsf"] in out.stdout, (out.stdout, ref["input_bsf"])
    assert "final distance :%f" % ref["final_distance"] in out.stdout
    o = oracle.motif_discovery(a, c["m"], c["w"])
    assert abs(o.seed_bsf - float(re.search(r"input of bsf:\s*(\S+)", out.stdout).group(1))) <= 5e-4
    assert "first min is %d sec min is %d" % (o.i + 1, o.j + 1) in out.stdout
    if o.found_in_sweep:
        assert (o.i + 1, o.j + 1) == (int(ref["first_min"]), int(ref["sec_min"]))
