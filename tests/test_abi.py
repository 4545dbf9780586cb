"""C ABI checks that need no GPU: the library loads, exports every declared symbol, and
host-side validation / helpers behave (errors are codes, never exits)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, PKG_DIR


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "dtwj.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = re.findall(r"\b(dtwj_\w+)\s*\(", hdr)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(dtwj):
    lib = ctypes.CDLL(dtwj.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), nm
    assert sorted(dtwj.EXPORTED) == names


def test_band_rule(dtwj):
    # w = MAX(round(m * r), 1) (GPU_version1/src/AAAA.cu:34)
    assert dtwj.band_from_ratio(128, 0.05) == 6
    assert dtwj.band_from_ratio(256, 0.01) == 3
    assert dtwj.band_from_ratio(256, 0.05) == 13
    assert dtwj.band_from_ratio(1024, 0.1) == 102
    assert dtwj.band_from_ratio(32, 0.01) == 1
    assert dtwj.band_from_ratio(64, 0.02) == 1


@pytest.mark.parametrize("n,m,w,msg", [(100, 4, 1, "m must"), (100, 16, -1, "w must"), (20, 16, 1, "too short")])
def test_argument_validation(dtwj, n, m, w, msg):
    with pytest.raises(dtwj.DTWJError) as e:
        dtwj.top1(np.cumsum(np.ones(n)), m, w)
    assert e.value.code == dtwj.DTWJ_EINVAL and msg in e.value.message


def test_nan_rejected(dtwj):
    a = np.cumsum(np.random.default_rng(0).standard_normal(500))
    a[77] = np.nan
    with pytest.raises(dtwj.DTWJError) as e:
        dtwj.top1(a, 32, 2)
    assert e.value.code == dtwj.DTWJ_EINVAL


def test_no_silent_cpu_fallback(dtwj):
    if dtwj.device_count() > 0:
        pytest.skip("a GPU is visible")
    a = np.cumsum(np.random.default_rng(0).standard_normal(500))
    with pytest.raises(dtwj.DTWJError) as e:
        dtwj.top1(a, 32, 2)
    assert e.value.code == dtwj.DTWJ_EDEVICE


def test_loader(dtwj, tmp_path):
    p = tmp_path / "ts.txt"
    vals = np.random.default_rng(1).standard_normal(1000) * 1e3
    p.write_text("\n".join("%.17g" % v for v in vals[:500]) + " " + " ".join("%.17g" % v for v in vals[500:]))
    got = dtwj.loadfile(str(p))
    assert np.array_equal(got, vals)
    with pytest.raises(dtwj.DTWJError) as e:
        dtwj.loadfile(str(tmp_path / "missing.txt"))
    assert e.value.code == dtwj.DTWJ_EIO
    q = tmp_path / "stop.txt"
    q.write_text("1 2 3 x 4 5")           # like operator>>, reading stops at the first non-number
    assert list(dtwj.loadfile(str(q))) == [1.0, 2.0, 3.0]


def test_generator_deterministic_and_planted(dtwj):
    a1, s1, d1 = dtwj.generate(20000, 5, n_plants=2, plant_len=128)
    a2, s2, d2 = dtwj.generate(20000, 5, n_plants=2, plant_len=128)
    assert np.array_equal(a1, a2) and list(s1) == list(s2) and list(d1) == list(d2)
    a3, _, _ = dtwj.generate(20000, 6, n_plants=2, plant_len=128)
    assert not np.array_equal(a1, a3)
    for s, d in zip(s1, d1):
        assert d - s >= 2 * 128
        zs = (a1[s:s + 128] - a1[s:s + 128].mean()) / a1[s:s + 128].std()
        zd = (a1[d:d + 128] - a1[d:d + 128].mean()) / a1[d:d + 128].std()
        assert np.corrcoef(zs, zd)[0, 1] > 0.9
    # regime shifts: step std alternates every regime_len samples
    r, _, _ = dtwj.generate(40000, 7, step_std=0.5, step_std_alt=2.0, regime_len=10000)
    st = np.diff(r)
    assert abs(st[:9999].std() - 0.5) < 0.03 and abs(st[10000:19999].std() - 2.0) < 0.1


def test_cli_reports_errors(dtwj, tmp_path):
    exe = dtwj.CLI_PATH
    assert os.path.exists(exe)
    r = subprocess.run([exe, str(tmp_path / "missing.txt"), str(tmp_path / "t.txt"), "64", "0.05"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "Error opening file" in r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2


def test_cli_grid_order_without_a_device(dtwj, tmp_path):
    """--grid runs the reference CPU main's 24 configurations (CPU_version1/AAAA.cpp:30-38) in its
    order; without a HIP device every join reports an error and the exit status is non-zero."""
    if dtwj.device_count() > 0:
        pytest.skip("a HIP device is visible: the GPU test covers the grid")
    data = tmp_path / "ts.txt"
    np.savetxt(data, np.cumsum(np.random.default_rng(5).standard_normal(3000)), fmt="%.17g")
    r = subprocess.run([dtwj.CLI_PATH, str(data), str(tmp_path / "t.txt"), "--grid"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 5
    got = re.findall(r"subsequence length m=(\d+), band radius w=(\d+), n=3000", r.stdout)
    want = [(m, max(int(np.floor(m * q + 0.5)), 1)) for q in (0.01, 0.02, 0.05, 0.1)
            for m in (32, 64, 128, 256, 512, 1024)]
    assert [(int(m), int(w)) for m, w in got] == want
    assert r.stderr.count("error ") == 24
    r = subprocess.run([dtwj.CLI_PATH, str(data), "--grid"], capture_output=True, text=True)
    assert r.returncode == 2
