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


def test_envelope_radius_limit(dtwj):
    with pytest.raises(dtwj.DTWJError) as e:
        dtwj.envelope(np.cumsum(np.ones(10000)), 2000)
    assert e.value.code == dtwj.DTWJ_EINVAL and "1536" in e.value.message


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


# the reference reads with `while (inputFile >> value)` (CPU_version1/loadfile.cpp:24-28); this
# program does exactly that with the C++ standard library, the loader's independent oracle
_STREAM_READER = r"""
#include <cstdio>
#include <fstream>
int main(int argc, char **argv) {
    std::ifstream f(argv[1]);
    double v;
    while (f >> v) std::printf("%.17g\n", v);
    return 0;
}
"""

LOADER_CASES = [
    "1 2 3 x 4 5",          # stops at the first non-number
    "1.5 nan 2",            # nan is not a number for operator>>
    "1 inf 2", "1 -Infinity 2",
    "0x1p3 4",              # hex float: reads 0, then fails at 'x'
    "1 2 1e 3",             # incomplete exponent: the extraction fails, 1 is not read
    "1 2 1e+ 3", "1 2 1e5x 3", "1 2 3.5e-2 4E+1",
    "1.5.3 7",              # two numbers without a separator
    "+.5 -.25 . 3",         # a lone point fails
    "- 3", "+7 -8",
    "1e400 2",              # overflow fails
    "1e-400 2",             # underflow reads a denormal or zero
    "  \t\n 12\r\n13\v14\f15 ",
    "00012.500 -0 1.",
    "5e 6",
]


@pytest.mark.skipif(subprocess.run(["which", "g++"], capture_output=True).returncode != 0, reason="g++ missing")
def test_loader_matches_the_standard_stream_extraction(dtwj, tmp_path):
    src = tmp_path / "rd.cpp"
    src.write_text(_STREAM_READER)
    exe = tmp_path / "rd"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True, capture_output=True)
    for k, text in enumerate(LOADER_CASES):
        p = tmp_path / ("case%d.txt" % k)
        p.write_text(text)
        out = subprocess.run([str(exe), str(p)], check=True, capture_output=True, text=True).stdout
        want = [float(x) for x in out.split()]
        got = list(dtwj.loadfile(str(p)))
        assert got == want, (text, got, want)


def _items_closed_form(n, m, rows):
    """Sweep items of a plan: per group g of 256 diagonal offsets, ceil(rows_g / rows) items,
    rows_g = (n - m + 1) - (m + 256 g) rows."""
    nw, K = n - m + 1, n - 2 * m + 1
    G = (K + 255) // 256
    g = np.arange(G, dtype=np.int64)
    rows_g = nw - (m + 256 * g)
    return int(((rows_g + rows - 1) // rows).sum())


@pytest.mark.parametrize("n,m,w", [(16384, 128, 6), (1_000_000, 256, 2), (1_000_000, 256, 13),
                                   (4_000_000, 1024, 102), (60_000_000, 256, 2), (2_000_000_000, 64, 1)])
def test_plan_item_count_is_64_bit(dtwj, n, m, w):
    """The sweep's work items are claimed from a 64-bit device counter: the plan's count is the
    closed form, also past 2^31 (n = 6e7 at w = 2 gives 2.3e9 items; n = 2e9, 6e15)."""
    p = dtwj.plan(n, m, w)
    assert p.rows_per_item == (3072 if w <= 3 else 2048)
    assert p.items == _items_closed_form(n, m, p.rows_per_item)
    K = n - 2 * m + 1
    assert p.pairs_total == K * (K + 1) // 2
    assert p.diag_groups == (K + 255) // 256
    if n >= 60_000_000:
        assert p.items > 2 ** 31
    assert p.device_bytes > 180 * n  # 180 bytes of tables per sample, plus buffers


def test_plan_rejects_what_the_join_rejects(dtwj):
    with pytest.raises(dtwj.DTWJError):
        dtwj.plan(2 ** 31, 256, 2)
    with pytest.raises(dtwj.DTWJError):
        dtwj.plan(500, 256, 2)
    with pytest.raises(dtwj.DTWJError):
        dtwj.plan(100000, 1024, 600)


def test_plan_counts_the_segment_mean_table(dtwj):
    """The segment-mean stage's table (3 S FP32 values per window, S = 16 below m = 512) is in
    the plan when the stage is on: automatically for w >= 8 and m >= 128, or when forced; paa
    outside 0..2 is rejected."""
    n = 1_000_000
    nw = n - 256 + 1
    off = dtwj.plan(n, 256, 13, paa=2).device_bytes
    assert dtwj.plan(n, 256, 13).device_bytes - off >= 192 * nw
    assert dtwj.plan(n, 256, 13, paa=1).device_bytes == dtwj.plan(n, 256, 13).device_bytes
    assert dtwj.plan(n, 256, 2).device_bytes == dtwj.plan(n, 256, 2, paa=2).device_bytes
    assert dtwj.plan(n, 256, 2, paa=1).device_bytes - dtwj.plan(n, 256, 2).device_bytes >= 192 * nw
    assert dtwj.plan(n, 64, 13).device_bytes == dtwj.plan(n, 64, 13, paa=2).device_bytes
    nw = n - 1024 + 1
    assert dtwj.plan(n, 1024, 102).device_bytes - dtwj.plan(n, 1024, 102, paa=2).device_bytes >= 384 * nw
    for bad in (-1, 3):
        with pytest.raises(dtwj.DTWJError):
            dtwj.plan(n, 256, 13, paa=bad)


def test_result_mirror_maps_every_field(dtwj):
    """Every field of dtwj_result reaches the Python result object (no GPU involved)."""
    r = dtwj.Result()
    for k, (name, _) in enumerate(dtwj.Result._fields_):
        if not name.startswith("reserved"):
            setattr(r, name, k + 1)
    t = dtwj._result(r)
    for k, (name, _) in enumerate(dtwj.Result._fields_):
        if not name.startswith("reserved"):
            assert getattr(t, name) == k + 1, name
