"""GPU parity of the no-allocation entry point (dtwj_top1_device_noalloc) against the allocating
one it mirrors (dtwj_top1_device), through the C ABI: identical pair, FP64 distance and FP32
distance, and the same pairs swept, on two calls in a row (the second with a smaller series,
so that state left in the shared fixed workspace by the first call must not leak into it),
plus the size limit of the workspace and a call on an explicit stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert (a.i, a.j) == (b.i, b.j)
    assert a.dist == b.dist and a.dist_fp32 == b.dist_fp32
    assert a.pairs_total == b.pairs_total and a.pairs_swept == b.pairs_swept


@pytest.mark.parametrize("first,second", [((100000, 256, 13, 41), (30000, 128, 6, 42)),
                                          ((131072, 1024, 102, 43), (20000, 64, 2, 44))])
def test_noalloc_matches_the_allocating_entry_point(gpu, first, second):
    for n, m, w, seed in (first, second):
        a, _, _ = gpu.generate(n, seed, n_plants=1, plant_len=m)
        ctx = gpu.Context()
        try:
            ctx.upload(a)
            ref = ctx.top1(m, w)
            got = ctx.top1_noalloc(m, w)
            _same(got, ref)
            assert got.reruns == 0
        finally:
            ctx.close()


def test_noalloc_rejects_series_beyond_the_workspace(gpu):
    a = np.cumsum(np.random.default_rng(1).standard_normal(400000))
    ctx = gpu.Context()
    try:
        ctx.upload(a)
        with pytest.raises(gpu.DTWJError) as e:
            ctx.top1_noalloc(256, 13)
        assert e.value.code == gpu.DTWJ_EINVAL and "fixed workspace" in e.value.message
        # the workspace is still usable afterwards
        b = a[:50000]
        ctx.upload(b)
        _same(ctx.top1_noalloc(256, 13), ctx.top1(256, 13))
    finally:
        ctx.close()


def test_noalloc_on_an_explicit_stream(gpu):
    """The caller's stream, created here by the HIP runtime the library itself uses."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # already loaded as a dependency of the library
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    a, _, _ = gpu.generate(40000, 45, n_plants=1, plant_len=128)
    ctx = gpu.Context()
    try:
        ctx.upload(a)
        got = ctx.top1_noalloc(128, 6, stream=stream)
        _same(got, ctx.top1(128, 6))
    finally:
        ctx.close()
        hip.hipStreamDestroy(stream)
