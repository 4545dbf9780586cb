"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker or the timed CPU baseline; never by the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class JoinResult(ctypes.Structure):
    _fields_ = [("first_min", ctypes.c_int64), ("sec_min", ctypes.c_int64), ("found_in_sweep", ctypes.c_int32),
                ("i", ctypes.c_int64), ("j", ctypes.c_int64), ("bsf", ctypes.c_double),
                ("seed_bsf", ctypes.c_double), ("mp_seed", ctypes.c_double), ("fast_cnt", ctypes.c_double),
                ("global_cnt", ctypes.c_double), ("all_cnt", ctypes.c_int64), ("kk_cnt", ctypes.c_double),
                ("dtw_cnt", ctypes.c_double), ("fast_stage_rounds", ctypes.c_double),
                ("global_rounds", ctypes.c_double), ("seed_i", ctypes.c_int64), ("seed_j", ctypes.c_int64),
                ("pe_cnt", ctypes.c_double), ("local_rounds", ctypes.c_double),
                ("local_down_rounds", ctypes.c_double), ("local_cnt", ctypes.c_double), ("bad_bin", ctypes.c_double)]


def build():
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        lp = ctypes.POINTER(ctypes.c_int64)
        L.orc_mvmean.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, dp, dp]
        L.orc_lemire.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, dp, dp]
        L.orc_moving_mean.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, dp]
        L.orc_mpx.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, dp, ip]
        L.orc_min_v_k.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ip]
        L.orc_mon_dtw.argtypes = [dp, dp, dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]
        L.orc_mon_dtw.restype = ctypes.c_double
        L.orc_join.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, dp, ip,
                               ctypes.POINTER(JoinResult)]
        L.orc_motif_discovery.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(JoinResult)]
        L.orc_dtw_sq.argtypes = [dp, dp, ctypes.c_int32, ctypes.c_int32]
        L.orc_dtw_sq.restype = ctypes.c_double
        L.orc_znorm.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, dp]
        L.orc_bruteforce_top1.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, lp, lp, dp]
        L.orc_lb_keogh_sq.argtypes = [dp, dp, ctypes.c_int32, ctypes.c_int32]
        L.orc_lb_keogh_sq.restype = ctypes.c_double
        L.orc_lb_petitjean_pair.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                            ctypes.c_int64]
        L.orc_lb_petitjean_pair.restype = ctypes.c_double
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.orc_global_mask.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, dp, dp, u8p, dp, dp, ip]
        L.orc_masked_ball_diag.argtypes = [dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, dp, dp,
                                           ctypes.c_int64, dp, dp]
        _L = L
    return _L


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def mvmean(a, m):
    a = _arr(a)
    mu = np.empty(a.size - m + 1)
    sg = np.empty(a.size - m + 1)
    lib().orc_mvmean(_d(a), a.size, m, _d(mu), _d(sg))
    return mu, sg


def lemire(a, w):
    a = _arr(a)
    lo = np.empty(a.size)
    up = np.empty(a.size)
    lib().orc_lemire(_d(a), a.size, w, _d(lo), _d(up))
    return lo, up


def mpx(a, m, minlag=None):
    a = _arr(a)
    mp = np.empty(a.size - m + 1)
    mpi = np.empty(a.size - m + 1, dtype=np.int32)
    rc = lib().orc_mpx(_d(a), a.size, minlag if minlag is not None else m, m, _d(mp),
                       mpi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc:
        raise ValueError("orc_mpx failed (%d)" % rc)
    return mp, mpi


def motif_discovery(a, m, w):
    """Restated reference pipeline (ED matrix profile seed + join)."""
    a = _arr(a)
    r = JoinResult()
    rc = lib().orc_motif_discovery(_d(a), a.size, m, w, ctypes.byref(r))
    if rc:
        raise ValueError("orc_motif_discovery failed (%d)" % rc)
    return r


def join(a, m, w, mp, mpi):
    a = _arr(a)
    mp = _arr(mp)
    mpi = np.ascontiguousarray(np.asarray(mpi, dtype=np.int32))
    r = JoinResult()
    rc = lib().orc_join(_d(a), a.size, m, w, _d(mp), mpi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                        ctypes.byref(r))
    if rc:
        raise ValueError("orc_join failed (%d)" % rc)
    return r


def dtw_sq(x, y, w):
    x = _arr(x)
    y = _arr(y)
    return lib().orc_dtw_sq(_d(x), _d(y), x.size, w)


def znorm(a, i, m):
    a = _arr(a)
    z = np.empty(m)
    if lib().orc_znorm(_d(a), a.size, i, m, _d(z)):
        raise ValueError("constant or out-of-range window")
    return z


def bruteforce_top1(a, m, w):
    a = _arr(a)
    bi = ctypes.c_int64()
    bj = ctypes.c_int64()
    d = ctypes.c_double()
    if lib().orc_bruteforce_top1(_d(a), a.size, m, w, ctypes.byref(bi), ctypes.byref(bj), ctypes.byref(d)):
        raise ValueError("bruteforce failed")
    return bi.value, bj.value, d.value


def lb_petitjean_pair(a, m, w, i, j):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_lb_petitjean_pair(_d(a), a.size, m, w, int(i), int(j))


def lb_keogh_sq(zi, zj, w):
    zi = _arr(zi)
    zj = _arr(zj)
    return lib().orc_lb_keogh_sq(_d(zi), _d(zj), zi.size, w)


def mon_dtw(lines, cols, cb, w, bsf2):
    lines = _arr(lines)
    cols = _arr(cols)
    cb = _arr(cb)
    return lib().orc_mon_dtw(_d(lines), _d(cols), _d(cb), lines.size, w, bsf2)


def global_mask(a, m, w, mu=None, sig=None):
    """Global occupancy mask (uint8, n), occupancy statistic (n), threshold, table length.
    mu, sig: window statistics to use (None: the reference's running sums)."""
    a = _arr(a)
    n = a.size
    mask = np.zeros(n, dtype=np.uint8)
    diff = np.zeros(n)
    ths = ctypes.c_double(0)
    ln = ctypes.c_int32(0)
    mp = _d(_arr(mu)) if mu is not None else None
    sp = _d(_arr(sig)) if sig is not None else None
    rc = lib().orc_global_mask(_d(a), n, m, w, mp, sp, mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _d(diff),
                               ctypes.byref(ths), ctypes.byref(ln))
    if rc:
        raise ValueError("orc_global_mask failed: %d" % rc)
    return mask, diff, ths.value, ln.value


def masked_ball_diag(a, m, w, diag, mu=None, sig=None):
    """The reference's masked ball bound along diagonal diag (pairs (lo, lo + diag - 1)), both
    directions, nothing pruned: (lb_fir, lb_sec), each of length n - m + 2 - diag."""
    a = _arr(a)
    n = a.size
    cnt = n - m + 1 - diag + 1
    fir = np.zeros(max(cnt, 1))
    sec = np.zeros(max(cnt, 1))
    mp = _d(_arr(mu)) if mu is not None else None
    sp = _d(_arr(sig)) if sig is not None else None
    rc = lib().orc_masked_ball_diag(_d(a), n, m, w, mp, sp, int(diag), _d(fir), _d(sec))
    if rc:
        raise ValueError("orc_masked_ball_diag failed: %d" % rc)
    return fir[:cnt], sec[:cnt]
