"""Host-side mirror of the reference's top-1 DTW join interface, over the gfx950 C ABI.

The reference exposes the join as two C++ functions and a command line
(GPU_version1/src/AAAA.cu:17-36, GPU_version1/src/DTWMotifDiscoveryGUI.cu:12-28,
GPU_version1/src/new_dtw_motifGUI_malloc.cu:30-31):

    DTWMotifDiscoveryGUI(TS, subseqlen, maxwarp, output_time_file_path)
    new_dtw_motifGUI_malloc(a, subseqLen, maxwarp, mp_ed, mp_ed_index, file)
    GPU_dtw <data> <time_out> <m> <r>

This module provides the same names with the same argument meaning, backed by
lib/libdtwj.so (include/dtwj.h).  Differences: results are returned (and also printed
like the reference), indices are 0-based in the returned object, errors raise DTWJError
instead of exiting.  There is no CPU fallback: if the HIP library is missing or no GPU is
visible, every compute call raises.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdtwj.so")
CLI_PATH = os.path.join(_HERE, "bin", "gpu_dtw")

DTWJ_OK, DTWJ_EINVAL, DTWJ_ENOMEM, DTWJ_EDEVICE, DTWJ_EIO, DTWJ_EOVERFLOW = range(6)


class DTWJError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("dtwj error %d: %s" % (code, message))
        self.code = code
        self.message = message


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("seed_items", ctypes.c_int32), ("rows_per_item", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("verbose", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("autotune", ctypes.c_int32), ("sweep", ctypes.c_int32), ("cand_cap", ctypes.c_int32),
                ("paa", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2)]


class Result(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int64), ("j", ctypes.c_int64), ("dist", ctypes.c_double),
                ("dist_fp32", ctypes.c_double), ("seed_dist", ctypes.c_double),
                ("pairs_total", ctypes.c_int64), ("pairs_swept", ctypes.c_int64), ("pass_ball", ctypes.c_int64),
                ("pass_kim", ctypes.c_int64), ("pass_keogh", ctypes.c_int64), ("dtw_abandoned", ctypes.c_int64),
                ("dtw_cells", ctypes.c_int64), ("keogh_elems", ctypes.c_int64), ("candidates", ctypes.c_int64),
                ("t_total", ctypes.c_double), ("t_prep", ctypes.c_double), ("t_seed", ctypes.c_double),
                ("t_join", ctypes.c_double), ("t_verify", ctypes.c_double), ("join_kernel_ms", ctypes.c_double),
                ("items", ctypes.c_int32), ("rows_per_item", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("reruns", ctypes.c_int32), ("pairs_swept_one_sided", ctypes.c_int64),
                ("items_total", ctypes.c_int64), ("pass_paa", ctypes.c_int64)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("pairs_total", ctypes.c_int64), ("items", ctypes.c_int64), ("rows_per_item", ctypes.c_int32),
                ("diag_groups", ctypes.c_int32), ("device_bytes", ctypes.c_int64), ("reserved", ctypes.c_int64 * 4)]


MLP_IN, MLP_HID = 40, 64


class MLP(ctypes.Structure):
    """dtwj_mlp: the learned proposal's 40 -> 64 -> 1 network (FP32 master copy)."""
    _fields_ = [("w1", (ctypes.c_float * MLP_IN) * MLP_HID), ("b1", ctypes.c_float * MLP_HID),
                ("w2", ctypes.c_float * MLP_HID), ("b2", ctypes.c_float), ("feat_mean", ctypes.c_float * MLP_IN),
                ("feat_scale", ctypes.c_float * MLP_IN)]

    def to_dict(self):
        return {"w1": np.ctypeslib.as_array(self.w1).tolist(), "b1": list(self.b1), "w2": list(self.w2),
                "b2": float(self.b2), "feat_mean": list(self.feat_mean), "feat_scale": list(self.feat_scale)}

    @classmethod
    def from_dict(cls, d):
        m = cls()
        np.ctypeslib.as_array(m.w1)[:] = np.asarray(d["w1"], dtype=np.float32)
        for name in ("b1", "w2", "feat_mean", "feat_scale"):
            np.ctypeslib.as_array(getattr(m, name))[:] = np.asarray(d[name], dtype=np.float32)
        m.b2 = float(d["b2"])
        return m

    def arrays(self):
        return (np.ctypeslib.as_array(self.w1).copy(), np.ctypeslib.as_array(self.b1).copy(),
                np.ctypeslib.as_array(self.w2).copy(), float(self.b2))


class ProposalReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int64), ("verified", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("seed_dist", ctypes.c_double), ("seed_dist_ed", ctypes.c_double), ("regret", ctypes.c_double),
                ("regret_ed", ctypes.c_double), ("t_propose", ctypes.c_double), ("reserved", ctypes.c_double * 3)]


class Series(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)), ("n", ctypes.c_int64), ("m", ctypes.c_int32),
                ("w", ctypes.c_int32)]


class GenSpec(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("seed", ctypes.c_uint64), ("step_std", ctypes.c_double),
                ("step_std_alt", ctypes.c_double), ("regime_len", ctypes.c_int64), ("n_plants", ctypes.c_int32),
                ("plant_len", ctypes.c_int32), ("warp", ctypes.c_double), ("noise_std", ctypes.c_double),
                ("plant_src", ctypes.POINTER(ctypes.c_int64)), ("plant_dst", ctypes.POINTER(ctypes.c_int64))]


# Every symbol include/dtwj.h declares (checked by the ABI test).
EXPORTED = [
    "dtwj_default_options", "dtwj_band_from_ratio", "dtwj_top1", "dtwj_top1_seeded", "dtwj_context_create",
    "dtwj_context_destroy", "dtwj_top1_device", "dtwj_context_upload", "dtwj_motif_discovery", "dtwj_load_text",
    "dtwj_free", "dtwj_last_error", "dtwj_device_count", "dtwj_sliding_stats", "dtwj_envelope", "dtwj_dtw_pairs",
    "dtwj_gen_spec_default", "dtwj_generate", "dtwj_top1_batch",
    "dtwj_context_top1_batch", "dtwj_proposal_candidates", "dtwj_mlp_forward", "dtwj_mlp_gradient",
    "dtwj_mlp_train", "dtwj_top1_learned", "dtwj_plan", "dtwj_top1_device_noalloc", "dtwj_matrix_profile",
    "dtwj_paa_pairs", "dtwj_lb_petitjean_pairs", "dtwj_global_mask", "dtwj_masked_ball_pairs",
]

_lib = None


def lib():
    """Loads lib/libdtwj.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DTWJError(DTWJ_EDEVICE, "HIP library not built: %s (run build.py)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    d_p, i64_p, i32_p = P(ctypes.c_double), P(ctypes.c_int64), P(ctypes.c_int32)
    L.dtwj_last_error.restype = ctypes.c_char_p
    L.dtwj_band_from_ratio.argtypes = [ctypes.c_int32, ctypes.c_double]
    L.dtwj_band_from_ratio.restype = ctypes.c_int32
    L.dtwj_device_count.restype = ctypes.c_int32
    L.dtwj_default_options.argtypes = [P(Options)]
    L.dtwj_top1.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, P(Options), P(Result)]
    L.dtwj_top1_seeded.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, d_p, i32_p, ctypes.c_int64,
                                   P(Options), P(Result)]
    L.dtwj_motif_discovery.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                       P(Options), P(Result)]
    L.dtwj_context_create.argtypes = [ctypes.c_int32, P(ctypes.c_void_p)]
    L.dtwj_context_destroy.argtypes = [ctypes.c_void_p]
    L.dtwj_context_upload.argtypes = [ctypes.c_void_p, d_p, ctypes.c_int64, P(ctypes.c_void_p)]
    L.dtwj_top1_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                   P(Options), P(Result)]
    L.dtwj_load_text.argtypes = [ctypes.c_char_p, P(d_p), i64_p]
    L.dtwj_free.argtypes = [ctypes.c_void_p]
    L.dtwj_sliding_stats.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, d_p, d_p]
    L.dtwj_envelope.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, d_p, d_p]
    L.dtwj_dtw_pairs.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, i64_p, i64_p, ctypes.c_int64,
                                 ctypes.c_int32, d_p]
    L.dtwj_paa_pairs.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, i64_p, i64_p, ctypes.c_int64, d_p]
    L.dtwj_lb_petitjean_pairs.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, i64_p, i64_p,
                                          ctypes.c_int64, d_p]
    L.dtwj_global_mask.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, P(ctypes.c_uint8), d_p, d_p,
                                   i32_p]
    L.dtwj_masked_ball_pairs.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, i64_p, i64_p,
                                         ctypes.c_int64, d_p]
    L.dtwj_top1_batch.argtypes = [P(Series), ctypes.c_int32, P(Options), P(Result), i32_p]
    L.dtwj_context_top1_batch.argtypes = [ctypes.c_void_p, P(Series), ctypes.c_int32, P(Options), P(Result), i32_p]
    f_p = P(ctypes.c_float)
    L.dtwj_proposal_candidates.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int64, i64_p, i64_p, f_p, i64_p]
    L.dtwj_mlp_forward.argtypes = [P(MLP), f_p, ctypes.c_int64, f_p]
    L.dtwj_mlp_gradient.argtypes = [P(MLP), f_p, f_p, ctypes.c_int64, P(MLP), d_p]
    L.dtwj_mlp_train.argtypes = [P(MLP), f_p, f_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                 ctypes.c_uint64, d_p]
    L.dtwj_top1_learned.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, P(MLP), ctypes.c_int32,
                                    ctypes.c_int32, P(Options), P(Result), P(ProposalReport)]
    L.dtwj_plan.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, P(Options), P(PlanInfo)]
    L.dtwj_matrix_profile.argtypes = [d_p, ctypes.c_int64, ctypes.c_int32, d_p, i32_p]
    L.dtwj_top1_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                           P(Options), ctypes.c_void_p, P(Result)]
    L.dtwj_gen_spec_default.argtypes = [P(GenSpec)]
    L.dtwj_generate.argtypes = [P(GenSpec), d_p, i64_p, i64_p]
    _lib = L
    return L


def _check(rc):
    if rc != DTWJ_OK:
        raise DTWJError(rc, lib().dtwj_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _as_series(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if a.ndim != 1:
        raise ValueError("series must be one-dimensional")
    return a


@dataclass
class Top1:
    """Result of one join; i, j are 0-based (the reference prints i+1, j+1)."""
    i: int
    j: int
    dist: float
    dist_fp32: float
    seed_dist: float
    pairs_total: int
    pairs_swept: int
    pass_ball: int
    pass_kim: int
    pass_keogh: int
    dtw_abandoned: int
    dtw_cells: int
    keogh_elems: int
    candidates: int
    t_total: float
    t_prep: float
    t_seed: float
    t_join: float
    t_verify: float
    join_kernel_ms: float
    items: int
    rows_per_item: int = 0
    waves_per_cu: int = 0
    pairs_swept_one_sided: int = 0
    reruns: int = 0
    items_total: int = 0
    pass_paa: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def first_min(self):
        """1-based index of the first window, as the reference prints it."""
        return self.i + 1

    @property
    def sec_min(self):
        return self.j + 1


def _result(r):
    return Top1(**{k: getattr(r, k) for k, _ in Result._fields_ if not k.startswith("reserved")})


def make_options(device=0, seed_items=0, rows_per_item=0, waves_per_cu=0, verbose=False, graph=0, autotune=0,
                 sweep=0, cand_cap=0, paa=0):
    o = Options()
    lib().dtwj_default_options(ctypes.byref(o))
    o.device, o.seed_items, o.rows_per_item, o.waves_per_cu, o.verbose = (
        device, seed_items, rows_per_item, waves_per_cu, int(bool(verbose)))
    o.graph, o.autotune, o.sweep, o.cand_cap = int(graph), int(autotune), int(sweep), int(cand_cap)
    o.paa = int(paa)
    return o


def device_count():
    return int(lib().dtwj_device_count())


def band_from_ratio(m, r):
    """w = MAX(round(m*r), 1) as the reference CLI computes it (GPU_version1/src/AAAA.cu:34)."""
    return int(lib().dtwj_band_from_ratio(int(m), float(r)))


def plan(n, m, w, **opts):
    """Host-side work plan of a top-1 join (dtwj_plan): items, tile height, device bytes."""
    out = PlanInfo()
    _check(lib().dtwj_plan(int(n), int(m), int(w), ctypes.byref(make_options(**opts)), ctypes.byref(out)))
    return out


def top1(series, m, w, **opts):
    """Exact top-1 pair (smallest banded-DTW distance, j - i >= m) on the GPU."""
    a = _as_series(series)
    r = Result()
    _check(lib().dtwj_top1(_dptr(a), a.size, int(m), int(w), ctypes.byref(make_options(**opts)), ctypes.byref(r)))
    return _result(r)


def _batch_args(series_list, ms, ws):
    arrs = [_as_series(x) for x in series_list]
    k = len(arrs)
    if k == 0 or len(ms) != k or len(ws) != k:
        raise ValueError("need one m and one w per series")
    items = (Series * k)()
    for t, a in enumerate(arrs):
        items[t].data = _dptr(a)
        items[t].n = a.size
        items[t].m = int(ms[t])
        items[t].w = int(ws[t])
    return arrs, items, (Result * k)(), (ctypes.c_int32 * k)()


def _batch_results(out, status):
    res = []
    for t in range(len(status)):
        if status[t] == DTWJ_OK:
            res.append(_result(out[t]))
        else:
            res.append(DTWJError(int(status[t]), "series %d rejected" % t))
    return res


def top1_batch(series_list, ms, ws, **opts):
    """Batched form: independent exact top-1 joins of many series (uneven lengths, mixed
    (m, w)) in one persistent join launch fed from one device work queue, captured with the
    preprocessing and seeding in one HIP graph.  Returns a list with, per series, a Top1 or the
    DTWJError that series raised."""
    arrs, items, out, status = _batch_args(series_list, ms, ws)
    _check(lib().dtwj_top1_batch(items, len(arrs), ctypes.byref(make_options(**opts)), out, status))
    return _batch_results(out, status)


# ---- learned proposal ---------------------------------------------------------------
def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def proposal_candidates(series, m, w, n_items=4096):
    """Candidate pairs (closest ED pair of every lane of n_items sampled sweep items) and their
    raw features, computed on the GPU.  Returns (pi, pj, features[count, 40])."""
    a = _as_series(series)
    cap = 64 * max(int(n_items), 1)
    pi = np.zeros(cap, dtype=np.int64)
    pj = np.zeros(cap, dtype=np.int64)
    feats = np.zeros((cap, MLP_IN), dtype=np.float32)
    cnt = ctypes.c_int64()
    _check(lib().dtwj_proposal_candidates(_dptr(a), a.size, int(m), int(w), int(n_items), cap,
                                          pi.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                          pj.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _fptr(feats),
                                          ctypes.byref(cnt)))
    k = cnt.value
    return pi[:k].copy(), pj[:k].copy(), feats[:k].copy()


def _feats(x):
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if f.ndim != 2 or f.shape[1] != MLP_IN:
        raise ValueError("features must be (count, %d)" % MLP_IN)
    return f


def proposal_targets(features, cost, m):
    """Training target of the network for candidates with exact squared DTW cost `cost`:
    log(DTW distance / ED distance), ED^2 = m expm1(feature 36)."""
    f = np.asarray(features, dtype=np.float64)
    ed2 = m * np.expm1(f[:, 36])
    return (0.5 * (np.log(np.maximum(cost, 1e-30)) - np.log(np.maximum(ed2, 1e-30)))).astype(np.float32)


def proposal_rank_key(features, y, m):
    """Ranking key of the learned proposal: log ED distance + network output."""
    f = np.asarray(features, dtype=np.float64)
    return 0.5 * np.log(m * np.expm1(f[:, 36]) + 1e-30) + np.asarray(y, dtype=np.float64)


def mlp_forward(model, features):
    """Network output (predicted log(DTW / ED) distance ratio) of every feature row, on the GPU."""
    f = _feats(features)
    out = np.zeros(f.shape[0], dtype=np.float32)
    _check(lib().dtwj_mlp_forward(ctypes.byref(model), _fptr(f), f.shape[0], _fptr(out)))
    return out


def mlp_gradient(model, features, targets):
    """Mean-squared-error gradient over all rows in one GPU pass (the training kernels).
    Returns (grad MLP, loss)."""
    f = _feats(features)
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.float32))
    g = MLP()
    loss = ctypes.c_double()
    _check(lib().dtwj_mlp_gradient(ctypes.byref(model), _fptr(f), _fptr(t), f.shape[0], ctypes.byref(g),
                                   ctypes.byref(loss)))
    return g, loss.value


def mlp_train(model, features, targets, epochs=4, batch=4096, lr=1e-3, seed=0):
    """Trains model in place on the GPU (Adam); returns the final full-batch loss."""
    f = _feats(features)
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.float32))
    loss = ctypes.c_double()
    _check(lib().dtwj_mlp_train(ctypes.byref(model), _fptr(f), _fptr(t), f.shape[0], int(epochs), int(batch),
                                float(lr), int(seed), ctypes.byref(loss)))
    return loss.value


def mlp_init(features, seed=0, targets=None):
    """A fresh network whose feature normalisation is fitted to `features`; with `targets`,
    the output starts near their mean (small output weights, bias = mean)."""
    f = _feats(features).astype(np.float64)
    m = MLP()
    rng = np.random.default_rng(seed)
    np.ctypeslib.as_array(m.w1)[:] = (rng.standard_normal((MLP_HID, MLP_IN)) * np.sqrt(2.0 / MLP_IN)).astype(np.float32)
    w2_scale = np.sqrt(1.0 / MLP_HID) * (0.05 if targets is not None else 1.0)
    np.ctypeslib.as_array(m.w2)[:] = (rng.standard_normal(MLP_HID) * w2_scale).astype(np.float32)
    if targets is not None:
        m.b2 = float(np.mean(targets))
    mean = f.mean(axis=0)
    std = f.std(axis=0)
    np.ctypeslib.as_array(m.feat_mean)[:] = mean.astype(np.float32)
    np.ctypeslib.as_array(m.feat_scale)[:] = (1.0 / np.where(std > 1e-12, std, 1.0)).astype(np.float32)
    return m


def load_mlp(path):
    import json
    with open(path) as fh:
        return MLP.from_dict(json.load(fh))


def save_mlp(model, path):
    import json
    with open(path, "w") as fh:
        json.dump(model.to_dict(), fh)


def top1_learned(series, m, w, model, n_items=4096, n_verify=1024, **opts):
    """The learned form of the join: the network orders candidate pairs, the n_verify best are
    verified by exact DTW and seed the plain join (whose answer is returned).  Returns
    (Top1, report dict)."""
    a = _as_series(series)
    r = Result()
    rep = ProposalReport()
    _check(lib().dtwj_top1_learned(_dptr(a), a.size, int(m), int(w), ctypes.byref(model), int(n_items),
                                   int(n_verify), ctypes.byref(make_options(**opts)), ctypes.byref(r),
                                   ctypes.byref(rep)))
    report = {k: getattr(rep, k) for k, _ in ProposalReport._fields_ if not k.startswith("reserved")}
    return _result(r), report


def top1_seeded(series, m, w, mp_ed, mp_idx_1based, **opts):
    a = _as_series(series)
    mp = np.ascontiguousarray(np.asarray(mp_ed, dtype=np.float64))
    mi = np.ascontiguousarray(np.asarray(mp_idx_1based, dtype=np.int32))
    r = Result()
    _check(lib().dtwj_top1_seeded(_dptr(a), a.size, int(m), int(w), _dptr(mp),
                                  mi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), mp.size,
                                  ctypes.byref(make_options(**opts)), ctypes.byref(r)))
    return _result(r)


def matrix_profile(series, m):
    """ED matrix profile with exclusion m on the GPU (mpx_v2, GPU_version1/src/mpx_v2.cu:14-150):
    returns (mp_ed, mp_idx_1based), n - m + 1 entries each."""
    a = _as_series(series)
    nw = a.size - int(m) + 1
    mp = np.empty(max(nw, 1))
    mpi = np.empty(max(nw, 1), dtype=np.int32)
    _check(lib().dtwj_matrix_profile(_dptr(a), a.size, int(m), _dptr(mp),
                                     mpi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return mp, mpi


def DTWMotifDiscoveryGUI(TS, subseqlen, maxwarp, output_time_file_path, **opts):
    """Mirror of DTWMotifDiscoveryGUI (GPU_version1/src/DTWMotifDiscoveryGUI.cu:12-28): the ED
    matrix profile, the reference's seed from its 100 smallest entries, then the join for one
    (m, w); prints the reference's result lines and appends the seconds to
    output_time_file_path.  Returns the Top1 result (seed_dist: the printed "input of bsf")."""
    a = _as_series(TS)
    r = Result()
    path = None if output_time_file_path is None else os.fsencode(output_time_file_path)
    _check(lib().dtwj_motif_discovery(_dptr(a), a.size, int(subseqlen), int(maxwarp), path,
                                      ctypes.byref(make_options(**opts)), ctypes.byref(r)))
    return _result(r)


def new_dtw_motifGUI_malloc(a, subseqLen, maxwarp, mp_ed, mp_ed_index, file, **opts):
    """Mirror of new_dtw_motifGUI_malloc (GPU_version1/src/new_dtw_motifGUI_malloc.cu:30-31):
    the join seeded with the ED matrix profile (mp_ed) and its 1-based neighbour index."""
    res = top1_seeded(a, subseqLen, maxwarp, mp_ed, mp_ed_index, **opts)
    print("first min is %d sec min is %d" % (res.first_min, res.sec_min))
    print("total time is %fs   input of bsf:%5.3f  ,final distance :%f  " % (res.t_total, res.seed_dist, res.dist))
    if file is not None:
        with open(file, "a") as f:
            f.write("%f\n" % res.t_total)
    return res


def loadfile(filename):
    """Mirror of loadfile (GPU_version1/src/loadfile.cu:14-33); FP64, raises on a bad file."""
    p = ctypes.POINTER(ctypes.c_double)()
    n = ctypes.c_int64()
    _check(lib().dtwj_load_text(os.fsencode(filename), ctypes.byref(p), ctypes.byref(n)))
    try:
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0)
    finally:
        lib().dtwj_free(p)


def sliding_stats(series, m):
    a = _as_series(series)
    mu = np.empty(a.size - m + 1)
    sg = np.empty(a.size - m + 1)
    _check(lib().dtwj_sliding_stats(_dptr(a), a.size, int(m), _dptr(mu), _dptr(sg)))
    return mu, sg


def envelope(series, w):
    a = _as_series(series)
    lo = np.empty(a.size)
    up = np.empty(a.size)
    _check(lib().dtwj_envelope(_dptr(a), a.size, int(w), _dptr(lo), _dptr(up)))
    return lo, up


def dtw_pairs(series, m, w, pi, pj, prec=32):
    """Squared banded-DTW cost of z-normalised windows for each (pi[k], pj[k])."""
    a = _as_series(series)
    pi = np.ascontiguousarray(np.asarray(pi, dtype=np.int64))
    pj = np.ascontiguousarray(np.asarray(pj, dtype=np.int64))
    out = np.empty(pi.size)
    i64 = ctypes.POINTER(ctypes.c_int64)
    _check(lib().dtwj_dtw_pairs(_dptr(a), a.size, int(m), int(w), pi.ctypes.data_as(i64), pj.ctypes.data_as(i64),
                                pi.size, int(prec), _dptr(out)))
    return out


def lb_petitjean_pairs(series, m, w, pi, pj):
    """LB_Keogh of window pj[k] against window pi[k]'s envelope plus the LB_Petitjean
    increments, FP64 on the GPU, for each window pair (pi[k], pj[k])."""
    a = _as_series(series)
    pi = np.ascontiguousarray(np.asarray(pi, dtype=np.int64))
    pj = np.ascontiguousarray(np.asarray(pj, dtype=np.int64))
    out = np.empty(pi.size)
    i64 = ctypes.POINTER(ctypes.c_int64)
    _check(lib().dtwj_lb_petitjean_pairs(_dptr(a), a.size, int(m), int(w), pi.ctypes.data_as(i64),
                                         pj.ctypes.data_as(i64), pi.size, _dptr(out)))
    return out


def paa_pairs(series, m, w, pi, pj):
    """Segment-mean lower bound of LB_Keogh (the join's stage between LB_Kim and LB_Keogh, as
    the kernel computes it in FP32) for each window pair (pi[k], pj[k])."""
    a = _as_series(series)
    pi = np.ascontiguousarray(np.asarray(pi, dtype=np.int64))
    pj = np.ascontiguousarray(np.asarray(pj, dtype=np.int64))
    out = np.empty(pi.size)
    i64 = ctypes.POINTER(ctypes.c_int64)
    _check(lib().dtwj_paa_pairs(_dptr(a), a.size, int(m), int(w), pi.ctypes.data_as(i64), pj.ctypes.data_as(i64),
                                pi.size, _dptr(out)))
    return out


def global_mask(series, m, w):
    """The reference's global occupancy mask on the GPU (compute_shared_data_global +
    compute_MASK_global): (mask uint8[n], occupancy statistic float64[n], threshold, table length)."""
    a = _as_series(series)
    mask = np.zeros(a.size, dtype=np.uint8)
    occ = np.zeros(a.size)
    ths = ctypes.c_double(0)
    ln = ctypes.c_int32(0)
    _check(lib().dtwj_global_mask(_dptr(a), a.size, int(m), int(w), mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  _dptr(occ), ctypes.byref(ths), ctypes.byref(ln)))
    return mask, occ, ths.value, ln.value


def masked_ball_pairs(series, m, w, pi, pj):
    """The reference's masked ball bound (diag_mask_global) of explicit pairs on the GPU, FP64:
    array (npairs, 2): envelope of window pi against window pj, then the reverse."""
    a = _as_series(series)
    pi = np.ascontiguousarray(pi, dtype=np.int64)
    pj = np.ascontiguousarray(pj, dtype=np.int64)
    out = np.zeros((pi.size, 2))
    P = ctypes.POINTER
    _check(lib().dtwj_masked_ball_pairs(_dptr(a), a.size, int(m), int(w), pi.ctypes.data_as(P(ctypes.c_int64)),
                                        pj.ctypes.data_as(P(ctypes.c_int64)), pi.size, _dptr(out)))
    return out


def generate(n, seed, step_std=1.0, step_std_alt=0.0, regime_len=0, n_plants=0, plant_len=128, warp=0.05,
             noise_std=0.1, plant_src=None, plant_dst=None):
    """Seeded random walk with planted time-warped motif pairs (see csrc/series_gen.cpp).
    Returns (series, src_offsets, dst_offsets)."""
    s = GenSpec()
    lib().dtwj_gen_spec_default(ctypes.byref(s))
    s.n, s.seed, s.step_std, s.step_std_alt, s.regime_len = int(n), int(seed), step_std, step_std_alt, int(regime_len)
    s.n_plants, s.plant_len, s.warp, s.noise_std = int(n_plants), int(plant_len), warp, noise_std
    keep = []
    if plant_src is not None:
        ps = np.ascontiguousarray(np.asarray(plant_src, dtype=np.int64))
        pd = np.ascontiguousarray(np.asarray(plant_dst, dtype=np.int64))
        keep += [ps, pd]
        i64 = ctypes.POINTER(ctypes.c_int64)
        s.plant_src, s.plant_dst = ps.ctypes.data_as(i64), pd.ctypes.data_as(i64)
    out = np.empty(int(n))
    src = np.zeros(max(int(n_plants), 1), dtype=np.int64)
    dst = np.zeros(max(int(n_plants), 1), dtype=np.int64)
    i64 = ctypes.POINTER(ctypes.c_int64)
    _check(lib().dtwj_generate(ctypes.byref(s), _dptr(out), src.ctypes.data_as(i64), dst.ctypes.data_as(i64)))
    return out, src[:int(n_plants)], dst[:int(n_plants)]


def batch_spec(seed):
    """(n, m, band ratio) of the batched workload's series `seed`: a random walk whose length is
    uniform in [8192, 262144], m drawn from {64, 128, 256, 512} and the band ratio from
    {0.01, 0.05, 0.1}, all from numpy's default_rng(seed); one planted motif of length m."""
    r = np.random.default_rng(seed)
    n = int(r.integers(8192, 262145))
    m = int(r.choice([64, 128, 256, 512]))
    ratio = float(r.choice([0.01, 0.05, 0.1]))
    return n, m, ratio


def batch_series(seeds):
    """The batched workload's series for the given seeds: (list of FP32-valued series, ms, ws)."""
    series, ms, ws = [], [], []
    for s in seeds:
        n, m, ratio = batch_spec(s)
        a, _, _ = generate(n, s, n_plants=1, plant_len=m)
        series.append(a.astype(np.float32).astype(np.float64))
        ms.append(m)
        ws.append(band_from_ratio(m, ratio))
    return series, ms, ws


class Context:
    """A device context that keeps buffers between joins (for repeated, device-resident runs)."""

    def __init__(self, device=0):
        self._c = ctypes.c_void_p()
        _check(lib().dtwj_context_create(int(device), ctypes.byref(self._c)))
        self._d = ctypes.c_void_p()
        self.n = 0

    def upload(self, series):
        a = _as_series(series)
        _check(lib().dtwj_context_upload(self._c, _dptr(a), a.size, ctypes.byref(self._d)))
        self.n = a.size

    def top1(self, m, w, **opts):
        r = Result()
        _check(lib().dtwj_top1_device(self._c, self._d, self.n, int(m), int(w), ctypes.byref(make_options(**opts)),
                                      ctypes.byref(r)))
        return _result(r)

    def top1_noalloc(self, m, w, stream=None, **opts):
        """The same join on the uploaded series through dtwj_top1_device_noalloc (the library's
        fixed per-device workspace; no device allocation; the context only holds the input)."""
        r = Result()
        _check(lib().dtwj_top1_device_noalloc(self._d, self.n, int(m), int(w), ctypes.byref(make_options(**opts)),
                                              stream, ctypes.byref(r)))
        return _result(r)

    def top1_batch(self, series_list, ms, ws, **opts):
        """top1_batch on this context (keeps device memory and autotuner choices)."""
        arrs, items, out, status = _batch_args(series_list, ms, ws)
        _check(lib().dtwj_context_top1_batch(self._c, items, len(arrs), ctypes.byref(make_options(**opts)), out,
                                             status))
        return _batch_results(out, status)

    def close(self):
        if self._c:
            lib().dtwj_context_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
