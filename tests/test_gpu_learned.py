"""GPU checks of the learned proposal: candidate features against a NumPy restatement, the
MFMA forward and gradient kernels against the FP64 oracle (oracle/mlp_oracle.py), training,
exactness of the learned join, and an ablation showing what the trained network gains over
the ED ordering it replaces (rank correlation with exact DTW, distance of its first picks).

Tolerances: features relative 1e-3 (FP32 on the GPU); network output absolute 2e-3 (FP32
accumulation over BF16 operands, the oracle on the same BF16 operands); gradient relative
5e-3 of each block's norm; the learned join's answer bit-identical to the plain join's.
"""
import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _np_features(a, i, j, m, w):
    n = a.size
    pad_hi = np.concatenate([np.full(w, -np.inf), a, np.full(w, -np.inf)])
    pad_lo = np.concatenate([np.full(w, np.inf), a, np.full(w, np.inf)])
    from numpy.lib.stride_tricks import sliding_window_view as swv
    U = swv(pad_hi, 2 * w + 1).max(1)
    L = swv(pad_lo, 2 * w + 1).min(1)

    def znorm(x, k):
        win = x[k:k + m]
        return win, a[k:k + m].mean(), a[k:k + m].std()

    ai, mi, si = znorm(a, i)
    aj, mj, sj = znorm(a, j)
    zi, zj = (ai - mi) / si, (aj - mj) / sj
    Ui, Li = (U[i:i + m] - mi) / si, (L[i:i + m] - mi) / si
    Uj, Lj = (U[j:j + m] - mj) / sj, (L[j:j + m] - mj) / sj
    f = np.zeros(40)
    for s in range(16):
        lo, hi = s * m // 16, (s + 1) * m // 16
        f[s] = zi[lo:hi].mean()
        f[16 + s] = zj[lo:hi].mean()
    lbA = np.sum(np.maximum(np.maximum(zj - Ui, Li - zj), 0) ** 2)
    lbB = np.sum(np.maximum(np.maximum(zi - Uj, Lj - zi), 0) ** 2)
    Ri, Rj = np.linalg.norm(Ui - Li) / 2, np.linalg.norm(Uj - Lj) / 2
    ballA = max(np.linalg.norm(zj - (Ui + Li) / 2) - Ri, 0) ** 2
    ballB = max(np.linalg.norm(zi - (Uj + Lj) / 2) - Rj, 0) ** 2
    ed = np.sum((zi - zj) ** 2)
    f[32:37] = np.log1p(np.array([lbA, lbB, ballA, ballB, ed]) / m)
    f[38] = (j - i) / n
    f[39] = np.mean(np.abs(f[:16] - f[16:32]))
    return f


def test_candidate_features_match_numpy(gpu):
    a, _, _ = gpu.generate(20000, 7, n_plants=1, plant_len=128)
    m, w = 128, 6
    pi, pj, F = gpu.proposal_candidates(a, m, w, n_items=64)
    assert len(pi) > 500
    assert np.all(pj - pi >= m) and np.all(pj <= a.size - m)
    for k in range(0, len(pi), max(1, len(pi) // 25)):
        ref = _np_features(a, int(pi[k]), int(pj[k]), m, w)
        got = F[k].astype(np.float64)
        sel = [c for c in range(40) if c != 37]  # 37: LB_Kim, checked below
        assert np.allclose(got[sel], ref[sel], rtol=1e-3, atol=1e-4), (k, np.max(np.abs(got[sel] - ref[sel])))
        # LB_Kim never exceeds the exact DTW of the pair
        cost = gpu.dtw_pairs(a, m, w, pi[k:k + 1], pj[k:k + 1], 64)[0]
        assert np.expm1(got[37]) * m <= cost * (1 + 1e-4) + 1e-6


@pytest.mark.parametrize("n,n_items", [(20000, 32), (20000, 100), (400000, 4096)])
def test_candidates_come_sorted_and_distinct(gpu, n, n_items):
    """The device sort (a bitonic network over the padded power-of-two key buffer: 2048, 8192 and
    262,144 keys here, the last with stages merged through HBM) returns the candidates in
    strictly increasing (i, j) order; the features above pin the pairs themselves."""
    a, _, _ = gpu.generate(n, 11, n_plants=1, plant_len=128)
    pi, pj, _ = gpu.proposal_candidates(a, 128, 6, n_items=n_items)
    assert 0.1 * 64 * n_items < len(pi) <= 64 * n_items
    key = pi.astype(np.int64) * (1 << 32) + pj
    assert np.all(np.diff(key) > 0)
    assert np.all(pj - pi >= 128) and np.all(pj <= a.size - 128)


def _random_model(gpu, rng, feats):
    model = gpu.mlp_init(feats, seed=int(rng.integers(1 << 30)))
    np.ctypeslib.as_array(model.b1)[:] = rng.standard_normal(64).astype(np.float32) * 0.2
    model.b2 = 0.3
    return model


def test_mlp_forward_matches_oracle(gpu):
    rng = np.random.default_rng(11)
    feats = rng.standard_normal((1000, 40)).astype(np.float32) * 2 + 1
    model = _random_model(gpu, rng, feats)
    w1, b1, w2, b2 = model.arrays()
    y_gpu = gpu.mlp_forward(model, feats)
    y_ref, _, _ = mo.forward(feats, w1, b1, w2, b2, np.ctypeslib.as_array(model.feat_mean),
                             np.ctypeslib.as_array(model.feat_scale))
    assert np.max(np.abs(y_gpu - y_ref)) < 2e-3 * (1 + np.max(np.abs(y_ref)))


@pytest.mark.parametrize("count", [16, 1000, 4099])
def test_mlp_gradient_matches_oracle(gpu, count):
    rng = np.random.default_rng(count)
    feats = rng.standard_normal((count, 40)).astype(np.float32)
    targets = rng.standard_normal(count).astype(np.float32)
    model = _random_model(gpu, rng, feats)
    w1, b1, w2, b2 = model.arrays()
    g, loss = gpu.mlp_gradient(model, feats, targets)
    rl, dW1, db1, dw2, db2 = mo.loss_and_grad(feats, targets, w1, b1, w2, b2,
                                              np.ctypeslib.as_array(model.feat_mean),
                                              np.ctypeslib.as_array(model.feat_scale))
    gw1, gb1, gw2, gb2 = g.arrays()
    assert abs(loss - rl) <= 1e-4 * (1 + rl)
    for got, ref in [(gw1, dW1), (gb1, db1), (gw2, dw2)]:
        assert np.linalg.norm(got - ref) <= 5e-3 * (np.linalg.norm(ref) + 1e-6), (np.linalg.norm(got - ref),
                                                                                   np.linalg.norm(ref))
    assert abs(gb2 - db2) <= 5e-3 * (abs(db2) + 1e-6) + 1e-6


def test_training_fits_a_learnable_target(gpu):
    rng = np.random.default_rng(3)
    feats = rng.standard_normal((20000, 40)).astype(np.float32)
    coef = rng.standard_normal(40)
    targets = (np.tanh(feats.astype(np.float64) @ coef / 6) + 0.1 * feats[:, 0] ** 2).astype(np.float32)
    model = gpu.mlp_init(feats, seed=1)
    _, loss0 = gpu.mlp_gradient(model, feats, targets)
    loss1 = gpu.mlp_train(model, feats, targets, epochs=8, batch=1024, lr=3e-3, seed=2)
    assert loss1 < 0.2 * loss0, (loss0, loss1)


def tail_biased_sample(F, k, seed):
    """k candidate indices: a third from the lowest-ED tenth (where the top-1 lives), the rest
    uniform over all candidates."""
    rng = np.random.default_rng(seed)
    c = F.shape[0]
    k = min(k, c)
    low = np.argsort(F[:, 36])[: max(c // 10, 1)]
    a = rng.choice(low, size=min(k // 3, low.size), replace=False)
    rest = np.setdiff1d(np.arange(c), a)
    b = rng.choice(rest, size=min(k - a.size, rest.size), replace=False)
    return np.concatenate([a, b])


def _training_set(gpu, seeds, n, m, w, per_series):
    X, T = [], []
    for s in seeds:
        a, _, _ = gpu.generate(n, s, n_plants=1, plant_len=m)
        pi, pj, F = gpu.proposal_candidates(a, m, w, n_items=1024)
        take = tail_biased_sample(F, per_series, s)
        cost = gpu.dtw_pairs(a, m, w, pi[take], pj[take], 64)
        X.append(F[take])
        T.append(gpu.proposal_targets(F[take], cost, m))
    return np.concatenate(X), np.concatenate(T)


def test_learned_join_is_exact_and_ranks_better_than_ed(gpu):
    """Ablation: the same candidates ranked by the trained network (log ED + predicted
    log(DTW / ED)) versus by ED alone.  The network must rank held-out candidates closer to
    their exact DTW order (Spearman), put closer pairs among its first 64, and the learned join
    must return the plain join's exact answer."""
    from scipy.stats import spearmanr
    m, w, n = 128, 6, 60000
    X, T = _training_set(gpu, range(1000, 1012), n, m, w, 6000)
    model = gpu.mlp_init(X, seed=0, targets=T)
    _, loss0 = gpu.mlp_gradient(model, X, T)
    loss1 = gpu.mlp_train(model, X, T, epochs=30, batch=1024, lr=1e-3, seed=0)
    assert loss1 < 0.5 * T.var(), (loss0, loss1, T.var())
    rho_l, rho_e, top_l, top_e = [], [], [], []
    for s in range(2000, 2004):
        a, _, _ = gpu.generate(n, s, n_plants=1, plant_len=m)
        pi, pj, F = gpu.proposal_candidates(a, m, w, n_items=1024)
        y = gpu.mlp_forward(model, F)
        key = gpu.proposal_rank_key(F, y, m)
        edk = gpu.proposal_rank_key(F, np.zeros_like(y), m)
        # held-out labels: the 2000 lowest-ED candidates and 1000 random ones
        sel = np.unique(np.concatenate([np.argsort(edk)[:2000],
                                        np.random.default_rng(s).choice(len(pi), 1000, replace=False)]))
        dist = np.sqrt(gpu.dtw_pairs(a, m, w, pi[sel], pj[sel], 64))
        rho_l.append(spearmanr(key[sel], dist).correlation)
        rho_e.append(spearmanr(edk[sel], dist).correlation)
        top_l.append(np.mean(dist[np.argsort(key[sel])[:64]]))
        top_e.append(np.mean(dist[np.argsort(edk[sel])[:64]]))
        plain = gpu.top1(a, m, w)
        learned, rep = gpu.top1_learned(a, m, w, model, n_items=1024, n_verify=64)
        assert (learned.i, learned.j, learned.dist) == (plain.i, plain.j, plain.dist)
        assert rep["verified"] == 64 and rep["candidates"] > 1000
        assert rep["regret"] >= 0.0 and rep["seed_dist"] >= learned.dist * (1 - 1e-12)
    print("spearman learned %.3f ed %.3f; mean top-64 distance learned %.4f ed %.4f" % (
        np.mean(rho_l), np.mean(rho_e), np.mean(top_l), np.mean(top_e)))
    assert np.mean(rho_l) > np.mean(rho_e) + 0.02
    assert np.mean(top_l) < np.mean(top_e)


@pytest.mark.parametrize("name", ["n16384_m256_w13", "n16384_m256_w26", "n16384_m256_w3", "n16384_m256_w2"])
def test_committed_model_join_matches_the_reference(gpu, oracle, name):
    """The learned join with the shipped network (models/proposal_m256_w13.json, trained on
    generator seeds 1000..1099) on the m = 256 golden series: the compiled reference's pair and
    distance (tests/golden/golden.json) and the oracle's FP64 distance."""
    import os
    from conftest import PKG_DIR, golden_cases, golden_series
    c = golden_cases()[name]
    a = golden_series(name)
    model = gpu.load_mlp(os.path.join(PKG_DIR, "models", "proposal_m256_w13.json"))
    r, rep = gpu.top1_learned(a, c["m"], c["w"], model, n_items=4096, n_verify=256)
    o = oracle.motif_discovery(a, c["m"], c["w"])
    assert (r.i, r.j) == (o.i, o.j)
    assert abs(r.dist - o.bsf) <= 1e-9 * o.bsf
    assert abs(r.dist - c["ref"]["final_distance"]) <= 5e-7 + 1e-5 * r.dist
    if o.found_in_sweep:
        assert (r.i + 1, r.j + 1) == (int(c["ref"]["first_min"]), int(c["ref"]["sec_min"]))
    assert rep["verified"] > 0 and rep["seed_dist"] >= r.dist * (1 - 1e-9)
