"""Trains the learned proposal's network on the GPU and writes it to models/.

Training data come from the plain method itself: for every series of the generator's seeds
(default 1000..1099, n = 1,000,000, 5 planted motifs, m = 256, w = 13), the GPU proposes
candidate pairs (the closest ED pair of every lane of 4096 sampled sweep items), computes their
40 features, and labels a subset (a third from the lowest-ED tenth, the rest uniform) with the exact FP64
banded DTW distance (the library's
own DTW kernel).  The network (40 -> 64 -> 1, BF16 MFMA forward/backward, Adam) then learns
log(DTW / ED) distance ratio (candidates are ranked by log ED + prediction).

Usage: python train_proposal.py [--seeds 1000:1100] [--per-series 20000] [--n 1000000]
                                [--m 256] [--w 13] [--epochs 12] [--out models/proposal_m256_w13.json]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    spec = importlib.util.spec_from_file_location("dtwj_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


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


def training_set(d, seeds, n, m, w, per_series, plants, log=print):
    X, T = [], []
    for s in seeds:
        a, _, _ = d.generate(n, s, n_plants=plants, plant_len=m)
        pi, pj, F = d.proposal_candidates(a, m, w, n_items=4096)
        take = tail_biased_sample(F, per_series, s)
        cost = d.dtw_pairs(a, m, w, pi[take], pj[take], 64)
        X.append(F[take])
        T.append(d.proposal_targets(F[take], cost, m))
        log("seed %d: %d candidates, %d labelled" % (s, len(pi), len(take)))
    return np.concatenate(X), np.concatenate(T)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="1000:1100")
    ap.add_argument("--per-series", type=int, default=20000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--w", type=int, default=13)
    ap.add_argument("--plants", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "models", "proposal_m256_w13.json"))
    args = ap.parse_args()
    lo, hi = (int(x) for x in args.seeds.split(":"))
    d = load_pkg()
    t0 = time.time()
    X, T = training_set(d, range(lo, hi), args.n, args.m, args.w, args.per_series, args.plants,
                        log=lambda s: print(s, flush=True))
    t1 = time.time()
    model = d.mlp_init(X, seed=0, targets=T)
    _, loss0 = d.mlp_gradient(model, X, T)
    loss = d.mlp_train(model, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=0)
    t2 = time.time()
    out = model.to_dict()
    out["meta"] = {"seeds": [lo, hi], "n": args.n, "m": args.m, "w": args.w, "plants": args.plants,
                   "samples": int(X.shape[0]), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
                   "loss_initial": loss0, "loss_final": loss, "target": "log(DTW / ED) distance ratio",
                   "seconds_data": t1 - t0, "seconds_train": t2 - t1}
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as fh:
        json.dump(out, fh)
    print(json.dumps(out["meta"]), flush=True)


if __name__ == "__main__":
    sys.exit(main())
