"""Trains the learned pruner on exact runs and saves the network.

    python -m gpu_inscy_amd.train_pruner [--train 1-8] [--holdout 9] [--out PATH]

Training pairs come from INSCY_MODE_COLLECT runs (exact, level-batched) on
data of the config-2 generator (n=100,000, d=20, 20 Zipf-sized clusters in 4-10
dimensions, 15% noise) with seeds 2000 + r for the given repeats r.  Only
pairs of trees with at least --min-depth restricted dimensions are learned
(the others are never skipped).  The network is trained on the GPU with the
engine's own kernels (pruner.Trainer); the threshold is the largest one that
keeps a given recall of the positive pairs of the held-out repeats, chosen
among the recall targets --recalls as the largest threshold whose learned runs
on every held-out data set still reach --f1 cluster F1 against their exact
runs.
Prints one JSON line with the training figures.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, datagen, pruner
from . import inscy as INSCY


def repeats(spec: str) -> list:
    out = []
    for part in spec.split(","):
        if "-" in part:
            a, b = part.split("-")
            out += list(range(int(a), int(b) + 1))
        else:
            out.append(int(part))
    return out


def collect(rep: int, min_depth: int, log=print, keep_run: bool = False):
    import torch
    w = datagen.config(2, rep)
    X = torch.from_numpy(w.X).to("cuda:0")
    t = time.perf_counter()
    res = INSCY.run(X, mode=_lib.MODE_COLLECT, **w.params())
    el = time.perf_counter() - t
    P = res.pairs
    keep = P["info"][:, 0] >= min_depth
    out = {k: np.array(v[keep]) for k, v in P.items()}
    log(f"repeat {rep}: {len(keep)} pairs, {int(keep.sum())} learnable, positive {out['labels'].mean():.3f}, "
        f"collect run {el:.2f} s")
    if keep_run:
        out["run"] = (w, X, res, el)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--train", default="1-8")
    ap.add_argument("--holdout", default="9-11")
    ap.add_argument("--min-depth", type=int, default=2)
    ap.add_argument("--recalls", default="0.999,0.995,0.99,0.98,0.95,0.9")
    ap.add_argument("--f1", type=float, default=0.995)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pos-weight", type=float, default=3.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=pruner.DEFAULT_WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)

    t0 = time.perf_counter()
    parts = [collect(r, args.min_depth, log) for r in repeats(args.train)]
    holds = [collect(r, args.min_depth, log, keep_run=True) for r in repeats(args.holdout)]
    hold = {k: np.concatenate([h[k] for h in holds]) for k in ("features", "labels", "info", "cost")}
    t1 = time.perf_counter()
    F = np.concatenate([p["features"] for p in parts])
    Y = np.concatenate([p["labels"] for p in parts]).astype(np.float32)
    W = np.concatenate([pruner.features_weights(p, args.pos_weight) for p in parts])
    tr = pruner.Trainer(pruner.init_params(args.seed))
    losses = tr.fit(F, Y, W, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    t2 = time.perf_counter()
    prob = tr.forward(hold["features"]).cpu().numpy()
    y = hold["labels"].astype(bool)
    c0 = hold["cost"][:, 0].astype(np.float64)
    # threshold: the learned runs on the held-out sets against their exact runs
    import torch
    weights = tr.weights()
    sweep = []
    for rc in sorted({float(x) for x in args.recalls.split(",")}, reverse=True):
        thr_rc = pruner.choose_threshold(prob, y, rc)
        _lib.set_pruner(weights, thr_rc, args.min_depth)
        f1s, secs = [], []
        for h in holds:
            hw, hX, hexact, _ = h["run"]
            torch.cuda.synchronize()
            t = time.perf_counter()
            le = INSCY.run(hX, mode=_lib.MODE_LEARNED, **hw.params())
            secs.append(round(time.perf_counter() - t, 4))
            f1s.append(pruner.cluster_f1(hexact, le)["f1"])
        sk = prob < thr_rc
        sweep.append({"recall_target": rc, "threshold": thr_rc, "f1": min(f1s), "f1_per_holdout": f1s,
                      "seconds": secs, "exact_seconds": [round(h["run"][3], 4) for h in holds],
                      "pairs_skipped": float(sk.mean()), "work_skipped": float(c0[sk].sum() / max(c0.sum(), 1.0))})
        log(json.dumps(sweep[-1]))
    _lib.set_pruner(None, 0.0)
    ok = [e for e in sweep if e["f1"] >= args.f1]
    best = max(ok, key=lambda e: e["threshold"]) if ok else max(sweep, key=lambda e: e["f1"])
    thr = best["threshold"]
    skip = prob < thr
    train_prob = tr.forward(F).cpu().numpy()
    meta = {
        "generator": "config 2 (n=100000 d=20, 20 Zipf clusters in 4-10 dims, 15% noise), seeds 2000+r",
        "train_repeats": repeats(args.train), "holdout_repeats": repeats(args.holdout), "min_depth": args.min_depth,
        "recall_target": best["recall_target"], "f1_target": args.f1, "holdout_sweep": sweep,
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "pos_weight": args.pos_weight, "train_pairs": int(len(Y)), "train_positive": float(Y.mean()),
        "train_loss_first_last": [losses[0], losses[-1]],
        "train_accuracy": float(((train_prob >= 0.5) == (Y > 0.5)).mean()),
        "holdout_pairs": int(len(y)), "threshold": thr,
        "holdout_positive_recall": float((~skip[y]).mean()) if y.any() else 1.0,
        "holdout_pairs_skipped": float(skip.mean()),
        "holdout_work_skipped": float(c0[skip].sum() / max(c0.sum(), 1.0)),
        "collect_seconds": round(t1 - t0, 2), "train_seconds": round(t2 - t1, 2),
    }
    pruner.save(args.out, tr.weights(), thr, args.min_depth, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
