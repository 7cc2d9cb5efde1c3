"""Tuning of the engine's launch shapes on a workload: child subspaces per
level-synchronous batch and, for many data sets at once, run_many's worker
streams.

    python -m gpu_inscy_amd.autotune --config 2 [--phi] [--runs 3] [--out tuning.json]
    python -m gpu_inscy_amd.autotune --config 3 --out tuning.json

Each candidate setting runs the workload `runs` times after one warm-up run
and is timed by the wall clock around the call (labels back in host memory).
Every setting's result must equal the first setting's: the settings change
launch shapes, not results, and a difference raises.  The fastest setting
(lowest median) is applied to this process and written, with every
measurement, to the JSON file; apply(path) applies a saved file in another
process.  The reference has no counterpart (its batch shapes are constants).
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import statistics
import time

import numpy as np

from . import _lib, datagen
from . import inscy as INSCY

BATCHES = (256, 512, 1024, 2048)
WORKERS = (2, 4, 8)


def _same(a, b) -> bool:
    return a.subspaces == b.subspaces and np.array_equal(a.labels, b.labels)


def _single(w, X, settings, runs):
    """(setting, times) per setting for one data set on the level-synchronous driver."""
    out, ref = [], None
    for mb, wk in settings:
        _lib.set_tuning(mb, wk)
        run = lambda: INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular,
                                w.entropy_order, mode=_lib.MODE_BATCHED)
        res = run()
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            res = run()
            times.append(time.perf_counter() - t0)
        if ref is None:
            ref = res
        elif not _same(ref, res):
            raise RuntimeError(f"max_batch={mb} changed the result")
        out.append(((mb, wk), times))
    return out


def _many(sets, jobs, settings, runs):
    """(setting, times) per setting for run_many over the data sets."""
    out, ref = [], None
    for mb, wk in settings:
        _lib.set_tuning(mb, wk)
        res = INSCY.run_many(jobs, mode=_lib.MODE_BATCHED)
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            res = INSCY.run_many(jobs, mode=_lib.MODE_BATCHED)
            times.append(time.perf_counter() - t0)
        if ref is None:
            ref = res
        elif not all(_same(a, b) for a, b in zip(ref, res)):
            raise RuntimeError(f"max_batch={mb} workers={wk} changed a result")
        out.append(((mb, wk), times))
    return out


def tune(config: int, phi: bool = False, runs: int = 3, batches=BATCHES, workers=WORKERS) -> dict:
    """Measures every setting on the workload (configs of datagen: 3 is the 64
    data sets through run_many) and applies the fastest; returns the record."""
    import torch
    if config == 3:
        sets = datagen.multi_instance()
        if phi:
            sets = [dataclasses.replace(w, rectangular=False) for w in sets]
        jobs = [(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular,
                 w.entropy_order) for w in sets]
        settings = [(mb, wk) for mb in batches for wk in workers]
        meas = _many(sets, jobs, settings, runs)
        name = "64 data sets (run_many)"
    else:
        w = datagen.config(config)
        if phi:
            w = dataclasses.replace(w, rectangular=False)
        X = torch.from_numpy(w.X).cuda()
        settings = [(mb, 0) for mb in batches]
        meas = _single(w, X, settings, runs)
        name = w.name
    best = min(meas, key=lambda m: statistics.median(m[1]))[0]
    _lib.set_tuning(*best)
    return {"workload": name, "config": config, "rectangular": not phi, "runs": runs,
            "results": [{"max_batch": s[0], "workers": s[1], "seconds": [round(t, 5) for t in ts],
                         "median": round(statistics.median(ts), 5)} for s, ts in meas],
            "best": {"max_batch": best[0], "workers": best[1]}}


def apply(path: str) -> dict:
    """Applies a tuning file written by this module; returns its best setting."""
    with open(path) as f:
        best = json.load(f)["best"]
    _lib.set_tuning(int(best["max_batch"]), int(best["workers"]))
    return best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", type=int, default=2, help="datagen configuration (3: the 64 data sets)")
    ap.add_argument("--phi", action="store_true", help="rectangular=False (the phi density)")
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--out", default=None, help="JSON file for the measurements and the best setting")
    a = ap.parse_args(argv)
    rec = tune(a.config, a.phi, a.runs)
    text = json.dumps(rec, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text, flush=True)
    return rec


if __name__ == "__main__":
    main()
