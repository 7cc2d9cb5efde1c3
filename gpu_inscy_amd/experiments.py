"""Parameter sweeps of the reference's experiment harness on the engine.

    python -m gpu_inscy_amd.experiments METHOD EXPERIMENT REPEATS REAL_NO_CLUSTERS [EXPERIMENT_2ND] [LARGE]
                                        [--out experiments_data]

Same command line, parameter grids and per-run records as
experiments/run.py:1-175: METHOD is INSCY (exact, depth-first driver),
GPU-INSCY (level-batched), GPU-INSCY* or GPU-INSCY-memory (level-batched, the
reference's variants of its GPU driver), or GPU-INSCY-learned (the learned
pruner, gpu_inscy_amd/pruner_weights.npz).  EXPERIMENT (and EXPERIMENT_2ND)
name the swept parameter: real_cls, n, d, c, N_size, F, r, num_obj, min_size,
order.  Every run draws its data with load_synt(d, n, clusters, repeat) and
calls the method with rectangular=True, as run.py does.

Each finished grid point writes `<out>/runs/<method>n..d..c..N_size..F..r..
num_obj..min_size..order..[cl..].npz` holding `times` and `no_clusters`
(one value per repeat) and, per repeat i, `subspaces_i` (int8 [#subspaces, d]
dimension masks) and `clusterings_i` (int32 [#subspaces, n] canonical
labels).  run.py stores the same lists as pickled object arrays; plain
arrays load without unpickling.  Grid points with a file are skipped, as in
run.py.  Grid points the engine does not run are reported and skipped:
d > 128 (the reference's grids stop at d = 100), and entropy_order != 0 with the
learned pruner (the permuted level order runs in the exact and level-batched
modes).
"""

from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import _lib
from . import inscy as INSCY

BASE = {"n": [1500], "c": [4], "d": [15], "N_size": [0.01], "F": [1.], "r": [1.], "num_obj": [8],
        "min_size": [0.05], "order": [0]}
# experiments/run.py:49-117
SWEEPS = {
    "n": [500, 1000, 2000, 4000, 8000],
    "d": [5, 10, 15, 20, 25, 30],
    "c": [2, 4, 6, 8, 10],
    "N_size": [0.001, 0.005, 0.01, 0.02],
    "F": [.5, 1., 1.5, 2., 2.5],
    "r": [0., .2, .5, .7, 1.],
    "num_obj": [2, 4, 8, 16],
    "min_size": [0.01, 0.05, 0.10],
    "order": [0, 1, -1],
}
SWEEPS_2ND = dict(SWEEPS, n=[500, 1000, 2000, 2500, 5000, 10000], d=[2, 5, 10, 15, 25],
                  N_size=[0.001, 0.005, 0.01, 0.02, 0.05])
METHODS = {"INSCY": ("INSCY", _lib.MODE_EXACT), "GPU-INSCY": ("GPU_INSCY", _lib.MODE_BATCHED),
           "GPU-INSCY*": ("GPU_INSCY_star", _lib.MODE_BATCHED),
           "GPU-INSCY-memory": ("GPU_INSCY_memory", _lib.MODE_BATCHED),
           "GPU-INSCY-learned": ("GPU_INSCY_learned", _lib.MODE_LEARNED)}


def grid(experiment: str, real_no_clusters: int, experiment_2nd=None, large: bool = False):
    """(name suffix, parameter lists) of run.py's sweep."""
    p = {k: list(v) for k, v in BASE.items()}
    p["real_cls"] = [real_no_clusters]
    name = ""
    if experiment == "real_cls":
        name += "_real_cls"
        p["real_cls"] = [5, 10, 20, 30, 40]
        if large:
            name += "_large"
            p["n"] = [25000]
    elif experiment in SWEEPS:
        name += "_" + experiment
        p[experiment] = list(SWEEPS[experiment])
        if large and experiment == "n":
            name += "_large"
            p["n"] = SWEEPS["n"] + [i * 8000 for i in range(2, 26)]
        if large and experiment == "d":
            name += "_large"
            p["n"] = [25000]
            p["d"] = [i * 5 for i in range(1, 21)]
    else:
        raise ValueError(f"unknown experiment {experiment!r}")
    if experiment_2nd:
        if experiment_2nd not in SWEEPS_2ND:
            raise ValueError(f"unknown experiment {experiment_2nd!r}")
        name += "_" + experiment_2nd
        p[experiment_2nd] = list(SWEEPS_2ND[experiment_2nd])
    return name, p


def run_file(method_name, n, d, c, N_size, F, r, num_obj, min_size, order, cls):
    f = (f"{method_name}n{n}d{d}c{c}N_size{N_size}F{F}r{r}num_obj{num_obj}min_size{min_size}order{order}")
    if cls != 4:
        f += f"cl{cls}"
    return f + ".npz"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("method", choices=list(METHODS))
    ap.add_argument("experiment")
    ap.add_argument("repeats", type=int)
    ap.add_argument("real_no_clusters", type=int)
    ap.add_argument("experiment_2nd", nargs="?", default=None)
    ap.add_argument("large", nargs="?", type=int, default=0)
    ap.add_argument("--out", default="experiments_data")
    a = ap.parse_args(argv)
    method_name, mode = METHODS[a.method]
    suffix, p = grid(a.experiment, a.real_no_clusters, a.experiment_2nd, bool(a.large))
    os.makedirs(os.path.join(a.out, "runs"), exist_ok=True)
    if mode == _lib.MODE_LEARNED:
        from . import pruner
        pruner.use()
    np.savez(os.path.join(a.out, method_name + suffix + ".npz"),
             **{k: np.asarray(v) for k, v in p.items()}, repeats=a.repeats, real_no_clusters=a.real_no_clusters)
    done = []
    for n in p["n"]:
        for d in p["d"]:
            for c in p["c"]:
                for N_size in p["N_size"]:
                    for F in p["F"]:
                        for r in p["r"]:
                            for num_obj in p["num_obj"]:
                                for min_size in p["min_size"]:
                                    for order in p["order"]:
                                        for cls in p["real_cls"]:
                                            path = os.path.join(a.out, "runs", run_file(
                                                method_name, n, d, c, N_size, F, r, num_obj, min_size, order, cls))
                                            tag = (f"{a.method} n {n} d {d} c {c} N_size {N_size} F {F} r {r} "
                                                   f"num_obj {num_obj} min_size {min_size} order {order}")
                                            if os.path.exists(path):
                                                print("Experiment already performed!", tag, flush=True)
                                                continue
                                            if d > 128:
                                                print("Skipped (the engine runs d <= 128):", tag, flush=True)
                                                continue
                                            if order != 0 and mode == _lib.MODE_LEARNED:
                                                print("Skipped (entropy_order != 0 is not run by the learned "
                                                      "pruner):", tag, flush=True)
                                                continue
                                            print("Running experiment...", tag, flush=True)
                                            rec = {"times": [], "no_clusters": []}
                                            for i in range(a.repeats):
                                                X = INSCY.load_synt(d, n, cls, i)
                                                t0 = time.time()
                                                res = INSCY.run(X, N_size, F, num_obj, int(n * min_size), r, c,
                                                                True, order, mode=mode)
                                                t = time.time() - t0
                                                rec["times"].append(t)
                                                no = INSCY.count_number_of_clusters(res.subspaces, res.labels)
                                                rec["no_clusters"].append(no)
                                                masks = np.zeros((len(res.subspaces), d), dtype=np.int8)
                                                for j, s in enumerate(res.subspaces):
                                                    masks[j, s] = 1
                                                rec[f"subspaces_{i}"] = masks
                                                rec[f"clusterings_{i}"] = np.array(res.labels, dtype=np.int32)
                                                print(f"Finished {method_name}{suffix}, took: {t:.4f}s", i + 1, "/",
                                                      a.repeats, "number of clusters:", no, flush=True)
                                            np.savez(path, **{k: np.asarray(v) for k, v in rec.items()})
                                            done.append(path)
    return done


if __name__ == "__main__":
    main()
