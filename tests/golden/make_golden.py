"""Writes the golden fixtures of the measured workloads from the oracle.

    python tests/golden/make_golden.py 0          # config 0: full canonical labels (seconds)
    python tests/golden/make_golden.py 1          # config 1: per-subspace label digests (~15 min, 1 core)
    python tests/golden/make_golden.py 0 phi      # the same with rectangular=False (phi density)
    python tests/golden/make_golden.py 1 phi
    python tests/golden/make_golden.py 2 --threads 6   # config 2 (n=100k, d=20), set restatement
    python tests/golden/make_golden.py 3 --threads 6   # config 3: the 64 data sets, set restatement
    python tests/golden/make_golden.py 2small     # config 2's parameters at n=25,000, pointer-tree oracle
    python tests/golden/make_golden.py 1ent --threads 4   # config 1, ascending entropy level order
    python tests/golden/make_golden.py 1ent- --threads 4  # config 1, descending entropy level order
    python tests/golden/make_golden.py check1 --threads 8  # set restatement on config 1 vs config1.json
    python tests/golden/make_golden.py wide --threads 4    # d = 100 and 127 (dimension ids beyond 64)
    python tests/golden/make_golden.py wide_ent --threads 2  # d = 127 in both entropy level orders

The oracle is the reference's CPU INSCY restated (oracle/inscy_ref.cpp): the
pointer-tree restatement for configs 0, 1 and 2small, and the set restatement
(the same algorithm over point sets, multithreaded, identical results on
config 1 and the fuzz cases of tests/test_oracle.py) for configs 2 and 3,
which the pointer-tree walk cannot finish in hours.  These files let GPU tests
check the full-size workloads without re-running it.
"""

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpu_inscy_amd import datagen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(labels: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()


def canonical(labels: np.ndarray) -> np.ndarray:
    """O.canonical, vectorised: negatives -> -1, clusters 0.. by smallest member."""
    lab = np.asarray(labels)
    out = np.full(lab.shape, -1, dtype=np.int32)
    pos = np.flatnonzero(lab >= 0)
    if pos.size == 0:
        return out
    vals = lab[pos]
    ids, first = np.unique(vals, return_index=True)  # first: index into pos of each id's smallest member
    rank = np.empty(len(ids), dtype=np.int32)
    rank[np.argsort(pos[first], kind="stable")] = np.arange(len(ids), dtype=np.int32)
    out[pos] = rank[np.searchsorted(ids, vals)]
    return out


def rows(run) -> list:
    """[subspace, clusters, clustered points, sha256 of the canonical labels] per result subspace."""
    out = []
    for s, lab in zip(run.subspaces, run.labels):
        c = canonical(lab)
        out.append([s, int(c.max() + 1), int((c >= 0).sum()), digest(c)])
    return out


def meta_of(w, run) -> dict:
    return dict(workload=w.name, n=w.n, d=w.d, eps=w.eps, cells=w.cells, F=w.F, num_obj=w.num_obj,
                min_size=w.min_size, r=w.r, rectangular=w.rectangular, calls=run.calls, trees=run.trees,
                X_sha256=hashlib.sha256(w.X.tobytes()).hexdigest())


def run_workload(w, engine: str, threads: int):
    if engine == "sets":
        return O.run_sets(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, threads=threads)
    return O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order)


def write_json(name: str, obj) -> None:
    tmp = os.path.join(HERE, name + ".tmp")
    with open(tmp, "w") as f:
        json.dump(obj, f)
    os.replace(tmp, os.path.join(HERE, name))


WIDE_CASES = [("ref", datagen.wide(300, 100, cells=4)), ("sets", datagen.wide(600, 127, cells=2)),
              ("sets", datagen.wide(2000, 100, cells=2))]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("config")
    ap.add_argument("mode", nargs="?", default="")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    phi = a.mode == "phi"
    tag = "_phi" if phi else ""
    t0 = time.time()
    if a.config in ("0", "1"):
        cfg = int(a.config)
        w = datagen.config(cfg)
        if phi:
            import dataclasses
            w = dataclasses.replace(w, rectangular=False)
        ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order,
                        trace=True)
        meta = meta_of(w, ref)
        if cfg == 0:
            canon = [canonical(l) for l in ref.labels]
            np.savez_compressed(os.path.join(HERE, f"config0{tag}.npz"), subspace_masks=np.array(
                [sum(1 << k for k in s) for s in ref.subspaces], dtype=np.int64), labels=np.stack(canon),
                trace=ref.trace)
            with open(os.path.join(HERE, f"config0{tag}.json"), "w") as f:
                json.dump(meta, f, indent=1)
        else:
            meta["subspaces"] = rows(ref)
            write_json(f"config1{tag}.json", meta)
    elif a.config in ("1ent", "1ent-"):
        # config 1 with the SCY-tree levels in ascending (1ent) or descending
        # (1ent-) entropy order, from the set restatement of a permuted level
        # order (tests/test_oracle.py pins it against the pointer tree)
        order = 1 if a.config == "1ent" else -1
        import dataclasses
        w = dataclasses.replace(datagen.config(1), entropy_order=order)
        run = O.run_sets_order(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order, threads=a.threads)
        meta = meta_of(w, run)
        meta["entropy_order"] = order
        meta["oracle"] = "set restatement, permuted level order"
        meta["subspaces"] = rows(run)
        write_json(f"config1_entropy{'' if order == 1 else '_desc'}.json", meta)
    elif a.config == "check1":
        w = datagen.config(1)
        run = run_workload(w, "sets", a.threads)
        with open(os.path.join(HERE, "config1.json")) as f:
            g = json.load(f)
        got = rows(run)
        print("set restatement == config1.json:", got == g["subspaces"], run.trees == g["trees"], flush=True)
    elif a.config == "2":
        w = datagen.config(2)
        run = run_workload(w, "sets", a.threads)
        meta = meta_of(w, run)
        meta["oracle"] = "set restatement"
        meta["subspaces"] = rows(run)
        write_json("config2.json", meta)
    elif a.config == "2small":
        w = datagen.config2_small()
        run = run_workload(w, "ref", 1)
        meta = meta_of(w, run)
        meta["oracle"] = "pointer-tree restatement"
        meta["generator"] = "datagen.config2_small()"
        meta["subspaces"] = rows(run)
        write_json("config2_small.json", meta)
    elif a.config == "3":
        sets = datagen.multi_instance()
        path = os.path.join(HERE, "config3.json")
        done = {}
        if os.path.exists(path):
            with open(path) as f:
                done = {m["workload"]: m for m in json.load(f)["sets"]}
        out = []
        for i, w in enumerate(sets):
            if w.name in done:
                out.append(done[w.name])
                continue
            t1 = time.time()
            run = run_workload(w, "sets", a.threads)
            m = meta_of(w, run)
            m["subspaces"] = rows(run)
            out.append(m)
            write_json("config3.json", dict(oracle="set restatement", sets=out, complete=len(out) == len(sets)))
            print(f"set {i} {w.name}: {len(run.subspaces)} subspaces, {run.trees} trees, "
                  f"{time.time() - t1:.1f} s", flush=True)
        write_json("config3.json", dict(oracle="set restatement", sets=out, complete=True))
    elif a.config == "wide":
        # dimension sets beyond 64 bits: the pointer tree at the large-d
        # sweep's parameters (c=4, eps=0.01, mu=8) and the set restatement at
        # c=2 (its prefix keys hold 2^127 cells)
        out = []
        for engine, w in WIDE_CASES:
            t1 = time.time()
            run = run_workload(w, engine, a.threads)
            m = meta_of(w, run)
            m["generator"] = f"datagen.wide({w.n}, {w.d}, cells={w.cells}, rectangular={w.rectangular})"
            m["oracle"] = "pointer-tree restatement" if engine == "ref" else "set restatement"
            m["subspaces"] = rows(run)
            out.append(m)
            print(f"{w.name}: {len(run.subspaces)} subspaces, {run.trees} trees, {time.time() - t1:.1f} s", flush=True)
        write_json("wide.json", dict(sets=out))
    elif a.config == "wide_ent":
        # d = 127 in ascending and descending entropy level order (set
        # restatement of a permuted level order)
        import dataclasses
        out = []
        for order in (1, -1):
            w = dataclasses.replace(datagen.wide(600, 127, cells=2), entropy_order=order)
            run = O.run_sets_order(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order,
                                   threads=a.threads)
            m = meta_of(w, run)
            m["generator"] = "datagen.wide(600, 127, cells=2)"
            m["entropy_order"] = order
            m["oracle"] = "set restatement, permuted level order"
            m["subspaces"] = rows(run)
            out.append(m)
        write_json("wide_entropy.json", dict(sets=out))
    else:
        raise SystemExit(f"unknown config {a.config!r}")
    print(f"done in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
