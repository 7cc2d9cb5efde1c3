"""Real-data inputs of the reference's published comparison, and their golden
results from the oracle.

    python tests/golden/real/make_real.py inputs            # writes real_inputs.npz
    python tests/golden/real/make_real.py golden NAME...     # adds NAME to real_golden.json

The reference times INSCY on four public data sets (experiments/run_real.py:15-34,
published in experiments/plot_real.py:13-14): glass (UCI), vowel, pendigits
(UCI) and tiles of an SDSS SkyServer photometry query.  `inputs` parses the
files under /root/reference/data/real the way run_real.py does (np.loadtxt, then
float32) and stores the raw float32 arrays; gpu_inscy_amd.datagen.real() applies
run_real.py's normalisation and column selection.  Only this script reads the
reference tree; everything else reads the .npz.

Sky workloads: sky_0.5x0.5 is the tile ra 0-0.5, dec 0-0.5; sky_1x1 the four
tiles of ra 0-1, dec 0-1 in run_real.py's order; sky_1x2 the eight tiles of
ra 0-2, dec 0-1.

`golden` runs the oracle's set restatement (all host cores) with run_real.py's
parameters and records per-subspace label digests.
"""

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, ROOT)

SRC = "/root/reference/data/real"
SKY_TILES = ["(0_0.5_0_0.5)", "(0_0.5_0.5_1)", "(0.5_1_0_0.5)", "(0.5_1_0.5_1)",
             "(1_1.5_0_0.5)", "(1_1.5_0.5_1)", "(1.5_2_0_0.5)", "(1.5_2_0.5_1)"]


def make_inputs():
    arrs = {
        "glass": np.loadtxt(os.path.join(SRC, "glass.data"), delimiter=",", skiprows=0),
        "vowel": np.loadtxt(os.path.join(SRC, "vowel.dat"), delimiter=",", skiprows=0),
        "pendigits": np.loadtxt(os.path.join(SRC, "pendigits.tra"), delimiter=",", skiprows=0),
    }
    for i, t in enumerate(SKY_TILES):
        arrs[f"sky_tile{i}"] = np.loadtxt(os.path.join(SRC, "skyserver", f"result {t}.csv"), delimiter=",",
                                          skiprows=1)
    arrs = {k: v.astype(np.float32) for k, v in arrs.items()}
    np.savez_compressed(os.path.join(HERE, "real_inputs.npz"), **arrs)
    for k, v in arrs.items():
        print(k, v.shape)


def digest(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()


def make_golden(names):
    from gpu_inscy_amd import datagen
    from oracle import oracle as O
    path = os.path.join(HERE, "real_golden.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for name in names:
        w = datagen.real(name, os.path.join(HERE, "real_inputs.npz"))
        t = time.time()
        ref = O.run_sets(w.X, w.eps, w.F, w.num_obj, w.min_size, r=w.r, cells=w.cells, rectangular=w.rectangular)
        el = time.time() - t
        canon = [O.canonical(l) for l in ref.labels]
        out[name] = dict(n=w.n, d=w.d, calls=ref.calls, trees=ref.trees, oracle_seconds=round(el, 1),
                         X_sha256=hashlib.sha256(w.X.tobytes()).hexdigest(),
                         subspaces=[[s, int(c.max() + 1), int((c >= 0).sum()), digest(c)]
                                    for s, c in zip(ref.subspaces, canon)])
        print(name, w.n, w.d, "subspaces", len(ref.subspaces), "trees", ref.trees, f"{el:.1f} s", flush=True)
        json.dump(out, open(path, "w"))


if __name__ == "__main__":
    if sys.argv[1] == "inputs":
        make_inputs()
    else:
        make_golden(sys.argv[2:])
