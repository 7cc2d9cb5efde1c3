"""Writes the golden fixtures of the measured workloads from the oracle.

    python tests/golden/make_golden.py 0      # config 0: full canonical labels (seconds)
    python tests/golden/make_golden.py 1      # config 1: per-subspace label digests (~15 min, 1 core)
    python tests/golden/make_golden.py 0 phi  # the same with rectangular=False (phi density)
    python tests/golden/make_golden.py 1 phi

The oracle is the pointer-tree restatement of the reference's CPU INSCY
(oracle/inscy_ref.cpp); these files let GPU tests check the full-size
workloads without re-running it.
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpu_inscy_amd import datagen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(labels: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()


def main(cfg: int, phi: bool = False) -> None:
    w = datagen.config(cfg)
    tag = ""
    if phi:
        import dataclasses
        w = dataclasses.replace(w, rectangular=False)
        tag = "_phi"
    ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order,
                    trace=True)
    canon = [O.canonical(l) for l in ref.labels]
    meta = dict(workload=w.name, n=w.n, d=w.d, eps=w.eps, cells=w.cells, F=w.F, num_obj=w.num_obj,
                min_size=w.min_size, r=w.r, rectangular=w.rectangular, calls=ref.calls, trees=ref.trees,
                X_sha256=hashlib.sha256(w.X.tobytes()).hexdigest())
    if cfg == 0:
        np.savez_compressed(os.path.join(HERE, f"config0{tag}.npz"), subspace_masks=np.array(
            [sum(1 << k for k in s) for s in ref.subspaces], dtype=np.int64), labels=np.stack(canon),
            trace=ref.trace)
        with open(os.path.join(HERE, f"config0{tag}.json"), "w") as f:
            json.dump(meta, f, indent=1)
    else:
        meta["subspaces"] = [[s, int(c.max() + 1), int((c >= 0).sum()), digest(c)] for s, c in zip(ref.subspaces, canon)]
        with open(os.path.join(HERE, f"config{cfg}{tag}.json"), "w") as f:
            json.dump(meta, f)


if __name__ == "__main__":
    main(int(sys.argv[1]), len(sys.argv) > 2 and sys.argv[2] == "phi")
