"""Seeded synthetic subspace-cluster generator and the reference's `normalize`.

The reference generates its synthetic sets with a prebuilt binary
(python/inscy.py:235-250 shells out to data/cluster_generator); this module is
a from-scratch numpy (PCG64) generator for the workloads the engine is
measured on, so every figure is reproducible from a seed.

Each cluster lives in a random axis-parallel subspace: Gaussian around a centre
in its own dimensions, uniform in the others.  Noise points are uniform in all
dimensions.  Values are clipped to [0, 1] and min-max normalised per column
exactly as python/inscy.py:37-42 (`(x - min) / (max - min)` in float32).
"""

from __future__ import annotations

import dataclasses

import numpy as np


def normalize(x: np.ndarray) -> np.ndarray:
    """Per-column min-max scaling in float32 (python/inscy.py:37-42)."""
    x = np.asarray(x, dtype=np.float32)
    lo = x.min(axis=0, keepdims=True)
    hi = x.max(axis=0, keepdims=True)
    return ((x - lo) / (hi - lo)).astype(np.float32)


@dataclasses.dataclass
class Workload:
    """One INSCY run: the data and the reference call's parameters."""

    name: str
    X: np.ndarray            # float32 [n, d], row-major, values in [0, 1]
    truth: np.ndarray        # int32 [n], generating cluster id, -1 for noise
    subspaces: list          # generating subspace (sorted dims) per cluster
    eps: float
    cells: int
    F: float
    num_obj: int
    min_size: int
    r: float = 1.0
    rectangular: bool = True
    entropy_order: int = 0

    @property
    def n(self) -> int:
        return int(self.X.shape[0])

    @property
    def d(self) -> int:
        return int(self.X.shape[1])

    def params(self) -> dict:
        return dict(neighborhood_size=self.eps, F=self.F, num_obj=self.num_obj, min_size=self.min_size,
                    r=self.r, number_of_cells=self.cells, rectangular=self.rectangular,
                    entropy_order=self.entropy_order)


def make_clusters(rng: np.random.Generator, n: int, d: int, n_clusters: int, sub_dims, sigma,
                  noise_frac: float, sizes: str = "equal", zipf_a: float = 1.2):
    """Draws an [n, d] float32 set; returns (X_normalised, truth, subspaces).

    sub_dims: int, or (lo, hi) inclusive range drawn per cluster.
    sigma:    float, or (lo, hi) range drawn per cluster.
    sizes:    "equal" splits the cluster points evenly (remainder to the first
              clusters); "zipf" draws weights k^-zipf_a over a random ranking.
    """
    n_noise = int(round(noise_frac * n))
    n_clu = n - n_noise
    if sizes == "equal":
        base, extra = divmod(n_clu, n_clusters)
        counts = np.array([base + (1 if i < extra else 0) for i in range(n_clusters)], dtype=np.int64)
    elif sizes == "zipf":
        w = np.arange(1, n_clusters + 1, dtype=np.float64) ** (-zipf_a)
        w = w[rng.permutation(n_clusters)]
        counts = np.floor(w / w.sum() * n_clu).astype(np.int64)
        counts[: n_clu - counts.sum()] += 1
    else:
        raise ValueError(f"unknown sizes mode {sizes!r}")

    X = rng.random((n, d), dtype=np.float64)
    truth = np.full(n, -1, dtype=np.int32)
    subspaces = []
    at = 0
    for k in range(n_clusters):
        m = int(counts[k])
        if isinstance(sub_dims, tuple):
            kd = int(rng.integers(sub_dims[0], sub_dims[1] + 1))
        else:
            kd = int(sub_dims)
        kd = min(kd, d)
        dims = np.sort(rng.choice(d, size=kd, replace=False))
        s = float(rng.uniform(sigma[0], sigma[1])) if isinstance(sigma, tuple) else float(sigma)
        centre = rng.uniform(0.2, 0.8, size=kd)
        X[at:at + m, dims] = centre + s * rng.standard_normal((m, kd))
        truth[at:at + m] = k
        subspaces.append([int(x) for x in dims])
        at += m
    perm = rng.permutation(n)
    X = np.clip(X[perm], 0.0, 1.0).astype(np.float32)
    truth = truth[perm]
    return normalize(X), truth, subspaces


def config(idx: int, repeat: int = 0) -> Workload:
    """The measured workloads, seeded with 1000*idx + repeat.

    0: n=1500 d=5, 4 clusters sigma=0.05 in 3 dims, 10% noise; eps=0.01 c=4 mu=2 min_size=75
    1: n=25000 d=15, 10 clusters sigma=0.03 in 4-8 dims, 10% noise; eps=0.015 c=6 mu=4 min_size=500
    2: n=100000 d=20, 20 Zipf(1.2) clusters sigma~U[0.02,0.05] in 4-10 dims, 15% noise;
       eps=0.02 c=8 mu=4 min_size=1000
    4: n=200000 d=32, 30 clusters sigma=0.03 in 4-12 dims, 15% noise; eps=0.02 c=8 mu=4
       min_size=2000 (the learned-prune evaluation set)
    """
    rng = np.random.default_rng(1000 * idx + repeat)
    if idx == 0:
        X, t, s = make_clusters(rng, 1500, 5, 4, 3, 0.05, 0.10)
        return Workload("n1500_d5_4cl", X, t, s, eps=0.01, cells=4, F=1.0, num_obj=2, min_size=75)
    if idx == 1:
        X, t, s = make_clusters(rng, 25000, 15, 10, (4, 8), 0.03, 0.10)
        return Workload("n25000_d15_10cl", X, t, s, eps=0.015, cells=6, F=1.0, num_obj=4, min_size=500)
    if idx == 2:
        X, t, s = make_clusters(rng, 100000, 20, 20, (4, 10), (0.02, 0.05), 0.15, sizes="zipf")
        return Workload("n100000_d20_20cl", X, t, s, eps=0.02, cells=8, F=1.0, num_obj=4, min_size=1000)
    if idx == 4:
        X, t, s = make_clusters(rng, 200000, 32, 30, (4, 12), 0.03, 0.15)
        return Workload("n200000_d32_30cl", X, t, s, eps=0.02, cells=8, F=1.0, num_obj=4, min_size=2000)
    raise ValueError(f"no single-dataset workload {idx}")


def config2_small() -> Workload:
    """Config 2's generator and parameters (20 Zipf(1.2) clusters in 4-10 of
    20 dims, sigma~U[0.02,0.05], 15% noise; eps=0.02 c=8 mu=4) at n=25,000,
    min_size=0.01 n, seed 2500: small enough for the single-threaded
    pointer-tree oracle."""
    rng = np.random.default_rng(2500)
    X, t, s = make_clusters(rng, 25000, 20, 20, (4, 10), (0.02, 0.05), 0.15, sizes="zipf")
    return Workload("n25000_d20_20cl_cfg2params", X, t, s, eps=0.02, cells=8, F=1.0, num_obj=4, min_size=250)


def wide(n: int, d: int, eps: float = 0.01, cells: int = 4, num_obj: int = 8, sigma: float = 0.005,
         rectangular: bool = True, seed: int = 11) -> Workload:
    """n points in d dimensions (up to 128) with 3 clusters in 2-4 dimensional
    subspaces drawn from all d dimensions (5% noise), min_size = 0.05 n:
    dimension ids beyond 64, as in the reference's large-d sweep
    (experiments/run.py:62-66: d up to 100, eps 0.01, mu 8), at sizes the
    oracle finishes."""
    rng = np.random.default_rng(seed)
    X, t, s = make_clusters(rng, n, d, 3, (2, 4), sigma, 0.05)
    name = f"wide_n{n}_d{d}_c{cells}{'' if rectangular else '_phi'}"
    return Workload(name, X, t, s, eps=eps, cells=cells, F=1.0, num_obj=num_obj, min_size=int(0.05 * n),
                    rectangular=rectangular)


def multi_instance(seed: int = 3000, count: int = 64):
    """64 independent datasets: n log-uniform in [2k, 50k], d in 8..24, 3-12
    clusters (subspace dim 3-8, sigma 0.03, 10% noise), eps=0.015 c=6 mu=4,
    min_size = 0.02 n."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(50000))))
        d = int(rng.integers(8, 25))
        k = int(rng.integers(3, 13))
        X, t, s = make_clusters(rng, n, d, k, (3, min(8, d)), 0.03, 0.10)
        out.append(Workload(f"multi{i}_n{n}_d{d}", X, t, s, eps=0.015, cells=6, F=1.0, num_obj=4,
                            min_size=int(0.02 * n)))
    return out


# The reference's real-data workloads (experiments/run_real.py:15-66): data set
# -> (arrays in the inputs file, columns kept after normalisation).  Parameters
# are run_real.py's: eps=0.01, c=4, F=1, num_obj=8, r=1, min_size=0.05 n,
# rectangular, identity dimension order.
REAL = {
    "glass": (["glass"], slice(1, -1)),
    "vowel": (["vowel"], slice(None)),
    "pendigits": (["pendigits"], slice(None, -1)),
    "sky_0.5x0.5": (["sky_tile0"], slice(None)),
    "sky_1x1": ([f"sky_tile{i}" for i in range(4)], slice(None)),
    "sky_1x2": ([f"sky_tile{i}" for i in range(8)], slice(None)),
}


def real(name: str, inputs_path: str) -> Workload:
    """One of the reference's real-data workloads from the committed inputs
    file (tests/golden/real/real_inputs.npz): concatenated, normalised per
    column as python/inscy.py:37-42, columns selected as run_real.py."""
    keys, cols = REAL[name]
    with np.load(inputs_path, allow_pickle=False) as z:
        raw = np.concatenate([z[k] for k in keys], axis=0).astype(np.float32)
    X = np.ascontiguousarray(normalize(raw)[:, cols], dtype=np.float32)
    n = X.shape[0]
    return Workload(f"real_{name}", X, np.full(n, -1, dtype=np.int32), [], eps=0.01, cells=4, F=1.0, num_obj=8,
                    min_size=int(n * 0.05))


def small(seed: int, n: int = 200, d: int = 4, n_clusters: int = 2, eps: float = 0.05, cells: int = 3,
          num_obj: int = 3, min_size: int = 10, rectangular: bool = True, noise_frac: float = 0.2,
          sigma: float = 0.05, sub_dims=(2, 3)) -> Workload:
    """Small seeded case for exhaustive oracle/engine cross-checks."""
    rng = np.random.default_rng(seed)
    sub = (min(sub_dims[0], d), min(sub_dims[1], d)) if isinstance(sub_dims, tuple) else min(sub_dims, d)
    X, t, s = make_clusters(rng, n, d, n_clusters, sub, sigma, noise_frac)
    return Workload(f"small{seed}_n{n}_d{d}", X, t, s, eps=eps, cells=cells, F=1.0, num_obj=num_obj,
                    min_size=min_size, rectangular=rectangular)
