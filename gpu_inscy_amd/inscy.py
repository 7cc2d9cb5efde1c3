"""Drop-in replacement of the reference's Python surface (python/inscy.py).

Same function names, argument order and defaults, same `(subspaces,
clusterings)` return: lists of ascending dimension lists and, per subspace, a
list of n labels.  Every entry point runs on the HIP engine
(libinscy_hip.so); there is no CPU fallback.

Results equal the reference's CPU INSCY (INSCY.CPU -> run_cpu3_weak) with
labels in canonical form: -1 for noise, clusters numbered 0.. by smallest
point id (the reference's raw labels are a renaming of these; its -2 border
marks are noise).

Mapping of the reference's entry points:
  CPU (python/inscy.py:121-125)          -> exact mode (same results, on the GPU)
  GPU (:189-192, run_gpu_4)              -> exact mode, one restricted tree at a time
  GPU5 (:194-197, run_gpu_5)             -> batched mode (same results)
  GPU3, GPU_star, GPU_memory, GPU_memory2 (:183-214) -> batched mode
  run_cpu, run_cpu_weak, run_gpu, run_gpu_* (7-argument ops, :110-186) -> exact
      mode with rectangular=False, entropy_order=0, the defaults those ops had
  run_cmp (:133-134, inscy_map.cpp:273-320)  -> exact mode, tensor output
  GPU_learned (no counterpart)            -> the learned pruner (approximate)
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .datagen import normalize as _normalize_np

__all__ = [
    "normalize", "clean_up", "count_number_of_clusters", "run", "CPU", "GPU", "GPU5", "GPU3", "GPU_star",
    "GPU_memory", "GPU_memory2", "run_cpu", "run_cpu_weak", "run_gpu", "run_gpu_weak", "run_gpu_reduced",
    "run_gpu_multi", "run_gpu_multi2", "run_gpu_multi2_cl_all", "run_gpu_multi2_cl_re_all",
    "run_gpu_multi2_cl_multi", "run_gpu_multi2_cl_multi_mem", "run_cpu_gpu_mix", "run_cpu_gpu_mix_cl_steam",
    "run_gpu_stream", "run_cmp", "GPU_learned", "load_synt",
]


def normalize(x):
    """Per-column min-max scaling (python/inscy.py:37-42); torch or numpy in, same type out."""
    try:
        import torch
        if isinstance(x, torch.Tensor):
            mn = x.min(0, keepdim=True)[0]
            mx = x.max(0, keepdim=True)[0]
            return (x - mn) / (mx - mn)
    except ImportError:
        pass
    return _normalize_np(x)


def run(X, neighborhood_size, F, num_obj, min_size, r=1.0, number_of_cells=3, rectangular=False, entropy_order=0,
        mode=_lib.MODE_EXACT) -> _lib.Result:
    """Runs INSCY and returns a Result (numpy labels [#subspaces, n], stats).

    X: float32 [n, d] numpy array, CPU torch tensor, or CUDA torch tensor (then
    the engine reads it in place on the tensor's device and current stream).
    """
    params = _lib.make_params(neighborhood_size, F, num_obj, min_size, r, number_of_cells, rectangular,
                              entropy_order)
    try:
        import torch
    except ImportError:
        torch = None
    if torch is not None and isinstance(X, torch.Tensor):
        if X.dim() != 2:
            raise ValueError("X must be [n, d]")
        if X.is_cuda:
            Xc = X.detach().to(torch.float32).contiguous()
            with torch.cuda.device(Xc.device):
                stream = torch.cuda.current_stream(Xc.device).cuda_stream
                return _lib.run_device(Xc.data_ptr(), Xc.shape[0], Xc.shape[1], params, mode, stream)
        X = X.detach().to(torch.float32).contiguous().numpy()
    return _lib.run_host(np.asarray(X, dtype=np.float32), params, mode)


def run_many(datasets, mode=_lib.MODE_BATCHED):
    """Runs INSCY on many independent data sets at once; returns a list of
    Results in input order.

    datasets: iterable of (X, neighborhood_size, F, num_obj, min_size, r,
    number_of_cells, rectangular, entropy_order) tuples.  All X must be CUDA
    tensors on one device, or all host arrays.
    """
    import torch
    xs, shapes, params = [], [], []
    for (X, eps, F, num_obj, min_size, r, cells, rect, order) in datasets:
        if isinstance(X, torch.Tensor):
            X = X.detach().to(torch.float32).contiguous()
        else:
            X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2:
            raise ValueError("every X must be [n, d]")
        xs.append(X)
        shapes.append(tuple(X.shape))
        params.append(_lib.make_params(eps, F, num_obj, min_size, r, cells, rect, order))
    on_device = [isinstance(X, torch.Tensor) and X.is_cuda for X in xs]
    if any(on_device) and not all(on_device):
        raise ValueError("mix of device and host inputs")
    if all(on_device):
        with torch.cuda.device(xs[0].device):
            torch.cuda.current_stream().synchronize()
            return _lib.run_many([X.data_ptr() for X in xs], shapes, params, mode, device=True)
    hosts = [X.numpy() if isinstance(X, torch.Tensor) else X for X in xs]
    return _lib.run_many([h.ctypes.data for h in hosts], shapes, params, mode, device=False)


def _lists(res: _lib.Result):
    return [list(s) for s in res.subspaces], [row.tolist() for row in res.labels]


def CPU(X, neighborhood_size, F, num_obj, min_size, r=1., number_of_cells=3, rectangular=False, entropy_order=0):
    return _lists(run(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells, rectangular, entropy_order,
                      _lib.MODE_EXACT))


def GPU(X, neighborhood_size, F, num_obj, min_size, r=1., number_of_cells=3, rectangular=False, entropy_order=0):
    return _lists(run(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells, rectangular, entropy_order,
                      _lib.MODE_EXACT))


def GPU5(X, neighborhood_size, F, num_obj, min_size, r=1., number_of_cells=3, rectangular=False, entropy_order=0):
    return _lists(run(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells, rectangular, entropy_order,
                      _lib.MODE_BATCHED))


GPU3 = GPU5
GPU_star = GPU5
GPU_memory = GPU5
GPU_memory2 = GPU5


def _seven(X, neighborhood_size, F, num_obj, min_size, r=1., number_of_cells=3):
    return CPU(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells, False, 0)


run_cpu = _seven
run_cpu_weak = _seven
run_gpu = _seven
run_gpu_weak = _seven
run_gpu_reduced = _seven
run_gpu_multi = _seven
run_gpu_multi2 = _seven
run_gpu_multi2_cl_all = _seven
run_gpu_multi2_cl_re_all = _seven
run_gpu_multi2_cl_multi = _seven
run_gpu_multi2_cl_multi_mem = _seven
run_cpu_gpu_mix = _seven
run_cpu_gpu_mix_cl_steam = _seven


def GPU_learned(X, neighborhood_size, F, num_obj, min_size, r=1., number_of_cells=3, rectangular=False,
                entropy_order=0):
    """Level-batched INSCY with the learned pruner (approximate; the shipped
    network gpu_inscy_amd/pruner_weights.npz unless another is loaded with
    gpu_inscy_amd.pruner.use)."""
    from . import pruner
    if not getattr(GPU_learned, "_loaded", False):
        pruner.use()
        GPU_learned._loaded = True
    return _lists(run(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells, rectangular, entropy_order,
                      _lib.MODE_LEARNED))


def run_cmp(X, neighborhood_size, F, num_obj, min_size, number_of_cells=3):
    """python/inscy.py:133-134 -> inscy_map.cpp:273-320: the result as two
    int32 tensors, subspace masks [#subspaces, d] (1 at the subspace's
    dimensions) and labels [#subspaces, n].  The reference computes it with
    INSCYCompare, a development driver that checks its GPU trees against its
    CPU trees on the way and passes r = 0.5 to its clustering; here it is the
    exact engine with r = 0.5 and the 7-argument ops' rectangular=False."""
    import torch
    res = run(X, neighborhood_size, F, num_obj, min_size, 0.5, number_of_cells, False, 0, _lib.MODE_EXACT)
    d = int(X.shape[1])
    masks = torch.zeros((len(res.subspaces), d), dtype=torch.int32)
    for i, s in enumerate(res.subspaces):
        masks[i, s] = 1
    return [masks, torch.from_numpy(np.array(res.labels, dtype=np.int32))]


def run_gpu_stream(X, neighborhood_size, F, num_obj, min_size, number_of_cells=3):
    """python/inscy.py:221-225 (no r argument)."""
    return _seven(X, neighborhood_size, F, num_obj, min_size, 1.0, number_of_cells)


def clean_up(subspaces, clusterings, min_size):
    """python/inscy.py:51-85: drop clusters smaller than min_size and the
    subspaces left without clusters."""
    subspaces = np.array(subspaces, dtype=object)
    clusterings = np.array(clusterings)
    keep = np.zeros(len(clusterings), dtype=bool)
    for i in range(len(clusterings)):
        row = clusterings[i]
        ids, sizes = np.unique(row[row >= 0], return_counts=True)
        for c, s in zip(ids, sizes):
            if s < min_size:
                row[row == c] = -1
        keep[i] = bool(np.any(sizes >= min_size))
    return subspaces[keep], clusterings[keep]


def count_number_of_clusters(subspaces, clusterings):
    """python/inscy.py:88-107: number of distinct non-negative labels summed over subspaces."""
    total = 0
    for row in clusterings:
        row = np.asarray(row)
        total += int(np.unique(row[row >= 0]).size)
    return total


def load_synt(d, n, cl, re):
    """Seeded synthetic set of the reference's shape parameters
    (python/inscy.py:235-250 calls a prebuilt generator; this draws cl clusters
    in 0.8*d-dimensional subspaces with the project's own generator, seed re)."""
    import torch

    from .datagen import make_clusters
    rng = np.random.default_rng(int(re))
    X, _, _ = make_clusters(rng, int(n), int(d), int(cl), max(1, int(0.8 * d)), 0.05, 0.01)
    return torch.from_numpy(X)
