"""ctypes binding of libinscy_hip.so (include/inscy.h).

The library is built in-tree by gpu_inscy_amd.build; loading fails with an
ImportError naming the missing file rather than falling back to anything else.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# INSCY_LIBRARY: another build of the same library (same C ABI) to load instead
LIB_PATH = os.environ.get("INSCY_LIBRARY") or os.path.join(HERE, "libinscy_hip.so")

INSCY_OK = 0
INSCY_ERR_ARGUMENT = 1
INSCY_ERR_DEVICE = 2
INSCY_ERR_MEMORY = 3
INSCY_ERR_UNSUPPORTED = 4

MODE_EXACT = 1
MODE_BATCHED = 2
MODE_LEARNED = 3
MODE_COLLECT = 4

MLP_INPUTS = 64
MLP_HIDDEN = 128
MLP_PARAMS = 128 * 64 + 128 + 128 * 128 + 128 + 128 + 1

EXPORTED = [
    "inscy_result_labels_all", "inscy_run_many", "inscy_run_many_device", "inscy_release_memory",
    "inscy_run", "inscy_run_device", "inscy_result_num_subspaces", "inscy_result_subspace",
    "inscy_result_labels", "inscy_result_num_points", "inscy_result_stats", "inscy_result_free",
    "inscy_last_error", "inscy_discretise", "inscy_neighbour_counts", "inscy_neighbour_counts_noalloc",
    "inscy_dense_components", "inscy_profile", "inscy_set_tuning",
    "inscy_set_pruner", "inscy_result_pairs", "inscy_mlp_forward", "inscy_mlp_gradient", "inscy_mlp_adam",
    "inscy_mlp_round",
]


vp_t = ctypes.c_void_p


class InscyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"INSCY error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("neighborhood_size", ctypes.c_float),
        ("F", ctypes.c_float),
        ("num_obj", ctypes.c_int32),
        ("min_size", ctypes.c_int32),
        ("r", ctypes.c_float),
        ("number_of_cells", ctypes.c_int32),
        ("rectangular", ctypes.c_int32),
        ("entropy_order", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("calls", ctypes.c_int64),
        ("trees", ctypes.c_int64),
        ("queries", ctypes.c_int64),
        ("candidates", ctypes.c_int64),
        ("subspaces", ctypes.c_int64),
        ("clusters", ctypes.c_int64),
        ("components", ctypes.c_int64),
        ("seconds", ctypes.c_double),
        ("count_kernel_ms", ctypes.c_double),
        ("count_launches", ctypes.c_int64),
        ("count_bytes", ctypes.c_int64),
        ("expand_seconds", ctypes.c_double),
        ("replay_seconds", ctypes.c_double),
        ("emit_seconds", ctypes.c_double),
        ("pairs_scored", ctypes.c_int64),
        ("pairs_skipped", ctypes.c_int64),
        ("sorts_kept", ctypes.c_int64),
        ("count_examined_bytes", ctypes.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gpu_inscy_amd.build` "
                          "(the engine has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.inscy_run.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Params), ctypes.c_int32,
                            ctypes.POINTER(vp)]
    L.inscy_run.restype = ctypes.c_int
    L.inscy_run_device.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Params), ctypes.c_int32, vp,
                                   ctypes.POINTER(vp)]
    L.inscy_run_device.restype = ctypes.c_int
    L.inscy_run_many.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(Params), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]
    L.inscy_run_many.restype = ctypes.c_int
    L.inscy_run_many_device.argtypes = L.inscy_run_many.argtypes
    L.inscy_run_many_device.restype = ctypes.c_int
    L.inscy_result_num_subspaces.argtypes = [vp]
    L.inscy_result_num_subspaces.restype = ctypes.c_int32
    L.inscy_result_subspace.argtypes = [vp, ctypes.c_int32, i32p]
    L.inscy_result_subspace.restype = ctypes.c_int32
    L.inscy_result_labels.argtypes = [vp, ctypes.c_int32]
    L.inscy_result_labels.restype = i32p
    L.inscy_result_labels_all.argtypes = [vp]
    L.inscy_result_labels_all.restype = i32p
    L.inscy_result_num_points.argtypes = [vp]
    L.inscy_result_num_points.restype = ctypes.c_int64
    L.inscy_result_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.inscy_result_stats.restype = ctypes.c_int
    L.inscy_result_free.argtypes = [vp]
    L.inscy_result_free.restype = None
    L.inscy_last_error.argtypes = []
    L.inscy_last_error.restype = ctypes.c_char_p
    L.inscy_discretise.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, vp, vp, vp]
    L.inscy_discretise.restype = ctypes.c_int
    L.inscy_neighbour_counts.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, i32p, ctypes.c_int32,
                                         vp, ctypes.c_int64, vp]
    L.inscy_neighbour_counts.restype = ctypes.c_int
    L.inscy_neighbour_counts_noalloc.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, i32p,
                                                 ctypes.c_int32, vp, ctypes.c_int64, vp, vp]
    L.inscy_neighbour_counts_noalloc.restype = ctypes.c_int
    L.inscy_dense_components.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Params), i32p,
                                         ctypes.c_int32, vp]
    L.inscy_dense_components.restype = ctypes.c_int
    L.inscy_release_memory.argtypes = []
    L.inscy_release_memory.restype = ctypes.c_int
    L.inscy_profile.argtypes = [ctypes.c_int]
    L.inscy_profile.restype = ctypes.c_int
    L.inscy_set_tuning.argtypes = [ctypes.c_int32, ctypes.c_int32]
    L.inscy_set_tuning.restype = ctypes.c_int32
    fp = ctypes.POINTER(ctypes.c_float)
    L.inscy_set_pruner.argtypes = [fp, ctypes.c_float, ctypes.c_int32]
    L.inscy_set_pruner.restype = ctypes.c_int
    L.inscy_result_pairs.argtypes = [vp, ctypes.POINTER(fp), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                     ctypes.POINTER(i32p), ctypes.POINTER(ctypes.POINTER(ctypes.c_int64))]
    L.inscy_result_pairs.restype = ctypes.c_int64
    L.inscy_mlp_forward.argtypes = [vp, vp, ctypes.c_int64, vp, vp]
    L.inscy_mlp_forward.restype = ctypes.c_int
    L.inscy_mlp_gradient.argtypes = [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_float, vp, ctypes.c_int32, vp, vp]
    L.inscy_mlp_gradient.restype = ctypes.c_int
    L.inscy_mlp_adam.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp]
    L.inscy_mlp_adam.restype = ctypes.c_int
    L.inscy_mlp_round.argtypes = [vp, vp, vp]
    L.inscy_mlp_round.restype = ctypes.c_int
    L.inscy_debug_trace.argtypes = [ctypes.c_int]
    L.inscy_debug_phi_exact.argtypes = [ctypes.c_int]
    L.inscy_debug_generic_walk.argtypes = [ctypes.c_int]
    L.inscy_debug_keep_budget.argtypes = [ctypes.c_int64]
    L.inscy_debug_thresholds.argtypes = [ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]
    L.inscy_debug_thresholds.restype = ctypes.c_int
    L.inscy_debug_scan_temp.argtypes = [ctypes.c_int64]
    L.inscy_debug_scan_temp.restype = ctypes.c_int64
    L.inscy_debug_exclusive_sum.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, vp]
    L.inscy_debug_exclusive_sum.restype = ctypes.c_int
    L.inscy_debug_sort_temp.argtypes = [ctypes.c_int64]
    L.inscy_debug_sort_temp.restype = ctypes.c_int64
    L.inscy_debug_sort_pairs.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int32, vp, ctypes.c_int64,
                                         vp]
    L.inscy_debug_sort_pairs.restype = ctypes.c_int
    L.inscy_debug_flag_compact_temp.argtypes = [ctypes.c_int64]
    L.inscy_debug_flag_compact_temp.restype = ctypes.c_int64
    L.inscy_debug_flag_compact.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, vp, vp]
    L.inscy_debug_trace_rows.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]
    L.inscy_debug_trace_rows.restype = ctypes.c_int64
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != INSCY_OK:
        raise InscyError(rc, lib().inscy_last_error().decode(errors="replace"))


def make_params(neighborhood_size, F, num_obj, min_size, r=1.0, number_of_cells=3, rectangular=False,
                entropy_order=0) -> Params:
    return Params(float(neighborhood_size), float(F), int(num_obj), int(min_size), float(r), int(number_of_cells),
                  int(bool(rectangular)), int(entropy_order))


class _Handle:
    """Owns an inscy_result; freed when the last array viewing it goes away."""

    def __init__(self, handle):
        self.handle = handle
        self._free = lib().inscy_result_free  # kept: module globals may be gone at exit

    def __del__(self):
        if self.handle:
            self._free(self.handle)
            self.handle = None


class Result:
    """An inscy_result: subspaces (ascending dims), canonical labels
    [#subspaces, n] int32 (a view of the library's contiguous label array, kept
    alive by the array itself), and the run counters."""

    def __init__(self, handle):
        L = lib()
        owner = _Handle(handle)
        self._owner = owner
        self._subspaces = None
        n = int(L.inscy_result_num_points(handle))
        m = int(L.inscy_result_num_subspaces(handle))
        self._m = m
        if m > 0 and n > 0:
            ptr = L.inscy_result_labels_all(handle)
            buf = (ctypes.c_int32 * (m * n)).from_address(ctypes.addressof(ptr.contents))
            buf._owner = owner  # the buffer keeps the result alive
            self.labels = np.frombuffer(buf, dtype=np.int32).reshape(m, n)
        else:
            self.labels = np.zeros((m, n), dtype=np.int32)
        st = Stats()
        check(L.inscy_result_stats(handle, ctypes.byref(st)))
        self.stats = {f: getattr(st, f) for f, _ in Stats._fields_}
        self.pairs = _pairs(handle, owner)

    @property
    def subspaces(self):
        """Dimension lists of the result subspaces (built on first access)."""
        if self._subspaces is None:
            L = lib()
            dims = (ctypes.c_int32 * 64)()
            out = []
            for i in range(self._m):
                k = L.inscy_result_subspace(self._owner.handle, i, dims)
                out.append(list(dims[:k]))
            self._subspaces = out
        return self._subspaces


def _pairs(handle, owner):
    """Training pairs of a collect-mode run (None otherwise): features
    [P, 64] float32, labels [P] uint8, info [P, 4] int32 (|S| of the tree, m,
    tree points, tree index), cost [P, 2] int64."""
    L = lib()
    f = ctypes.POINTER(ctypes.c_float)()
    y = ctypes.POINTER(ctypes.c_uint8)()
    info = ctypes.POINTER(ctypes.c_int32)()
    cost = ctypes.POINTER(ctypes.c_int64)()
    P = int(L.inscy_result_pairs(handle, ctypes.byref(f), ctypes.byref(y), ctypes.byref(info), ctypes.byref(cost)))
    if P == 0:
        return None

    def view(ptr, ctype, count, dtype, shape):
        buf = (ctype * count).from_address(ctypes.addressof(ptr.contents))
        buf._owner = owner
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    return {"features": view(f, ctypes.c_float, P * MLP_INPUTS, np.float32, (P, MLP_INPUTS)),
            "labels": view(y, ctypes.c_uint8, P, np.uint8, (P,)),
            "info": view(info, ctypes.c_int32, 4 * P, np.int32, (P, 4)),
            "cost": view(cost, ctypes.c_int64, 2 * P, np.int64, (P, 2))}


def set_pruner(weights, threshold: float, min_depth: int = 1) -> None:
    """Loads the learned pruner's network (MLP_PARAMS float32 values) for
    later learned-mode runs; weights=None unloads it."""
    L = lib()
    if weights is None:
        check(L.inscy_set_pruner(None, 0.0, 0))
        return
    w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
    if w.size != MLP_PARAMS:
        raise ValueError(f"the network has {MLP_PARAMS} parameters, got {w.size}")
    check(L.inscy_set_pruner(w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), float(threshold), int(min_depth)))


def run_host(X: np.ndarray, params: Params, mode: int = MODE_EXACT) -> Result:
    X = np.ascontiguousarray(X, dtype=np.float32)
    if X.ndim != 2:
        raise ValueError("X must be a 2-d [n, d] array")
    h = ctypes.c_void_p()
    check(lib().inscy_run(X.ctypes.data, X.shape[0], X.shape[1], ctypes.byref(params), mode, ctypes.byref(h)))
    return Result(h)


def run_device(ptr: int, n: int, d: int, params: Params, mode: int = MODE_EXACT, stream: int = 0) -> Result:
    h = ctypes.c_void_p()
    check(lib().inscy_run_device(ctypes.c_void_p(ptr), n, d, ctypes.byref(params), mode,
                                 ctypes.c_void_p(stream or None), ctypes.byref(h)))
    return Result(h)


def run_many(ptrs, shapes, params, mode: int = MODE_BATCHED, device: bool = True) -> list:
    """Runs independent data sets concurrently; ptrs are device (or host)
    addresses of row-major float32 arrays with the given (n, d) shapes."""
    count = len(ptrs)
    P = (vp_t * count)(*[ctypes.c_void_p(p) for p in ptrs])
    N = (ctypes.c_int64 * count)(*[int(s[0]) for s in shapes])
    D = (ctypes.c_int32 * count)(*[int(s[1]) for s in shapes])
    A = (Params * count)(*params)
    out = (vp_t * count)()
    fn = lib().inscy_run_many_device if device else lib().inscy_run_many
    check(fn(P, N, D, A, count, mode, out))
    return [Result(ctypes.c_void_p(out[i])) for i in range(count)]


def profile(on: bool) -> None:
    lib().inscy_profile(int(on))


def set_tuning(max_batch: int = 0, workers: int = 0) -> None:
    """Child subspaces per level-synchronous batch and run_many worker streams
    for later calls (0: the defaults).  Results do not depend on them."""
    check(lib().inscy_set_tuning(int(max_batch), int(workers)))


def debug_trace(on: bool) -> None:
    lib().inscy_debug_trace(int(on))


def debug_phi_exact(on: bool) -> None:
    lib().inscy_debug_phi_exact(int(on))


def debug_generic_walk(on: bool) -> None:
    """Test hook: run the neighbour walks for k known only at run time."""
    lib().inscy_debug_generic_walk(int(on))


def debug_keep_budget(nbytes: int) -> None:
    """Test hook: bytes a run keeps for index sorts and for sorted columns (each; < 0: default)."""
    lib().inscy_debug_keep_budget(int(nbytes))


def debug_flag_compact_temp(nb: int) -> int:
    """Test hook: device scratch bytes of debug_flag_compact over nb flags."""
    return int(lib().inscy_debug_flag_compact_temp(int(nb)))


def debug_flag_compact(flags_ptr: int, nb: int, temp_ptr: int, temp_bytes: int, out_ptr: int, nsel_ptr: int,
                       stream: int = 0) -> None:
    """Test hook: the ordered compaction of the query flags (device pointers)."""
    check(lib().inscy_debug_flag_compact(ctypes.c_void_p(flags_ptr), int(nb), ctypes.c_void_p(temp_ptr),
                                         int(temp_bytes), ctypes.c_void_p(out_ptr), ctypes.c_void_p(nsel_ptr),
                                         ctypes.c_void_p(stream or None)))


def debug_trace_rows() -> np.ndarray:
    L = lib()
    m = int(L.inscy_debug_trace_rows(None, 0))
    out = np.zeros((m, 7), dtype=np.int64)
    if m:
        L.inscy_debug_trace_rows(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), m)
    return out
