"""ctypes front-end of the oracle (TEST INFRASTRUCTURE ONLY).

Loads liboracle_inscy.so built from this folder by its Makefile.  Only tests,
the smoke check and the CPU-baseline leg of the benchmark import this module,
and only to check or to time against; the engine never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_inscy.so")
_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        i32p = ctypes.POINTER(ctypes.c_int)
        i64p = ctypes.POINTER(ctypes.c_longlong)
        L.oracle_ref_run.restype = ctypes.c_void_p
        L.oracle_ref_run.argtypes = [f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_ref_count.argtypes = [ctypes.c_void_p]
        L.oracle_ref_get.argtypes = [ctypes.c_void_p, ctypes.c_int, i32p, i32p]
        L.oracle_ref_stats.argtypes = [ctypes.c_void_p, i64p]
        L.oracle_ref_trace_len.argtypes = [ctypes.c_void_p]
        L.oracle_ref_trace.argtypes = [ctypes.c_void_p, i64p]
        L.oracle_ref_free.argtypes = [ctypes.c_void_p]
        L.oracle_unit_ball.restype = ctypes.c_float
        L.oracle_unit_ball.argtypes = [ctypes.c_int]
        L.oracle_alpha.restype = ctypes.c_float
        L.oracle_alpha.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.oracle_expden.restype = ctypes.c_float
        L.oracle_expden.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.oracle_omega.restype = ctypes.c_float
        L.oracle_omega.argtypes = [ctypes.c_int]
        L.oracle_neighbours.argtypes = [f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, i32p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, i32p]
        L.oracle_dimension_order.argtypes = [f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class OracleRun:
    """Result of one oracle run: subspaces (ascending dim lists) in the
    reference's output order, per-subspace label arrays, counters, trace."""

    def __init__(self, subspaces, labels, stats, trace):
        self.subspaces = subspaces
        self.labels = labels
        self.calls, self.trees, self.queries, self.nbr_total = stats
        self.trace = trace


def run_ref(X, eps, F, num_obj, min_size, r=1.0, cells=3, rectangular=False, entropy_order=0,
            literal=False, trace=False) -> OracleRun:
    """The reference CPU INSCY (run_cpu3_weak) restated; X float32 [n, d]."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    L = lib()
    h = L.oracle_ref_run(_fp(X), n, d, eps, F, num_obj, min_size, r, cells, int(rectangular), entropy_order,
                         int(literal), int(trace))
    try:
        cnt = L.oracle_ref_count(h)
        subs, labs = [], []
        dims = np.zeros(max(d, 1), dtype=np.int32)
        for i in range(cnt):
            lab = np.zeros(n, dtype=np.int32)
            k = L.oracle_ref_get(h, i, _ip(dims), _ip(lab))
            subs.append([int(x) for x in dims[:k]])
            labs.append(lab)
        st = np.zeros(4, dtype=np.int64)
        L.oracle_ref_stats(h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        tr = None
        if trace:
            m = L.oracle_ref_trace_len(h)
            tr = np.zeros((m, 7), dtype=np.int64)
            if m:
                L.oracle_ref_trace(h, tr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return OracleRun(subs, labs, [int(x) for x in st], tr)
    finally:
        L.oracle_ref_free(h)


def ref_rate(X, eps, F, num_obj, min_size, r=1.0, cells=3, rectangular=False, max_seconds=20.0):
    """Times the oracle on X for at most max_seconds: (trees evaluated, seconds, finished)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    L = lib()
    L.oracle_ref_rate.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_int)]
    trees, secs, fin = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_int()
    L.oracle_ref_rate(_fp(X), n, d, eps, F, num_obj, min_size, r, cells, int(rectangular), float(max_seconds),
                      ctypes.byref(trees), ctypes.byref(secs), ctypes.byref(fin))
    return int(trees.value), float(secs.value), bool(fin.value)


def neighbours(X, eps, S, p, literal=False):
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    S = np.ascontiguousarray(S, dtype=np.int32)
    out = np.zeros(n, dtype=np.int32)
    m = lib().oracle_neighbours(_fp(X), n, d, eps, _ip(S), len(S), p, int(literal), _ip(out))
    return out[:m].copy()


def phi_many(X, eps, S, ps, literal=False):
    """φ_S(p) of the points ps over their ε-neighbours in the neighbourhood
    tree's leaf order (ClusteringCpu.cu:116-124)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    S = np.ascontiguousarray(S, dtype=np.int32)
    ps = np.ascontiguousarray(ps, dtype=np.int32)
    out = np.zeros(len(ps), dtype=np.float32)
    L = lib()
    L.oracle_phi_many.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                  ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                  ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.oracle_phi_many(_fp(X), n, d, eps, _ip(S), len(S), _ip(ps), len(ps), int(literal), _fp(out))
    return out


def dimension_order(X, cells, entropy_order):
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    out = np.zeros(d, dtype=np.int32)
    lib().oracle_dimension_order(_fp(X), n, d, cells, entropy_order, _ip(out))
    return out


def thresholds(k, eps, n):
    L = lib()
    return dict(c=L.oracle_unit_ball(k), alpha=L.oracle_alpha(k, eps, n), expden=L.oracle_expden(k, eps, n),
                omega=L.oracle_omega(k))


def canonical(labels: np.ndarray) -> np.ndarray:
    """Negatives -> -1; clusters renumbered 0.. in order of their minimum point id."""
    lab = np.asarray(labels)
    out = np.full(lab.shape, -1, dtype=np.int32)
    pos = lab >= 0
    if not pos.any():
        return out
    ids = np.unique(lab[pos])
    first = np.array([np.flatnonzero(lab == c)[0] for c in ids])
    order = ids[np.argsort(first)]
    remap = {int(c): i for i, c in enumerate(order)}
    out[pos] = np.array([remap[int(c)] for c in lab[pos]], dtype=np.int32)
    return out


def _check_prefix_keys(cells, d):
    # the set restatement keys a cell prefix as a base-`cells` number in 128 bits
    if cells ** d >= 2 ** 128:
        raise ValueError(f"set restatement: cells**d must be below 2**128 (cells={cells}, d={d}); use run()")


def run_sets(X, eps, F, num_obj, min_size, r=1.0, cells=3, rectangular=False, threads=0,
             trace=False) -> OracleRun:
    """Set-based restatement of the same algorithm (identity dimension order)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    _check_prefix_keys(cells, d)
    L = lib()
    L.oracle_sets_run.restype = ctypes.c_void_p
    L.oracle_sets_run.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int]
    h = L.oracle_sets_run(_fp(X), n, d, eps, F, num_obj, min_size, r, cells, int(rectangular), threads,
                          int(trace))
    try:
        cnt = L.oracle_ref_count(h)
        subs, labs = [], []
        dims = np.zeros(max(d, 1), dtype=np.int32)
        for i in range(cnt):
            lab = np.zeros(n, dtype=np.int32)
            k = L.oracle_ref_get(h, i, _ip(dims), _ip(lab))
            subs.append([int(x) for x in dims[:k]])
            labs.append(lab)
        st = np.zeros(4, dtype=np.int64)
        L.oracle_ref_stats(h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        tr = None
        if trace:
            m = L.oracle_ref_trace_len(h)
            tr = np.zeros((m, 7), dtype=np.int64)
            if m:
                L.oracle_ref_trace(h, tr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return OracleRun(subs, labs, [int(x) for x in st], tr)
    finally:
        L.oracle_ref_free(h)


def run_sets_order(X, eps, F, num_obj, min_size, r=1.0, cells=3, rectangular=False, entropy_order=0, threads=0,
                   trace=False) -> OracleRun:
    """Set-based restatement for any SCY-tree level order (s-connection
    chains with their last level)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    _check_prefix_keys(cells, d)
    L = lib()
    L.oracle_sets_run_order.restype = ctypes.c_void_p
    L.oracle_sets_run_order.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    h = L.oracle_sets_run_order(_fp(X), n, d, eps, F, num_obj, min_size, r, cells, int(rectangular), entropy_order,
                                threads, int(trace))
    return _collect(h, n, d, trace)


def _collect(h, n, d, trace) -> OracleRun:
    L = lib()
    try:
        cnt = L.oracle_ref_count(h)
        subs, labs = [], []
        dims = np.zeros(max(d, 1), dtype=np.int32)
        for i in range(cnt):
            lab = np.zeros(n, dtype=np.int32)
            k = L.oracle_ref_get(h, i, _ip(dims), _ip(lab))
            subs.append([int(x) for x in dims[:k]])
            labs.append(lab)
        st = np.zeros(4, dtype=np.int64)
        L.oracle_ref_stats(h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        tr = None
        if trace:
            m = L.oracle_ref_trace_len(h)
            tr = np.zeros((m, 7), dtype=np.int64)
            if m:
                L.oracle_ref_trace(h, tr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return OracleRun(subs, labs, [int(x) for x in st], tr)
    finally:
        L.oracle_ref_free(h)


def same_result(a: OracleRun, b: OracleRun) -> tuple:
    """(equal, message): subspace lists and canonical labels identical."""
    if a.subspaces != b.subspaces:
        return False, f"subspaces differ: {a.subspaces} vs {b.subspaces}"
    for s, la, lb in zip(a.subspaces, a.labels, b.labels):
        ca, cb = canonical(la), canonical(lb)
        if not np.array_equal(ca, cb):
            return False, f"labels differ in {s}: {int((ca != cb).sum())} points"
    return True, "ok"


def _flat_result(result, n):
    ks = np.array([len(s) for s, _ in result], dtype=np.int32)
    dims = np.array([x for s, _ in result for x in s] or [0], dtype=np.int32)
    labs = np.ascontiguousarray(np.array([l for _, l in result], dtype=np.int32).reshape(len(result), n))
    return ks, dims, labs


def join_test(result, cl, S, min_size, r=1.0):
    """join (util.cu:655-743) on a hand-built result [(dims, labels)] and
    clustering cl of subspace S; returns the updated result in map order."""
    n = len(cl)
    L = lib()
    ks, dims, labs = _flat_result(result, n)
    m = len(result) + 1
    ok = np.zeros(m, dtype=np.int32)
    od = np.zeros(64 * m, dtype=np.int32)
    ol = np.zeros((m, n), dtype=np.int32)
    S = np.asarray(S, dtype=np.int32)
    cl = np.ascontiguousarray(cl, dtype=np.int32)
    L.oracle_join_test.restype = ctypes.c_int
    e = L.oracle_join_test(n, len(result), _ip(ks), _ip(dims), _ip(labs) if len(result) else None, _ip(cl),
                           len(S), _ip(S), min_size, ctypes.c_float(r), _ip(ok), _ip(od), _ip(ol))
    out, at = [], 0
    for i in range(e):
        out.append((od[at:at + ok[i]].tolist(), ol[i].tolist()))
        at += ok[i]
    return out


def redundancy_test(result, n, pts, S, r=1.0):
    """pruneRedundancy (ScyTreeNode.cu:669-734) of a region with points pts in S."""
    L = lib()
    ks, dims, labs = _flat_result(result, n)
    pts = np.ascontiguousarray(pts, dtype=np.int32)
    S = np.asarray(S, dtype=np.int32)
    L.oracle_redundancy_test.restype = ctypes.c_int
    return bool(L.oracle_redundancy_test(n, len(result), _ip(ks), _ip(dims), _ip(labs) if len(result) else None,
                                         _ip(pts), len(pts), len(S), _ip(S), ctypes.c_float(r)))


def mlp_forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    """FP64 logits of the learned pruner's network (oracle/mlp_ref.cpp)."""
    L = lib()
    L.oracle_mlp_forward.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                     ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.zeros(X.shape[0], dtype=np.float64)
    L.oracle_mlp_forward(_fp(P), _fp(X), X.shape[0], out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def mlp_grad(params: np.ndarray, X: np.ndarray, Y: np.ndarray, WT: np.ndarray, inv_wsum: float):
    """FP64 gradient and value of inv_wsum * sum_i WT_i * BCE(logit_i, Y_i)
    (oracle/mlp_ref.cpp); returns (grad [params], loss)."""
    L = lib()
    L.oracle_mlp_grad.argtypes = [ctypes.POINTER(ctypes.c_float)] * 4 + [ctypes.c_longlong, ctypes.c_double,
                                                                          ctypes.POINTER(ctypes.c_double)]
    L.oracle_mlp_params.restype = ctypes.c_int
    npar = int(L.oracle_mlp_params())
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    WT = np.ascontiguousarray(WT, dtype=np.float32)
    out = np.zeros(npar + 1, dtype=np.float64)
    L.oracle_mlp_grad(_fp(P), _fp(X), _fp(Y), _fp(WT), X.shape[0], float(inv_wsum),
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out[:npar], float(out[npar])
