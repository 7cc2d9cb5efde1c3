"""Parity of the HIP engine with the oracle (CPU restatement of the reference's
CPU INSCY) — exact for every integer output: cells, neighbour counts,
components, subspace sets, canonical labels, per-tree decisions."""

import ctypes
import hashlib
import os
import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen
from gpu_inscy_amd import inscy as INSCY
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    _lib.lib()


def host_cells(X, cells, eps):
    cs = np.float32(1.0) / np.float32(cells)
    ec = int(np.ceil(1.0 / np.float64(np.float32(eps))))
    ecs = np.float32(1.0) / np.float32(ec)
    c = np.minimum((X / cs).astype(np.int32), cells - 1)
    e = np.minimum((X / ecs).astype(np.int32), ec - 1)
    upper = (c + 1).astype(np.float32) * cs
    bound = (upper - np.float32(eps)).astype(np.float32)
    near = X >= bound
    return c, e, near


@pytest.mark.parametrize("seed,cells,eps,d", [(0, 4, 0.01, 7), (1, 6, 0.015, 7), (2, 8, 0.02, 7), (3, 3, 0.1, 7),
                                              (4, 255, 0.003, 7), (5, 4, 0.05, 64), (6, 4, 0.05, 100),
                                              (7, 6, 0.02, 128)])
def test_discretise_bit_exact(seed, cells, eps, d):
    # near marks: (d + 63) // 64 words per point, dimension k at bit k % 64 of word k // 64
    say("discretise", seed, cells, eps, d)
    rng = np.random.default_rng(seed)
    X = rng.random((3000, d), dtype=np.float32)
    X[:5] = 0.0
    X[5:10] = 1.0
    # values on exact cell edges
    X[10:20, 0] = np.arange(10, dtype=np.float32) / np.float32(cells)
    X = np.clip(X, 0, 1).astype(np.float32)
    n, d = X.shape
    dX = dev(X)
    cc = torch.zeros((d, n), dtype=torch.uint8, device="cuda")
    ce = torch.zeros((d, n), dtype=torch.int16, device="cuda")
    nw = (d + 63) // 64
    nm = torch.full((n, nw), -7, dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().inscy_discretise(dX.data_ptr(), n, d, cells, eps, cc.data_ptr(), ce.data_ptr(),
                                           nm.data_ptr()))
    torch.cuda.synchronize()
    c, e, near = host_cells(X, cells, eps)
    assert np.array_equal(cc.cpu().numpy().T.astype(np.int32), c)
    assert np.array_equal(ce.cpu().numpy().T.astype(np.int32) & 0xFFFF, e)
    bits = nm.cpu().numpy().view(np.uint64)
    want = np.zeros((n, nw), dtype=np.uint64)
    for k in range(d):
        want[:, k // 64] |= near[:, k].astype(np.uint64) << np.uint64(k % 64)
    assert near.any(axis=0).all()
    assert np.array_equal(bits, want)


def test_discretise_rejects_out_of_range():
    X = np.full((10, 2), 0.5, dtype=np.float32)
    X[3, 1] = 1.5
    dX = dev(X)
    cc = torch.zeros((2, 10), dtype=torch.uint8, device="cuda")
    ce = torch.zeros((2, 10), dtype=torch.int16, device="cuda")
    nm = torch.zeros(10, dtype=torch.int64, device="cuda")
    rc = _lib.lib().inscy_discretise(dX.data_ptr(), 10, 2, 4, 0.1, cc.data_ptr(), ce.data_ptr(), nm.data_ptr())
    assert rc == _lib.INSCY_ERR_ARGUMENT


@pytest.mark.parametrize("seed,S,eps", [(0, [0], 0.01), (1, [1, 3], 0.015), (2, [0, 2, 4], 0.05),
                                        (3, [0, 1, 2, 3, 4, 5], 0.1), (4, [5], 0.2)])
def test_neighbour_counts_match_oracle(seed, S, eps):
    say("counts", seed, S, eps)
    w = datagen.small(seed, n=1200, d=6, n_clusters=3, eps=eps, sigma=0.04)
    X = w.X
    n, d = X.shape
    rng = np.random.default_rng(seed)
    qs = rng.choice(n, size=200, replace=False).astype(np.int32)
    dims = (ctypes.c_int32 * len(S))(*S)
    out = torch.zeros(len(qs), dtype=torch.int32, device="cuda")
    dq = dev(qs)
    dX = dev(X)  # held: a freed temporary's block could be handed to a later allocation
    _lib.check(_lib.lib().inscy_neighbour_counts(dX.data_ptr(), n, d, eps, dims, len(S), dq.data_ptr(),
                                                 len(qs), out.data_ptr()))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = np.array([len(O.neighbours(X, eps, S, int(q))) for q in qs])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("S", [[70], [3, 64, 99], [63, 64], [1, 2, 65, 90, 127]])
def test_neighbour_counts_wide_dimensions(S):
    # dimension ids beyond 64 (d = 128).  The neighbour rule reads only the
    # dimensions of S, so the counts equal those of the columns X[:, S] as a
    # k-dimensional set, which the oracle checks on a sample of the queries
    say("counts wide", S)
    rng = np.random.default_rng(len(S))
    n, d, eps = 2000, 128, 0.08
    X = rng.random((n, d), dtype=np.float32)
    X[: n // 2, S] = (0.5 + 0.02 * rng.standard_normal((n // 2, len(S)))).clip(0, 1).astype(np.float32)
    qs = rng.choice(n, size=300, replace=False).astype(np.int32)

    def counts(Xc, dims_list):
        dims = (ctypes.c_int32 * len(dims_list))(*dims_list)
        out = torch.zeros(len(qs), dtype=torch.int32, device="cuda")
        dX, dq = dev(Xc), dev(qs)  # held until the call returns
        _lib.check(_lib.lib().inscy_neighbour_counts(dX.data_ptr(), n, Xc.shape[1], eps, dims, len(dims_list),
                                                     dq.data_ptr(), len(qs), out.data_ptr()))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    got = counts(X, S)
    Xs = np.ascontiguousarray(X[:, S])
    assert got.sum() > 0
    assert np.array_equal(got, counts(Xs, list(range(len(S)))))
    want = np.array([len(O.neighbours(Xs, eps, list(range(len(S))), int(q))) for q in qs[:40]])
    assert np.array_equal(got[:40], want)


@pytest.mark.parametrize("seed,S", [(0, [0, 1]), (1, [2]), (2, [0, 2, 3])])
def test_dense_components_match_bruteforce(seed, S):
    say("components", seed, S)
    w = datagen.small(seed, n=800, d=4, n_clusters=3, eps=0.05, num_obj=4, sigma=0.03)
    X = w.X
    n, d = X.shape
    p = _lib.make_params(w.eps, 1.0, w.num_obj, 0, 1.0, w.cells, True, 0)
    dims = (ctypes.c_int32 * len(S))(*S)
    comp = torch.zeros(n, dtype=torch.int32, device="cuda")
    dX = dev(X)
    _lib.check(_lib.lib().inscy_dense_components(dX.data_ptr(), n, d, ctypes.byref(p), dims, len(S),
                                                 comp.data_ptr()))
    torch.cuda.synchronize()
    got = comp.cpu().numpy()
    lists = [O.neighbours(X, w.eps, S, i) for i in range(n)]
    thr = max(1.0 * O.thresholds(len(S), w.eps, n)["expden"], float(w.num_obj))
    dense = np.array([np.float32(len(l)) >= np.float32(thr) for l in lists])
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(n):
        if dense[i]:
            for j in lists[i]:
                if dense[j]:
                    a, b = find(i), find(j)
                    if a != b:
                        parent[max(a, b)] = min(a, b)
    want = np.array([find(i) if dense[i] else -1 for i in range(n)])
    assert np.array_equal(got, want)


def compare_with_oracle(w, mode=_lib.MODE_EXACT):
    ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, 0, trace=True)
    _lib.debug_trace(True)
    try:
        got = INSCY.run(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, 0, mode=mode)
        trace = _lib.debug_trace_rows()
    finally:
        _lib.debug_trace(False)
    assert got.subspaces == ref.subspaces
    for s, a, b in zip(ref.subspaces, ref.labels, got.labels):
        assert np.array_equal(O.canonical(a), b), f"labels differ in subspace {s}"
    assert got.stats["trees"] == ref.trees
    assert got.stats["calls"] == ref.calls
    assert np.array_equal(trace, ref.trace)
    return got


@pytest.mark.parametrize("seed", list(range(12)))
def test_end_to_end_small_cases(seed):
    rng = np.random.default_rng(seed + 500)
    n = int(rng.integers(100, 700))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.02, 0.05, 0.1]))
    w = datagen.small(seed, n=n, d=d, n_clusters=int(rng.integers(1, 4)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(2, 6)), min_size=int(rng.integers(2, n // 10 + 3)),
                      sigma=float(rng.choice([0.02, 0.05])))
    say("e2e small", seed, n, d, cells, eps)
    compare_with_oracle(w)


def test_end_to_end_config0():
    w = datagen.config(0)
    say("e2e config0", w.name)
    got = compare_with_oracle(w)
    assert len(got.subspaces) > 0


@pytest.mark.parametrize("seed", list(range(12)))
def test_batched_small_cases_match_oracle(seed):
    # the same cases as the depth-first test, through the level-synchronous
    # driver: identical subspaces, labels and per-tree decisions (trace rows
    # in the reference's depth-first order)
    rng = np.random.default_rng(seed + 500)
    n = int(rng.integers(100, 700))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.02, 0.05, 0.1]))
    w = datagen.small(seed, n=n, d=d, n_clusters=int(rng.integers(1, 4)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(2, 6)), min_size=int(rng.integers(2, n // 10 + 3)),
                      sigma=float(rng.choice([0.02, 0.05])))
    say("batched small", seed, n, d, cells, eps)
    compare_with_oracle(w, mode=_lib.MODE_BATCHED)


@pytest.mark.parametrize("seed", list(range(10)))
@pytest.mark.parametrize("mode", [_lib.MODE_BATCHED, _lib.MODE_EXACT])
def test_phi_density_small_cases_match_oracle(seed, mode):
    # rectangular=False (the reference wrapper's default): phi density
    # against alpha/omega thresholds, summed in the neighbourhood tree's leaf
    # order; both modes, traces included
    import dataclasses
    rng = np.random.default_rng(seed + 900)
    n = int(rng.integers(100, 700))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.03, 0.05, 0.1]))
    w = datagen.small(seed + 40, n=n, d=d, n_clusters=int(rng.integers(1, 4)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(1, 5)), min_size=int(rng.integers(2, n // 10 + 3)),
                      sigma=float(rng.choice([0.02, 0.05])))
    w = dataclasses.replace(w, rectangular=False, F=float(rng.choice([0.5, 1.0, 2.0])))
    say("phi small", seed, mode, n, d, cells, eps, w.F)
    compare_with_oracle(w, mode=mode)


@pytest.mark.parametrize("seed", [0, 3, 7])
def test_phi_leaf_order_recomputation_matches(seed):
    # every phi decided by the exact leaf-order kernel (the band path off):
    # the same result, which pins that kernel against the oracle directly
    import dataclasses
    rng = np.random.default_rng(seed + 900)
    n = int(rng.integers(100, 700))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.03, 0.05, 0.1]))
    w = datagen.small(seed + 40, n=n, d=d, n_clusters=int(rng.integers(1, 4)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(1, 5)), min_size=int(rng.integers(2, n // 10 + 3)),
                      sigma=float(rng.choice([0.02, 0.05])))
    w = dataclasses.replace(w, rectangular=False, F=float(rng.choice([0.5, 1.0, 2.0])))
    say("phi leaf order", seed)
    _lib.debug_phi_exact(True)
    try:
        compare_with_oracle(w, mode=_lib.MODE_BATCHED)
    finally:
        _lib.debug_phi_exact(False)


def test_phi_density_config0_matches_golden():
    import dataclasses
    w = dataclasses.replace(datagen.config(0), rectangular=False)
    say("phi config0", w.name)
    got = compare_with_oracle(w, mode=_lib.MODE_BATCHED)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "config0_phi.npz"))
    assert [sum(1 << k for k in s) for s in got.subspaces] == list(g["subspace_masks"])
    assert np.array_equal(np.stack(got.labels) if len(got.labels) else got.labels, g["labels"])


def test_batched_config0_matches_oracle():
    w = datagen.config(0)
    say("batched config0", w.name)
    got = compare_with_oracle(w, mode=_lib.MODE_BATCHED)
    assert len(got.subspaces) > 0


def test_drop_in_lists_and_cuda_input():
    w = datagen.config(0)
    say("drop-in", w.name)
    subs, cls = INSCY.CPU(torch.from_numpy(w.X), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    subs2, cls2 = INSCY.GPU(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    assert subs == subs2 and cls == cls2
    assert len(cls[0]) == w.n
    assert INSCY.count_number_of_clusters(subs, cls) >= len(subs)


def digests(res):
    import hashlib
    return [hashlib.sha256(np.ascontiguousarray(l, dtype=np.int32).tobytes()).hexdigest() for l in res.labels]


def test_config1_matches_golden():
    import json
    import os
    say("config1 golden")
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))
    w = datagen.config(1)
    res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


def test_batched_config1_matches_golden():
    import json
    import os
    say("batched config1 golden")
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))
    w = datagen.config(1)
    res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0,
                    mode=_lib.MODE_BATCHED)
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


def test_run_many_equals_single_runs():
    # independent data sets on concurrent streams: every result equals the
    # single run of its data set (both density modes, both drivers)
    sets = datagen.multi_instance(count=10)
    jobs = []
    for i, w in enumerate(sets):
        rect = i % 3 != 2
        jobs.append((torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, 0))
    say("run_many", len(jobs))
    for mode in (_lib.MODE_BATCHED, _lib.MODE_EXACT):
        many = INSCY.run_many(jobs, mode=mode)
        assert len(many) == len(jobs)
        for job, got in zip(jobs, many):
            one = INSCY.run(*job, mode=mode)
            assert got.subspaces == one.subspaces
            assert np.array_equal(got.labels, one.labels)
            assert got.stats["trees"] == one.stats["trees"]


def test_run_many_host_input_and_errors():
    say("run_many host")
    sets = datagen.multi_instance(count=3)
    jobs = [(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0) for w in sets]
    many = INSCY.run_many(jobs)
    for job, got in zip(jobs, many):
        one = INSCY.run(*job, mode=_lib.MODE_BATCHED)
        assert got.subspaces == one.subspaces and np.array_equal(got.labels, one.labels)
    bad = list(jobs)
    X = np.array(bad[1][0], copy=True)
    X[3, 0] = 2.0
    bad[1] = (X,) + bad[1][1:]
    with pytest.raises(_lib.InscyError) as e:
        INSCY.run_many(bad)
    assert e.value.code == _lib.INSCY_ERR_ARGUMENT and "data set 1" in str(e.value)


def test_phi_density_config1_matches_golden():
    import json
    say("phi config1 golden")
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1_phi.json")))
    w = datagen.config(1)
    res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, False, 0,
                    mode=_lib.MODE_BATCHED)
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


@pytest.mark.parametrize("rect", [True, False])
def test_runtime_k_walks_match_golden(rect):
    # the walks compiled for a launch's k against the walks for k known only
    # at run time (the path of batches that mix dimensionalities): both give
    # the golden result, with the same counters
    import json
    say("runtime-k walks", rect)
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "config1.json" if rect else "config1_phi.json")))
    w = datagen.config(1)
    X = torch.from_numpy(w.X).cuda()
    run = lambda: INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, 0, mode=_lib.MODE_BATCHED)
    compiled = run()
    _lib.debug_generic_walk(True)
    try:
        generic = run()
    finally:
        _lib.debug_generic_walk(False)
    for res in (compiled, generic):
        assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
        assert digests(res) == [x[3] for x in g["subspaces"]]
        assert res.stats["trees"] == g["trees"]
    assert compiled.stats["queries"] == generic.stats["queries"]
    # most grid indexes of the run reuse a sort kept from an earlier batch
    assert compiled.stats["sorts_kept"] > 0


@pytest.mark.parametrize("budget", [0, 3 << 20, 64 << 20])
def test_kept_sorts_and_columns_budget(budget):
    # nothing kept, or the budget spent part way through the run (later
    # indexes are sorted and their columns gathered per batch, and a column
    # is never kept for an index whose sort was not): the golden config-1
    # result either way
    import json
    say("keep budget", budget)
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))
    w = datagen.config(1)
    _lib.debug_keep_budget(budget)
    try:
        res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0,
                        mode=_lib.MODE_BATCHED)
    finally:
        _lib.debug_keep_budget(-1)
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    say("sorts kept", res.stats["sorts_kept"])
    assert res.stats["trees"] == g["trees"]
    if budget == 0:
        assert res.stats["sorts_kept"] == 0


REAL_DIR = os.path.join(os.path.dirname(__file__), "golden", "real")


def _real_golden():
    import json
    path = os.path.join(REAL_DIR, "real_golden.json")
    return json.load(open(path)) if os.path.exists(path) else {}


@pytest.mark.parametrize("name", ["glass", "vowel", "pendigits", "sky_0.5x0.5", "sky_1x1", "sky_1x2"])
def test_real_data_matches_golden(name):
    # the reference's published real-data workloads (experiments/run_real.py)
    g = _real_golden().get(name)
    if g is None:
        pytest.skip(f"no golden result for {name}")
    say("real", name)
    w = datagen.real(name, os.path.join(REAL_DIR, "real_inputs.npz"))
    res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0,
                    mode=_lib.MODE_BATCHED)
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]
    if w.n <= 1000:
        ex = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0,
                       mode=_lib.MODE_EXACT)
        assert ex.subspaces == res.subspaces and np.array_equal(ex.labels, res.labels)


def test_out_of_range_input_is_an_error():
    say("out of range")
    X = np.full((50, 3), 0.5, dtype=np.float32)
    X[7, 2] = 1.25
    with pytest.raises(_lib.InscyError) as e:
        INSCY.run(torch.from_numpy(X).cuda(), 0.1, 1.0, 2, 5, 1.0, 3, True, 0)
    assert e.value.code == _lib.INSCY_ERR_ARGUMENT


def test_caller_stream_and_host_input_agree():
    say("streams")
    w = datagen.config(0)
    a = INSCY.run(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        X = torch.from_numpy(w.X).cuda()
        b = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    torch.cuda.synchronize()
    assert a.subspaces == b.subspaces and np.array_equal(a.labels, b.labels)


def check_invariants(res, w):
    masks = [sum(1 << k for k in s) for s in res.subspaces]
    assert masks == sorted(masks) and len(set(masks)) == len(masks)
    for lab in res.labels:
        assert lab.min() >= -1
        k = int(lab.max()) + 1
        assert k >= 1
        # canonical: cluster c's smallest member precedes cluster c+1's
        firsts = [int(np.flatnonzero(lab == c)[0]) for c in range(k)]
        assert firsts == sorted(firsts)
        sizes = np.bincount(lab[lab >= 0])
        assert sizes.min() >= w.min_size


GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    import json
    path = os.path.join(GOLDEN, name)
    return json.load(open(path)) if os.path.exists(path) else None


def assert_golden(res, g):
    assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
    assert [int(l.max()) + 1 for l in res.labels] == [x[1] for x in g["subspaces"]]
    assert digests(res) == [x[3] for x in g["subspaces"]]
    assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


@pytest.mark.parametrize("case", [0, 1, 2])
def test_wide_dimensions_match_oracle(case):
    # d = 100 and 127: subspaces, s-connection paths, near-edge marks and the
    # result order with dimension ids beyond 64, both drivers against the
    # oracle's digests (tests/golden/wide.json)
    g = _golden("wide.json")
    if g is None:
        pytest.skip("no oracle digests for the wide cases (tests/golden/make_golden.py wide)")
    m = g["sets"][case]
    w = datagen.wide(m["n"], m["d"], cells=m["cells"], rectangular=m["rectangular"])
    say("wide", w.name)
    assert m["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest()
    assert any(len(x[0]) > 1 and max(x[0]) >= 64 for x in m["subspaces"])
    X = torch.from_numpy(w.X).cuda()
    for mode in (_lib.MODE_BATCHED, _lib.MODE_EXACT):
        res = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, 0, mode=mode)
        assert_golden(res, m)


def _many_clusters(case):
    rng = np.random.default_rng(5 + case)
    if case == 0:
        # 300 clusters of 10 points along dimension 0: labels read back in two bytes
        k, per = 300, 10
        X = np.empty((k * per, 2), dtype=np.float32)
        X[:, 0] = np.repeat((np.arange(k) + 0.5) / k, per) + rng.uniform(-1e-4, 1e-4, k * per)
        X[:, 1] = rng.random(k * per)
        return X.clip(0, 1).astype(np.float32), 0.001, 3, 5
    # a 257 x 257 grid of two-point clusters: 66,049 clusters in {0, 1}, labels in four bytes
    g = 257
    a = (np.arange(g) + 0.5) / g
    cx, cy = np.meshgrid(a, a, indexing="ij")
    X = np.repeat(np.stack([cx.ravel(), cy.ravel()], 1), 2, axis=0) + rng.uniform(-2e-5, 2e-5, (2 * g * g, 2))
    return X.clip(0, 1).astype(np.float32), 0.0005, 1, 1


@pytest.mark.parametrize("case", [0, 1])
def test_rows_with_many_clusters_match_oracle(case):
    # the labels of rows with more than 255 (65,535) clusters leave the device
    # in two (four) bytes; fewer clusters, as in every other test, in one
    X, eps, mu, ms = _many_clusters(case)
    say("many clusters", case, X.shape)
    ref = O.run_ref(X, eps, 1.0, mu, ms, 1.0, 4, True)
    assert max(int(l.max()) + 1 for l in ref.labels) > (255 if case == 0 else 65535)
    for mode in (_lib.MODE_BATCHED, _lib.MODE_EXACT):
        res = INSCY.run(torch.from_numpy(X).cuda(), eps, 1.0, mu, ms, 1.0, 4, True, 0, mode=mode)
        assert res.subspaces == ref.subspaces
        for got, want in zip(res.labels, ref.labels):
            assert np.array_equal(got, O.canonical(want))


def test_config2_full_size_matches_oracle_and_is_deterministic():
    # n=100,000 d=20, c=8, eps=0.02: both drivers against the oracle's digests
    # (set restatement, tests/golden/config2.json), and run to run
    say("config2 full size")
    w = datagen.config(2)
    g = _golden("config2.json")
    if g is not None:
        assert g["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest()
    X = torch.from_numpy(w.X).cuda()
    c = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, mode=_lib.MODE_BATCHED)
    say("config2 batched", c.stats)
    check_invariants(c, w)
    if g is not None:
        assert_golden(c, g)
    b = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, mode=_lib.MODE_BATCHED)
    assert c.subspaces == b.subspaces and np.array_equal(c.labels, b.labels)
    del b
    # the depth-first driver gives the same result, trees and calls
    a = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, mode=_lib.MODE_EXACT)
    say("config2 depth-first", a.stats)
    assert a.subspaces == c.subspaces and np.array_equal(a.labels, c.labels)
    assert a.stats["trees"] == c.stats["trees"] and a.stats["calls"] == c.stats["calls"]
    if g is None:
        pytest.skip("no oracle digests for config 2 (tests/golden/make_golden.py 2): checked run to run only")


def test_config2_parameters_at_25k_match_pointer_oracle():
    # config 2's parameters (c=8, eps=0.02, d=20, Zipf sizes) at n=25,000,
    # against the pointer-tree restatement (tests/golden/config2_small.json)
    g = _golden("config2_small.json")
    if g is None:
        pytest.skip("no golden for config 2's parameters at n=25k")
    w = datagen.config2_small()
    assert g["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest()
    say("config2 params n=25k")
    X = torch.from_numpy(w.X).cuda()
    for mode in (_lib.MODE_BATCHED, _lib.MODE_EXACT):
        res = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, mode=mode)
        assert_golden(res, g)


def test_run_many_matches_config3_oracle():
    # the 64 independent data sets through one run_many call, each against
    # the oracle's digests (tests/golden/config3.json)
    # (the digests are written set by set; while the oracle's run is
    # incomplete the sets it has finished are compared)
    g = _golden("config3.json")
    sets = datagen.multi_instance()
    if g is None or not g.get("sets"):
        pytest.skip("no oracle digests for the 64 data sets")
    say("run_many 64 sets vs oracle", len(g["sets"]), "with digests")
    jobs = [(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0) for w in sets]
    many = INSCY.run_many(jobs, mode=_lib.MODE_BATCHED)
    assert len(many) == len(sets) == 64
    by_name = {gs["workload"]: gs for gs in g["sets"]}
    if g.get("complete"):
        assert len(by_name) == len(sets)
    checked = 0
    for w, res in zip(sets, many):
        gs = by_name.get(w.name)
        if gs is None:
            continue
        assert gs["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest()
        assert_golden(res, gs)
        checked += 1
    assert checked == len(by_name) > 0


def test_config4_full_size_deterministic_across_keep_budgets():
    # n=200,000 d=32 (the learned-prune evaluation set): the level-synchronous
    # run gives one result with its kept index sorts and columns, with none
    # kept, and with the budget spent part way; canonical labels, sizes
    # >= min_size and the result order hold
    say("config4 full size")
    w = datagen.config(4)
    X = torch.from_numpy(w.X).cuda()
    run = lambda: INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, mode=_lib.MODE_BATCHED)
    a = run()
    say("config4", a.stats)
    assert a.stats["sorts_kept"] > 0
    masks = [sum(1 << k for k in s) for s in a.subspaces]
    assert masks == sorted(masks) and len(set(masks)) == len(masks)
    lab = a.labels
    assert lab.min() >= -1
    for i in range(0, len(lab), 97):  # a spread sample of rows for the per-cluster checks
        row = lab[i]
        k = int(row.max()) + 1
        sizes = np.bincount(row[row >= 0], minlength=k)
        assert k >= 1 and sizes.min() >= w.min_size
        firsts = np.array([np.argmax(row == c) for c in range(k)])
        assert np.all(np.diff(firsts) > 0)
    for budget in (0, 1 << 30):
        _lib.debug_keep_budget(budget)
        try:
            b = run()
        finally:
            _lib.debug_keep_budget(-1)
        say("config4 keep budget", budget, b.stats["sorts_kept"])
        if budget == 0:
            assert b.stats["sorts_kept"] == 0
        assert b.subspaces == a.subspaces and np.array_equal(b.labels, a.labels)
        assert b.stats["trees"] == a.stats["trees"] and b.stats["calls"] == a.stats["calls"]
        del b


def degenerate_cases():
    # inputs at the edges of the algorithm: one point, one dimension,
    # duplicates only, a neighbourhood as wide as the data, one c-grid cell,
    # min_size above n, thresholds no point meets, r = 0
    rng = np.random.default_rng(7)
    X = rng.random((80, 3)).astype(np.float32)
    X[:30] = 0.2 + 0.01 * rng.random((30, 3)).astype(np.float32)
    X1 = np.concatenate([0.3 + 0.02 * rng.random((40, 1)), rng.random((20, 1))]).astype(np.float32)
    return {
        "n1_d1": (np.array([[0.5]], np.float32), 0.05, 1.0, 1, 1, 1.0, 3),
        "n2_same_d3": (np.full((2, 3), 0.25, np.float32), 0.05, 1.0, 1, 1, 1.0, 3),
        "dups50_d2": (np.full((50, 2), 0.5, np.float32), 0.05, 1.0, 2, 5, 1.0, 4),
        "eps1_d3": (X, 1.0, 1.0, 2, 5, 1.0, 3),
        "cells1_d3": (X, 0.05, 1.0, 2, 5, 1.0, 1),
        "min_size_above_n": (X, 0.05, 1.0, 2, 500, 1.0, 3),
        "F100": (X, 0.05, 100.0, 2, 5, 1.0, 3),
        "r0": (X, 0.05, 1.0, 2, 5, 0.0, 3),
        "d1_cluster": (X1, 0.02, 1.0, 2, 5, 1.0, 3),
    }


@pytest.mark.parametrize("name", sorted(degenerate_cases()))
@pytest.mark.parametrize("rect", [True, False])
def test_degenerate_inputs_match_oracle(name, rect):
    X, eps, F, num_obj, min_size, r, c = degenerate_cases()[name]
    w = datagen.Workload(name, X, np.full(len(X), -1, np.int32), [], eps=eps, cells=c, F=F, num_obj=num_obj,
                         min_size=min_size, r=r, rectangular=rect)
    say("degenerate", name, rect)
    for mode in (_lib.MODE_EXACT, _lib.MODE_BATCHED):
        compare_with_oracle(w, mode)
