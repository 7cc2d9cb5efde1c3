"""entropy_order != 0: the SCY-tree levels in the entropy order of the
dimensions (get_subspace_order, inscy_map.cpp:61-121).  The engine's
depth-first driver over s-connection chains against the oracle's pointer-tree
restatement (restrict_node's handling of a permuted level order,
ScyTreeNode.cu:227-320, mergeNodes :803-833, propergate_count2 :446-467):
subspaces, canonical labels, counters and the per-tree decisions."""

import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen
from gpu_inscy_amd import inscy as INSCY
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    _lib.lib()


def case(seed):
    rng = np.random.default_rng(20_000 + seed)
    d = int(rng.integers(2, 7))
    n = int(rng.integers(80, 420))
    cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.02, 0.04, 0.06, 0.1, 0.15]))
    return datagen.small(20_000 + seed, n=n, d=d, n_clusters=int(rng.integers(1, 5)), eps=eps, cells=cells,
                         num_obj=int(rng.integers(2, 6)), min_size=int(rng.integers(3, 20)),
                         sigma=float(rng.choice([0.02, 0.04, 0.07, 0.12])), sub_dims=(1, min(d, 4)))


def check(w, order, mode, rect=True):
    ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, order, trace=True)
    _lib.debug_trace(True)
    try:
        res = INSCY.run(torch.from_numpy(w.X).cuda(), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, order,
                        mode=mode)
        rows = _lib.debug_trace_rows()
    finally:
        _lib.debug_trace(False)
    assert res.subspaces == ref.subspaces
    for s, a, b in zip(ref.subspaces, ref.labels, res.labels):
        assert np.array_equal(O.canonical(a), b), f"labels differ in {s}"
    assert res.stats["trees"] == ref.trees and res.stats["calls"] == ref.calls
    assert np.array_equal(rows, ref.trace)


@pytest.mark.parametrize("seed", range(24))
def test_permuted_level_order_matches_oracle(seed):
    w = case(seed)
    for order in (1, -1):
        say("entropy", seed, order, w.n, w.d, w.cells)
        check(w, order, _lib.MODE_EXACT)


@pytest.mark.parametrize("seed", range(16))
def test_permuted_level_order_with_phi_matches_oracle(seed):
    # rectangular = false: phi summed in the neighbourhood tree's leaf order,
    # whose levels follow the same permuted order
    w = case(seed)
    for order in (1, -1):
        say("entropy phi", seed, order, w.n, w.d, w.cells)
        check(w, order, _lib.MODE_EXACT, rect=False)


@pytest.mark.parametrize("order", [1, -1])
def test_config0_with_entropy_order_and_phi(order):
    w = datagen.config(0)
    say("config0 entropy phi", order)
    check(w, order, _lib.MODE_EXACT, rect=False)
    check(w, order, _lib.MODE_BATCHED, rect=False)
    # every phi by the leaf-order recomputation: the same result as the
    # grid pass with its rounding band
    _lib.debug_phi_exact(True)
    try:
        check(w, order, _lib.MODE_EXACT, rect=False)
    finally:
        _lib.debug_phi_exact(False)


@pytest.mark.parametrize("order", [1, -1])
def test_config0_with_entropy_order(order):
    w = datagen.config(0)
    say("config0 entropy", order)
    assert list(O.dimension_order(w.X, w.cells, order)) != list(range(w.d))
    check(w, order, _lib.MODE_EXACT)
    # the level-synchronous mode names the same exact result
    check(w, order, _lib.MODE_BATCHED)


def test_entropy_order_limits():
    w = datagen.config(0)
    X = torch.from_numpy(w.X).cuda()
    with pytest.raises(_lib.InscyError) as e:
        INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 1, mode=_lib.MODE_COLLECT)
    assert e.value.code == _lib.INSCY_ERR_UNSUPPORTED


@pytest.mark.parametrize("order,name", [(1, "config1_entropy.json"), (-1, "config1_entropy_desc.json")])
def test_config1_with_entropy_order_matches_golden(order, name):
    # n=25,000 d=15 with the levels in ascending / descending entropy order:
    # both modes against the oracle's digests (set restatement of a permuted
    # level order, tests/golden/make_golden.py 1ent / 1ent-)
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", name)))
    w = datagen.config(1)
    assert g["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest() and g["entropy_order"] == order
    X = torch.from_numpy(w.X).cuda()
    for mode in (_lib.MODE_EXACT, _lib.MODE_BATCHED):
        say("config1 entropy", order, mode)
        res = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order, mode=mode)
        assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
        assert [int(l.max()) + 1 for l in res.labels] == [x[1] for x in g["subspaces"]]
        digests = [hashlib.sha256(np.ascontiguousarray(l, dtype=np.int32).tobytes()).hexdigest() for l in res.labels]
        assert digests == [x[3] for x in g["subspaces"]]
        assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


@pytest.mark.parametrize("case", [0, 1])
def test_wide_dimensions_with_entropy_order_match_golden(case):
    # d = 127 with the levels in entropy order: the permuted-order driver's
    # dimension sets beyond 64 (tests/golden/make_golden.py wide_ent)
    import hashlib
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "wide_entropy.json")
    if not os.path.exists(path):
        pytest.skip("no oracle digests for the wide entropy-order cases")
    g = json.load(open(path))["sets"][case]
    order = g["entropy_order"]
    w = datagen.wide(g["n"], g["d"], cells=g["cells"])
    assert g["X_sha256"] == hashlib.sha256(w.X.tobytes()).hexdigest()
    X = torch.from_numpy(w.X).cuda()
    for mode in (_lib.MODE_EXACT, _lib.MODE_BATCHED):
        say("wide entropy", order, mode)
        res = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order, mode=mode)
        assert res.subspaces == [s for s, _, _, _ in g["subspaces"]]
        digests = [hashlib.sha256(np.ascontiguousarray(l, dtype=np.int32).tobytes()).hexdigest() for l in res.labels]
        assert digests == [x[3] for x in g["subspaces"]]
        assert res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


@pytest.mark.parametrize("order", [1, -1])
def test_degenerate_inputs_with_entropy_order(order):
    # the inputs of test_gpu_parity's degenerate cases with the permuted level order
    from tests.test_gpu_parity import degenerate_cases
    for name, (X, eps, F, num_obj, min_size, r, c) in sorted(degenerate_cases().items()):
        w = datagen.Workload(name, X, np.full(len(X), -1, np.int32), [], eps=eps, cells=c, F=F, num_obj=num_obj,
                             min_size=min_size, r=r, rectangular=True)
        say("degenerate entropy", name, order)
        check(w, order, _lib.MODE_EXACT)
        check(w, order, _lib.MODE_BATCHED)
