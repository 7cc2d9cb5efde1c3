"""The learned pruner on the GPU: the network's kernels against the FP64
oracle, the collected training pairs against what the exact result fixes, and
the learned mode's limits (threshold 0 is exact; skipping everything leaves an
empty result)."""

import os
import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen, pruner
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
    yield
    _lib.set_pruner(None, 0.0)


def _problem(seed, N):
    rng = np.random.default_rng(seed)
    P = pruner.bf16_round(pruner.init_params(seed) + rng.normal(0, 0.01, _lib.MLP_PARAMS).astype(np.float32))
    X = rng.random((N, _lib.MLP_INPUTS), dtype=np.float32)
    Y = (rng.random(N) < 0.3).astype(np.float32)
    W = rng.uniform(0.5, 2.0, N).astype(np.float32)
    return P, X, Y, W


@pytest.mark.parametrize("N", [1, 31, 1000])
def test_mlp_forward_matches_fp64(N):
    say("mlp forward", N)
    P, X, _, _ = _problem(1, N)
    tr = pruner.Trainer(P)
    prob = tr.forward(X).cpu().numpy().astype(np.float64)
    ref = 1.0 / (1.0 + np.exp(-O.mlp_forward(P, X)))
    # fp32 accumulation of at most 129 bf16 * fp32 terms per unit
    np.testing.assert_allclose(prob, ref, rtol=0, atol=2e-5)


@pytest.mark.parametrize("N,nblocks", [(7, 4), (1000, 16), (5000, 128)])
def test_mlp_gradient_matches_fp64(N, nblocks):
    say("mlp gradient", N, nblocks)
    P, X, Y, W = _problem(2, N)
    tr = pruner.Trainer(P, nblocks=nblocks)
    dev = lambda a: torch.from_numpy(a).cuda()
    g = tr.gradient(dev(X), dev(Y), dev(W), 1.0 / N).cpu().numpy().astype(np.float64)
    ref, loss = O.mlp_grad(P, X, Y, W, 1.0 / N)
    assert abs(g[-1] - loss) <= 1e-5 * max(1.0, abs(loss))
    err = np.linalg.norm(g[:-1] - ref) / np.linalg.norm(ref)
    assert err < 2e-5, err
    np.testing.assert_allclose(g[:-1], ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    # fixed summation order: a second call is bit-identical
    g2 = tr.gradient(dev(X), dev(Y), dev(W), 1.0 / N).cpu().numpy()
    assert np.array_equal(g2, g.astype(np.float32))


def test_mlp_adam_step_and_bf16_copy():
    say("mlp adam")
    P, X, Y, W = _problem(3, 300)
    tr = pruner.Trainer(P)
    dev = lambda a: torch.from_numpy(a).cuda()
    m = np.zeros(_lib.MLP_PARAMS, np.float64)
    v = np.zeros(_lib.MLP_PARAMS, np.float64)
    w = P.astype(np.float64)
    for step in range(1, 4):
        g = tr.gradient(dev(X), dev(Y), dev(W), 1.0 / 300).cpu().numpy()[:-1].astype(np.float64)
        tr.adam(1e-3)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        w = w - 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
        master = tr.master.cpu().numpy()
        np.testing.assert_allclose(master, w, rtol=1e-5, atol=1e-6)
        w = master.astype(np.float64)
        bf = tr.params.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
        assert np.array_equal(bf.view(np.float32), pruner.bf16_round(master))


def test_training_separates_a_learnable_problem():
    say("mlp training")
    rng = np.random.default_rng(4)
    X = rng.random((20000, _lib.MLP_INPUTS), dtype=np.float32)
    # a disc in two of the inputs: not linearly separable
    Y = ((X[:, 3] - 0.5) ** 2 + (X[:, 17] - 0.5) ** 2 < 0.1).astype(np.float32)
    tr = pruner.Trainer(pruner.init_params(4))
    losses = tr.fit(X[:16000], Y[:16000], np.ones(16000, np.float32), epochs=30, batch=512, lr=3e-3)
    assert losses[-1] < 0.3 * losses[0]
    acc = ((tr.forward(X[16000:]).cpu().numpy() >= 0.5) == (Y[16000:] > 0.5)).mean()
    assert acc > 0.95, acc


def _small():
    return datagen.small(11, n=1500, d=6, n_clusters=3, eps=0.03, cells=4, num_obj=4, min_size=40,
                         sub_dims=(2, 4))


def _run(w, mode):
    return INSCY.run(torch.from_numpy(w.X).cuda(), mode=mode, **w.params())


@pytest.mark.parametrize("case", ["small", "config0", "config1"])
def test_collect_is_exact_and_labels_follow_the_result(case):
    say("collect", case)
    w = _small() if case == "small" else datagen.config(int(case[-1]))
    ex = _run(w, _lib.MODE_BATCHED)
    col = _run(w, _lib.MODE_COLLECT)
    assert col.subspaces == ex.subspaces and np.array_equal(col.labels, ex.labels)
    P = col.pairs
    assert P is not None and len(P["labels"]) == col.stats["pairs_scored"]
    info, y, cost, F = P["info"], P["labels"].astype(bool), P["cost"], P["features"]
    assert np.isfinite(F).all() and (cost[:, 1] >= cost[:, 0]).all()
    # pairs of the root tree: positive exactly for the smallest dimension of
    # some result subspace (the only way down to it)
    root = info[:, 0] == 0
    assert sorted(info[root, 1].tolist()) == list(range(w.d))
    assert set(info[root & y, 1].tolist()) == {s[0] for s in ex.subspaces}
    # pairs of one-dimensional trees: positive only towards a result subspace's second dimension
    one = info[:, 0] == 1
    second = {s[1] for s in ex.subspaces if len(s) > 1}
    assert set(info[one & y, 1].tolist()) <= second
    # root features follow from the data alone
    n, d = w.X.shape
    for j in np.nonzero(root)[0]:
        m = info[j, 1]
        f = F[j]
        h = np.bincount(np.minimum((w.X[:, m] * np.float32(16)).astype(np.int64), 15), minlength=16) / n
        np.testing.assert_allclose(f[8:24], np.sort(h)[::-1], atol=1e-6)
        np.testing.assert_allclose(f[[0, 3, 5, 6, 7, 31]], [1.0, 0.0, m / d, 1.0, 1.0, 0.0], atol=1e-6)
        assert f[32 + m % 32] == 1.0 and f[32:].sum() == 1.0


def test_learned_mode_limits():
    say("learned limits")
    w = datagen.config(1)
    ex = _run(w, _lib.MODE_BATCHED)
    rng = np.random.default_rng(5)
    W = pruner.init_params(5)
    # threshold 0 skips nothing: the exact result
    _lib.set_pruner(W, 0.0, 0)
    r0 = _run(w, _lib.MODE_LEARNED)
    assert r0.subspaces == ex.subspaces and np.array_equal(r0.labels, ex.labels)
    assert r0.stats["pairs_skipped"] == 0 and r0.stats["pairs_scored"] > 0
    # every pair skipped: nothing is counted
    _lib.set_pruner(W, 2.0, 0)
    r1 = _run(w, _lib.MODE_LEARNED)
    assert r1.subspaces == [] and r1.stats["pairs_skipped"] == r1.stats["pairs_scored"] == w.d
    # root pairs kept: one-dimensional subspaces only
    _lib.set_pruner(W, 2.0, 1)
    r2 = _run(w, _lib.MODE_LEARNED)
    assert all(len(s) == 1 for s in r2.subspaces)
    assert r2.stats["pairs_skipped"] == r2.stats["pairs_scored"] - w.d
    _lib.set_pruner(None, 0.0)
    with pytest.raises(_lib.InscyError):
        _run(w, _lib.MODE_LEARNED)


def test_trained_network_against_no_pruning():
    # the shipped network on a config-2 data set it was not trained on:
    # what it skips, what it costs in cluster F1, and the time of both runs
    if not os.path.exists(pruner.DEFAULT_WEIGHTS):
        pytest.skip("no trained network (python -m gpu_inscy_amd.train_pruner)")
    say("trained network")
    import time
    w = datagen.config(2, 12)  # trained on repeats 1-8, threshold from 9-11
    X = torch.from_numpy(w.X).cuda()
    torch.cuda.synchronize()
    t = time.perf_counter()
    ex = INSCY.run(X, mode=_lib.MODE_BATCHED, **w.params())
    t_exact = time.perf_counter() - t
    pruner.use()
    t = time.perf_counter()
    le = INSCY.run(X, mode=_lib.MODE_LEARNED, **w.params())
    t_learned = time.perf_counter() - t
    f = pruner.cluster_f1(ex, le)
    say(f"exact {t_exact:.3f}s learned {t_learned:.3f}s f1 {f['f1']:.5f} recall {f['recall']:.5f} "
        f"pairs skipped {le.stats['pairs_skipped']}/{le.stats['pairs_scored']} "
        f"trees {le.stats['trees']}/{ex.stats['trees']}")
    assert le.stats["pairs_skipped"] > 0 and le.stats["trees"] < ex.stats["trees"]
    assert f["f1"] >= 0.99
    _lib.set_pruner(None, 0.0)


def test_learned_mode_on_the_evaluation_set_keeps_f1_and_skips_work():
    # n=200,000 d=32, 30 clusters (the learned-prune evaluation set): the
    # shipped network against the exact level-synchronous run of the same data
    if not os.path.exists(pruner.DEFAULT_WEIGHTS):
        pytest.skip("no trained network (python -m gpu_inscy_amd.train_pruner)")
    import time
    say("learned on the n=200k evaluation set")
    w = datagen.config(4)
    X = torch.from_numpy(w.X).cuda()
    torch.cuda.synchronize()
    t = time.perf_counter()
    ex = INSCY.run(X, mode=_lib.MODE_BATCHED, **w.params())
    t_exact = time.perf_counter() - t
    pruner.use()
    t = time.perf_counter()
    le = INSCY.run(X, mode=_lib.MODE_LEARNED, **w.params())
    t_learned = time.perf_counter() - t
    f = pruner.cluster_f1(ex, le)
    say(f"exact {t_exact:.3f}s learned {t_learned:.3f}s ({t_exact / t_learned:.3f}x) f1 {f['f1']:.5f} "
        f"trees {le.stats['trees']}/{ex.stats['trees']} pairs skipped {le.stats['pairs_skipped']}")
    assert f["f1"] >= 0.99
    assert le.stats["pairs_skipped"] > 0 and le.stats["trees"] < ex.stats["trees"]
    assert le.stats["queries"] < ex.stats["queries"]
    _lib.set_pruner(None, 0.0)


def test_pruning_bound_from_the_exact_pair_labels():
    # What any pruner of (tree, dimension) pairs can save on the evaluation
    # set: a pair whose subtree holds a cluster of the exact result (label 1)
    # must be kept for F1 = 1, so the neighbour counts of the positive pairs
    # stay.  Their share of all counted points bounds the counting speed-up of
    # a perfect pruner; with it the 3x target is out of reach on this data.
    say("pruning bound (collect on the n=200k evaluation set)")
    w = datagen.config(4)
    X = torch.from_numpy(w.X).cuda()
    col = INSCY.run(X, mode=_lib.MODE_COLLECT, **w.params())
    P = col.pairs
    y = P["labels"].astype(bool)
    own = P["cost"][:, 0].astype(np.float64)   # points counted for the pair itself
    total = own.sum()
    kept = own[y].sum()
    bound = total / kept
    say(f"pairs {len(y)} positive {int(y.sum())} count work of positive pairs {kept / total:.4f} "
        f"-> perfect-pruner counting speed-up <= {bound:.3f}x; result subspaces {len(col.subspaces)}")
    assert y.any() and total > 0
    # the counts of pairs that reach the result are most of the work
    assert kept / total > 0.5
    assert bound < 3.0
