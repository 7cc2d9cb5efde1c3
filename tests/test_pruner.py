"""The learned pruner's host side and its FP64 oracle, on the CPU: the oracle's
gradient against an independent float64 autograd and against central finite
differences, bf16 rounding, threshold choice and the cluster F1 measure."""

import types

import numpy as np
import pytest

from gpu_inscy_amd import _lib, pruner
from oracle import oracle as O


def _problem(seed, N=40):
    rng = np.random.default_rng(seed)
    P = pruner.bf16_round(pruner.init_params(seed))
    X = rng.random((N, _lib.MLP_INPUTS), dtype=np.float32)
    Y = (rng.random(N) < 0.4).astype(np.float32)
    W = rng.uniform(0.5, 2.0, N).astype(np.float32)
    return P, X, Y, W


def test_oracle_gradient_matches_float64_autograd():
    torch = pytest.importorskip("torch")
    P, X, Y, W = _problem(0)
    g, loss = O.mlp_grad(P, X, Y, W, 1.0 / len(Y))
    p = torch.tensor(P.astype(np.float64), requires_grad=True)
    H, I = _lib.MLP_HIDDEN, _lib.MLP_INPUTS
    o = 0
    W1 = p[o:o + H * I].view(H, I); o += H * I
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].view(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    x = torch.tensor(X.astype(np.float64))
    z = torch.relu(torch.relu(x @ W1.T + b1) @ W2.T + b2) @ w3 + b3
    ref = (torch.nn.functional.binary_cross_entropy_with_logits(
        z, torch.tensor(Y, dtype=torch.float64), reduction="none") * torch.tensor(W, dtype=torch.float64)).mean()
    ref.backward()
    assert abs(ref.item() - loss) < 1e-12
    np.testing.assert_allclose(g, p.grad.numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.mlp_forward(P, X), z.detach().numpy(), rtol=0, atol=1e-12)


def test_oracle_gradient_matches_finite_differences():
    P, X, Y, W = _problem(1, N=8)
    g, _ = O.mlp_grad(P, X, Y, W, 1.0 / len(Y))
    rng = np.random.default_rng(2)
    # one parameter of every block: W1, b1, W2, b2, w3, b3
    idx = [int(rng.integers(0, 8192)), 8192 + 5, 8320 + int(rng.integers(0, 16384)), 24704 + 7, 24832 + 9, 24960]
    for i in idx:
        h = 2.0 ** -12
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += np.float32(h)
        Pm[i] -= np.float32(h)
        fd = (O.mlp_grad(Pp, X, Y, W, 1.0 / len(Y))[1] - O.mlp_grad(Pm, X, Y, W, 1.0 / len(Y))[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(pruner.bf16_round(x), want)


def test_choose_threshold_keeps_the_requested_recall():
    prob = np.array([0.1, 0.2, 0.3, 0.4, 0.9, 0.05], dtype=np.float32)
    y = np.array([1, 1, 1, 1, 0, 0])
    assert pruner.choose_threshold(prob, y, 1.0) == np.float32(0.1)
    assert pruner.choose_threshold(prob, y, 0.75) == np.float32(0.2)
    assert pruner.choose_threshold(prob, np.zeros(6), 0.999) == 0.0


def _res(subspaces, labels):
    return types.SimpleNamespace(subspaces=subspaces, labels=np.array(labels, dtype=np.int32).reshape(len(subspaces), -1))


def test_cluster_f1():
    a = _res([[0], [0, 1]], [[0, 0, 1, 1, -1, -1], [0, 0, 0, -1, -1, -1]])
    assert pruner.cluster_f1(a, a)["f1"] == 1.0
    # the same partition under other numbers is the same clustering
    b = _res([[0], [0, 1]], [[1, 1, 0, 0, -1, -1], [0, 0, 0, -1, -1, -1]])
    assert pruner.cluster_f1(a, b)["f1"] == 1.0
    # one subspace missing: recall 2/3 (its cluster scores 0), precision 1
    c = _res([[0]], [[0, 0, 1, 1, -1, -1]])
    f = pruner.cluster_f1(a, c)
    assert f["recall"] == pytest.approx(2 / 3) and f["precision"] == 1.0
    assert f["f1"] == pytest.approx(0.8)
    # a cluster losing one of its three points: F1 of that match 2*2/(3+2)
    d = _res([[0], [0, 1]], [[0, 0, 1, 1, -1, -1], [0, 0, -1, -1, -1, -1]])
    f = pruner.cluster_f1(a, d)
    assert f["recall"] == pytest.approx((1 + 1 + 0.8) / 3) and f["precision"] == pytest.approx((1 + 1 + 0.8) / 3)
