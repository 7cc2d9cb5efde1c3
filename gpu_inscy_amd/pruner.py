"""The learned pruner: training pairs from exact runs, the network's training
on the GPU with the engine's own kernels, and cluster F1 against exact.

The unit the pruner decides on is a (restricted tree T, dimension m) pair of
the level-batched driver: skipping it saves the neighbour counts of T's points
in S + {m} and everything below the restricted trees that pair would create.
An INSCY_MODE_COLLECT run labels every pair of an exact run (1 when a tree
below it holds a cluster of the exact result) with its 64 features.

Training (Trainer): minibatch Adam on the weighted binary cross-entropy, with
inscy_mlp_gradient / inscy_mlp_adam on device arrays; torch only carries the
arrays.  The threshold is then chosen on held-out pairs so that the positive
pairs keep a target recall, and learned-mode runs are judged by cluster F1
against the exact result (cluster_f1).
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "pruner_weights.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform initial weights, zero biases (float32 [MLP_PARAMS])."""
    rng = np.random.default_rng(seed)
    H, I = _lib.MLP_HIDDEN, _lib.MLP_INPUTS
    parts = [rng.uniform(-1, 1, H * I) * np.sqrt(6.0 / I), np.zeros(H),
             rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H), np.zeros(H),
             rng.uniform(-1, 1, H) * np.sqrt(3.0 / H), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def bf16_round(w: np.ndarray) -> np.ndarray:
    """float32 -> bf16 (round to nearest even) -> float32, as the engine stores weights."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class Trainer:
    """The network on one GPU: float32 master weights, bf16 copy, Adam state.

    features [N, 64] float32, labels [N] (0/1), weights [N] float32 may be
    numpy arrays or torch tensors; they are moved to `device` once."""

    def __init__(self, params: np.ndarray, device="cuda", nblocks: int = 128):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.master = torch.tensor(np.ascontiguousarray(params, dtype=np.float32), device=self.device)
        self.params = torch.zeros(_lib.MLP_PARAMS, dtype=torch.int16, device=self.device)
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.grad = torch.zeros(_lib.MLP_PARAMS + 1, dtype=torch.float32, device=self.device)
        self.nblocks = int(nblocks)
        self.part = torch.zeros(self.nblocks * (_lib.MLP_PARAMS + 1), dtype=torch.float32, device=self.device)
        self.step_no = 0
        self._stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.lib().inscy_mlp_round(_ptr(self.master), _ptr(self.params), ctypes.c_void_p(self._stream)))

    def _dev(self, a, dtype):
        t = self.torch
        if isinstance(a, t.Tensor):
            return a.to(device=self.device, dtype=dtype).contiguous()
        return t.tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)

    def forward(self, features):
        """Probabilities of the rows of `features` (torch tensor on the device)."""
        X = self._dev(features, self.torch.float32)
        out = self.torch.empty(X.shape[0], dtype=self.torch.float32, device=self.device)
        _lib.check(_lib.lib().inscy_mlp_forward(_ptr(self.params), _ptr(X), X.shape[0], _ptr(out),
                                                ctypes.c_void_p(self._stream)))
        return out

    def gradient(self, X, Y, W, inv_wsum: float):
        """Gradient (device tensor [MLP_PARAMS + 1], last entry the loss) of
        inv_wsum * sum W_i BCE(logit_i, Y_i) at the current bf16 weights."""
        nb = max(1, min(self.nblocks, (X.shape[0] + 31) // 32))
        _lib.check(_lib.lib().inscy_mlp_gradient(_ptr(self.params), _ptr(X), _ptr(Y), _ptr(W), X.shape[0],
                                                 float(inv_wsum), _ptr(self.part), nb, _ptr(self.grad),
                                                 ctypes.c_void_p(self._stream)))
        return self.grad

    def adam(self, lr: float):
        self.step_no += 1
        _lib.check(_lib.lib().inscy_mlp_adam(_ptr(self.master), _ptr(self.grad), _ptr(self.m), _ptr(self.v),
                                             _ptr(self.params), self.step_no, float(lr),
                                             ctypes.c_void_p(self._stream)))

    def fit(self, features, labels, weights, epochs: int = 20, batch: int = 4096, lr: float = 2e-3,
            seed: int = 0, log=None):
        """Minibatch Adam over shuffled data; the learning rate decays by
        cosine to lr/20.  Returns the mean loss of each epoch."""
        t = self.torch
        X = self._dev(features, t.float32)
        Y = self._dev(labels, t.float32)
        W = self._dev(weights, t.float32)
        N = X.shape[0]
        gen = t.Generator(device="cpu").manual_seed(int(seed))
        steps_per_epoch = (N + batch - 1) // batch
        total = epochs * steps_per_epoch
        losses = []
        k = 0
        for ep in range(epochs):
            perm = t.randperm(N, generator=gen).to(self.device)
            Xs, Ys, Ws = X[perm].contiguous(), Y[perm].contiguous(), W[perm].contiguous()
            ep_loss = t.zeros((), dtype=t.float64, device=self.device)
            for b0 in range(0, N, batch):
                xb, yb, wb = Xs[b0:b0 + batch], Ys[b0:b0 + batch], Ws[b0:b0 + batch]
                g = self.gradient(xb, yb, wb, 1.0 / xb.shape[0])
                ep_loss += g[_lib.MLP_PARAMS].double()
                step_lr = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * k / max(total, 1))))
                self.adam(step_lr)
                k += 1
            losses.append(float(ep_loss.item()) / steps_per_epoch)
            if log:
                log(f"epoch {ep + 1}/{epochs} loss {losses[-1]:.5f}")
        t.cuda.synchronize(self.device)
        return losses

    def weights(self) -> np.ndarray:
        """The trained weights as the engine holds them (bf16 values, float32)."""
        return bf16_round(self.master.detach().cpu().numpy())


def features_weights(pairs: dict, pos_weight: float = 1.0) -> np.ndarray:
    """Per-pair loss weights: a negative pair weighs by the work its skip
    saves (log of the points counted below it), a positive pair by
    pos_weight times the mean negative weight."""
    cost = pairs["cost"][:, 1].astype(np.float64)
    y = pairs["labels"].astype(bool)
    w = np.log2(2.0 + cost)
    neg_mean = w[~y].mean() if (~y).any() else 1.0
    w = np.where(y, pos_weight * neg_mean, w)
    return (w / w.mean()).astype(np.float32)


def choose_threshold(prob: np.ndarray, labels: np.ndarray, recall: float = 0.999) -> float:
    """Largest threshold that keeps at least `recall` of the positive pairs
    (their scores at or above it)."""
    pos = np.sort(prob[labels.astype(bool)])
    if pos.size == 0:
        return 0.0
    k = int(np.floor((1.0 - recall) * pos.size))
    return float(pos[k])


def save(path: str, weights: np.ndarray, threshold: float, min_depth: int, meta: dict) -> None:
    np.savez(path, weights=np.asarray(weights, dtype=np.float32), threshold=np.float32(threshold),
             min_depth=np.int32(min_depth), meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def load(path: str = DEFAULT_WEIGHTS):
    """(weights, threshold, min_depth, meta) of a saved network."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}
        return z["weights"].astype(np.float32), float(z["threshold"]), int(z["min_depth"]), meta


def use(path: str = DEFAULT_WEIGHTS, threshold: float | None = None, min_depth: int | None = None) -> None:
    """Loads a saved network into the engine for learned-mode runs."""
    w, thr, md, _ = load(path)
    _lib.set_pruner(w, thr if threshold is None else threshold, md if min_depth is None else min_depth)


def cluster_f1(exact: _lib.Result, approx: _lib.Result) -> dict:
    """Cluster F1 of `approx` against `exact`.  A cluster is a (subspace,
    label) pair; each exact cluster is matched to the approximate cluster of
    the same subspace with the highest F1 of their point sets (0 if the
    subspace is missing), and vice versa.  recall = mean over exact clusters,
    precision = mean over approximate clusters, f1 = their harmonic mean."""
    ka = {tuple(s): i for i, s in enumerate(exact.subspaces)}
    kb = {tuple(s): i for i, s in enumerate(approx.subspaces)}
    rec, prec = [], []
    for s in set(ka) | set(kb):
        la = exact.labels[ka[s]] if s in ka else None
        lb = approx.labels[kb[s]] if s in kb else None
        na = int(la.max()) + 1 if la is not None and la.size else 0
        nb = int(lb.max()) + 1 if lb is not None and lb.size else 0
        if la is None or lb is None or na == 0 or nb == 0:
            rec += [0.0] * na
            prec += [0.0] * nb
            continue
        sa = np.bincount(la[la >= 0], minlength=na).astype(np.float64)
        sb = np.bincount(lb[lb >= 0], minlength=nb).astype(np.float64)
        both = (la >= 0) & (lb >= 0)
        inter = np.bincount(la[both].astype(np.int64) * nb + lb[both], minlength=na * nb)
        inter = inter.reshape(na, nb).astype(np.float64)
        f = 2.0 * inter / (sa[:, None] + sb[None, :])
        rec += list(f.max(axis=1))
        prec += list(f.max(axis=0))
    r = float(np.mean(rec)) if rec else 1.0
    p = float(np.mean(prec)) if prec else 1.0
    return {"f1": 2 * p * r / (p + r) if p + r > 0 else 0.0, "recall": r, "precision": p,
            "clusters_exact": len(rec), "clusters_approx": len(prec)}
