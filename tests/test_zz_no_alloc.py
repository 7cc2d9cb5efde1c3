"""The allocation-free neighbour counts (inscy_neighbour_counts_noalloc)
against inscy_neighbour_counts on the same inputs: two calls in a row, the
second on a smaller data set, on a caller stream; ids outside [0, n) give -1;
inputs beyond the scratch array are refused before any device call."""

import ctypes
import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen

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


def _counts_both(X, S, eps, qs, stream):
    n, d = X.shape
    L = _lib.lib()
    dX = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    dq = torch.from_numpy(qs).cuda()
    dims = (ctypes.c_int32 * len(S))(*S)
    ref = torch.full((len(qs),), -7, dtype=torch.int32, device="cuda")
    _lib.check(L.inscy_neighbour_counts(dX.data_ptr(), n, d, eps, dims, len(S), dq.data_ptr(), len(qs),
                                        ref.data_ptr()))
    torch.cuda.synchronize()
    got = torch.full((len(qs),), -7, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(stream):
        _lib.check(L.inscy_neighbour_counts_noalloc(dX.data_ptr(), n, d, eps, dims, len(S), dq.data_ptr(), len(qs),
                                                    got.data_ptr(), ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_counts_equal_the_allocating_entry_point():
    stream = torch.cuda.Stream()
    # first call: 20,000 points, three dimensions of six
    w = datagen.small(31, n=20000, d=6, n_clusters=4, eps=0.02, sigma=0.04)
    rng = np.random.default_rng(31)
    qs = np.concatenate([np.arange(w.n), rng.integers(0, w.n, 5000)]).astype(np.int32)
    say("noalloc call 1: n", w.n, "dims [0, 2, 3]")
    ref, got = _counts_both(w.X, [0, 2, 3], 0.02, qs, stream)
    assert (ref >= 0).all() and ref.max() > 0
    assert np.array_equal(ref, got)
    # second call, smaller: 3,000 points, one dimension, another eps; the array
    # keeps the first call's contents, which the second must not read
    w2 = datagen.small(32, n=3000, d=4, n_clusters=2, eps=0.05, sigma=0.05)
    qs2 = np.random.default_rng(32).permutation(w2.n)[:1500].astype(np.int32)
    say("noalloc call 2: n", w2.n, "dims [1]")
    ref2, got2 = _counts_both(w2.X, [1], 0.05, qs2, stream)
    assert np.array_equal(ref2, got2)
    # and a third, larger again with all dimensions
    say("noalloc call 3: n", w.n, "dims 0..5")
    ref3, got3 = _counts_both(w.X, [0, 1, 2, 3, 4, 5], 0.03, qs[:4000], stream)
    assert np.array_equal(ref3, got3)


def test_noalloc_query_ids_outside_the_data_give_minus_one():
    w = datagen.small(33, n=2000, d=3, n_clusters=2, eps=0.05)
    qs = np.array([0, -1, 5, 2000, 1999, 1 << 30], dtype=np.int32)
    say("noalloc bad ids")
    stream = torch.cuda.Stream()
    n, d = w.X.shape
    L = _lib.lib()
    dX = torch.from_numpy(w.X).cuda()
    dq = torch.from_numpy(qs).cuda()
    dims = (ctypes.c_int32 * 2)(0, 2)
    out = torch.zeros(len(qs), dtype=torch.int32, device="cuda")
    _lib.check(L.inscy_neighbour_counts_noalloc(dX.data_ptr(), n, d, 0.05, dims, 2, dq.data_ptr(), len(qs),
                                                out.data_ptr(), ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    o = out.cpu().numpy()
    assert o[1] == -1 and o[3] == -1 and o[5] == -1
    assert (o[[0, 2, 4]] >= 0).all()
