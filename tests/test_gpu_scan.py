"""The engine's own exclusive prefix sum and stable radix sort (the passes
that replace library scans: grid-index offsets, query compaction, label
emission, the leaf order of the phi density) against numpy, through the test
hooks inscy_debug_exclusive_sum / inscy_debug_sort_pairs.  Sizes straddle the
scan tile (4096 items), its single-tile path, and the sort tile (2048)."""

import ctypes
import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib

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


@pytest.mark.parametrize("items", [1, 7, 4095, 4096, 4097, 8192 + 3, 100000, 4096 * 1025 + 17, 5_000_001])
def test_exclusive_sum_matches_numpy(items):
    say("scan", items)
    L = _lib.lib()
    rng = np.random.default_rng(items)
    a = rng.integers(0, 300, items).astype(np.int32)
    a[rng.integers(0, items, max(1, items // 10))] = 0
    dev = torch.from_numpy(a).cuda()
    out = torch.full((items,), -5, dtype=torch.int32, device="cuda")
    tb = int(L.inscy_debug_scan_temp(items))
    assert tb > 0
    temp = torch.empty(tb, dtype=torch.uint8, device="cuda")
    _lib.check(L.inscy_debug_exclusive_sum(dev.data_ptr(), out.data_ptr(), items, temp.data_ptr(), tb, None))
    torch.cuda.synchronize()
    want = np.concatenate([[0], np.cumsum(a, dtype=np.int64)[:-1]]).astype(np.int32)
    assert np.array_equal(out.cpu().numpy(), want)
    # in place, with garbage left in the scratch by the previous call
    _lib.check(L.inscy_debug_exclusive_sum(dev.data_ptr(), dev.data_ptr(), items, temp.data_ptr(), tb, None))
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), want)


@pytest.mark.parametrize("n,bits", [(1, 8), (2047, 7), (2048, 8), (2049, 12), (100003, 16), (65536, 9),
                                    (300000, 6)])
def test_stable_sort_pairs_matches_numpy(n, bits):
    say("sort", n, bits)
    L = _lib.lib()
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << bits, n).astype(np.uint16)
    keys[: n // 3] = keys[0]  # long runs of equal keys: stability matters
    vals = rng.permutation(n).astype(np.int32)
    dk = torch.from_numpy(keys.view(np.int16)).cuda()
    dv = torch.from_numpy(vals).cuda()
    kt = torch.empty_like(dk)
    ko = torch.empty_like(dk)
    vt = torch.empty_like(dv)
    vo = torch.empty_like(dv)
    tb = int(L.inscy_debug_sort_temp(n))
    temp = torch.empty(tb, dtype=torch.uint8, device="cuda")
    _lib.check(L.inscy_debug_sort_pairs(dk.data_ptr(), kt.data_ptr(), ko.data_ptr(), dv.data_ptr(), vt.data_ptr(),
                                        vo.data_ptr(), n, bits, temp.data_ptr(), tb, None))
    torch.cuda.synchronize()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ko.cpu().numpy().view(np.uint16), keys[order])
    assert np.array_equal(vo.cpu().numpy(), vals[order])
