"""The drop-in surface beyond the core run: run_cmp's tensor form, the
learned-mode wrapper and the experiment harness (experiments/run.py's
records), on the GPU."""

import os
import sys

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen, experiments
from gpu_inscy_amd import inscy as INSCY

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    _lib.lib()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_run_cmp_is_the_exact_result_as_tensors():
    say("run_cmp")
    w = datagen.config(0)
    masks, labels = INSCY.run_cmp(torch.from_numpy(w.X), w.eps, w.F, w.num_obj, w.min_size, w.cells)
    ref = INSCY.run(torch.from_numpy(w.X), w.eps, w.F, w.num_obj, w.min_size, 0.5, w.cells, False, 0)
    assert masks.dtype == torch.int32 and labels.dtype == torch.int32
    assert [list(np.nonzero(m)[0]) for m in masks.numpy()] == ref.subspaces
    assert np.array_equal(labels.numpy(), ref.labels)


def test_experiment_harness_records(tmp_path):
    say("experiments")
    done = experiments.main(["GPU-INSCY", "num_obj", "1", "4", "--out", str(tmp_path)])
    assert len(done) == len(experiments.SWEEPS["num_obj"])
    for path in done:
        with np.load(path, allow_pickle=False) as z:
            assert z["times"].shape == (1,) and z["no_clusters"].shape == (1,)
            sub, lab = z["subspaces_0"], z["clusterings_0"]
            assert sub.shape[0] == lab.shape[0] and lab.shape[1] == 1500 and sub.shape[1] == 15
            assert z["no_clusters"][0] == INSCY.count_number_of_clusters(None, lab)
    # a second call finds every record and runs nothing
    assert experiments.main(["GPU-INSCY", "num_obj", "1", "4", "--out", str(tmp_path)]) == []


def test_gpu_learned_wrapper():
    say("GPU_learned")
    w = datagen.config(1)
    subs, labels = INSCY.GPU_learned(torch.from_numpy(w.X), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True)
    assert len(subs) == len(labels) > 0 and all(len(l) == w.n for l in labels)
    _lib.set_pruner(None, 0.0)
    INSCY.GPU_learned._loaded = False


def test_autotune_settings_do_not_change_results():
    # the tuner checks every setting's result against the first one; one
    # child subspace per batch against the default bounds exercise the
    # batching, and the best setting is applied
    from gpu_inscy_amd import autotune
    say("autotune config 0")
    try:
        rec = autotune.tune(0, runs=1, batches=(1, 3, 1024), workers=(0,))
        assert [r["max_batch"] for r in rec["results"]] == [1, 3, 1024]
        assert rec["best"]["max_batch"] in (1, 3, 1024)
        rec3 = autotune.tune(0, phi=True, runs=1, batches=(2, 1024), workers=(0,))
        assert len(rec3["results"]) == 2
    finally:
        _lib.set_tuning(0, 0)
