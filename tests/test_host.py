"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol of include/inscy.h, argument errors come back as codes without
touching the device, the generator is deterministic, and the drop-in module
mirrors the reference's Python surface (python/inscy.py)."""

import ctypes
import os
import re

import numpy as np
import pytest

from gpu_inscy_amd import _lib, datagen
from gpu_inscy_amd import inscy as INSCY

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "inscy.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(inscy_[a-z_]+)\s*\(", src)))


def test_library_exports_every_header_function():
    lib = _lib.lib()
    names = header_functions()
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTED) <= set(names)


def test_argument_errors_are_codes_not_crashes():
    L = _lib.lib()
    h = ctypes.c_void_p()
    X = np.zeros((4, 2), dtype=np.float32)
    p = _lib.make_params(0.1, 1.0, 2, 1, 1.0, 3, True, 0)
    assert L.inscy_run(X.ctypes.data, 4, 2, None, 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    assert b"params" in L.inscy_last_error()
    assert L.inscy_run(X.ctypes.data, 0, 2, ctypes.byref(p), 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    # n above 2^30 - 1 is refused before any device call (the walks' 32-bit
    # column offsets hold 4 n bytes)
    assert L.inscy_run(X.ctypes.data, 1 << 30, 2, ctypes.byref(p), 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    assert b"n out of range" in L.inscy_last_error()
    assert L.inscy_run_device(X.ctypes.data, 1 << 30, 2, ctypes.byref(p), 2, None, ctypes.byref(h)) == \
        _lib.INSCY_ERR_ARGUMENT
    # d up to 128 (dimension sets are 128-bit masks)
    assert L.inscy_run(X.ctypes.data, 4, 129, ctypes.byref(p), 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    assert b"[1, 128]" in L.inscy_last_error()
    bad = _lib.make_params(0.0, 1.0, 2, 1, 1.0, 3, True, 0)
    assert L.inscy_run(X.ctypes.data, 4, 2, ctypes.byref(bad), 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    cells = _lib.make_params(0.1, 1.0, 2, 1, 1.0, 0, True, 0)
    assert L.inscy_run(X.ctypes.data, 4, 2, ctypes.byref(cells), 1, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    # a permuted level order runs in the exact and batched modes only
    ent = _lib.make_params(0.1, 1.0, 2, 1, 1.0, 3, False, 1)
    assert L.inscy_run(X.ctypes.data, 4, 2, ctypes.byref(ent), 4, ctypes.byref(h)) == _lib.INSCY_ERR_UNSUPPORTED
    ent = _lib.make_params(0.1, 1.0, 2, 1, 1.0, 3, True, -1)
    assert L.inscy_run(X.ctypes.data, 4, 2, ctypes.byref(ent), 3, ctypes.byref(h)) == _lib.INSCY_ERR_UNSUPPORTED
    assert L.inscy_run(X.ctypes.data, 4, 2, ctypes.byref(p), 7, ctypes.byref(h)) == _lib.INSCY_ERR_ARGUMENT
    assert not h.value
    with pytest.raises(_lib.InscyError):
        _lib.check(_lib.INSCY_ERR_ARGUMENT)


def test_flag_compaction_hook_rejects_bad_arguments_before_any_device_call():
    assert _lib.debug_flag_compact_temp(0) < 0
    assert _lib.debug_flag_compact_temp(1 << 31) < 0
    with pytest.raises(_lib.InscyError):
        _lib.debug_flag_compact(0, 16, 0, 0, 0, 0)  # null flags
    with pytest.raises(_lib.InscyError):
        _lib.debug_flag_compact(16, 0, 16, 1 << 20, 16, 16)  # no flags


def test_run_many_rejects_bad_arguments_before_any_device_call():
    L = _lib.lib()
    X = np.zeros((4, 2), dtype=np.float32)
    p = _lib.make_params(0.1, 1.0, 2, 1, 1.0, 3, True, 0)
    ptrs = (ctypes.c_void_p * 1)(X.ctypes.data)
    n = (ctypes.c_int64 * 1)(4)
    d = (ctypes.c_int32 * 1)(2)
    ps = (_lib.Params * 1)(p)
    out = (ctypes.c_void_p * 1)()
    assert L.inscy_run_many(ptrs, n, d, ps, 0, 2, out) == _lib.INSCY_ERR_ARGUMENT
    assert L.inscy_run_many(None, n, d, ps, 1, 2, out) == _lib.INSCY_ERR_ARGUMENT
    d0 = (ctypes.c_int32 * 1)(129)
    assert L.inscy_run_many(ptrs, n, d0, ps, 1, 2, out) == _lib.INSCY_ERR_ARGUMENT
    assert b"data set 0" in L.inscy_last_error()
    # the learned pruner belongs to the identity level order
    ent = (_lib.Params * 1)(_lib.make_params(0.1, 1.0, 2, 1, 1.0, 3, True, 1))
    assert L.inscy_run_many(ptrs, n, d, ent, 1, _lib.MODE_LEARNED, out) == _lib.INSCY_ERR_UNSUPPORTED


def test_noalloc_counts_refuse_what_the_scratch_cannot_hold_before_any_device_call():
    L = _lib.lib()
    dims = (ctypes.c_int32 * 8)(*range(8))
    X = np.zeros((4, 8), dtype=np.float32)
    # (12 k + 16) n bytes at k = 8, n = 2^24: far beyond the 64 MiB array
    rc = L.inscy_neighbour_counts_noalloc(X.ctypes.data, 1 << 24, 8, 0.01, dims, 8, X.ctypes.data, 1, X.ctypes.data,
                                          None)
    assert rc == _lib.INSCY_ERR_ARGUMENT and b"scratch" in L.inscy_last_error()
    assert L.inscy_neighbour_counts_noalloc(X.ctypes.data, 4, 8, 0.0, dims, 8, X.ctypes.data, 1, X.ctypes.data,
                                            None) == _lib.INSCY_ERR_ARGUMENT
    assert L.inscy_neighbour_counts_noalloc(X.ctypes.data, 4, 8, 0.01, dims, 9, X.ctypes.data, 1, X.ctypes.data,
                                            None) == _lib.INSCY_ERR_ARGUMENT
    bad = (ctypes.c_int32 * 2)(3, 1)  # not ascending
    assert L.inscy_neighbour_counts_noalloc(X.ctypes.data, 4, 8, 0.01, bad, 2, X.ctypes.data, 1, X.ctypes.data,
                                            None) == _lib.INSCY_ERR_ARGUMENT


def test_result_accessors_tolerate_null():
    L = _lib.lib()
    assert L.inscy_result_num_subspaces(None) == 0
    assert L.inscy_result_subspace(None, 0, None) == -1
    assert not L.inscy_result_labels(None, 0)
    assert L.inscy_result_num_points(None) == 0
    L.inscy_result_free(None)


def test_generator_is_deterministic_and_normalised():
    a = datagen.config(0)
    b = datagen.config(0)
    assert np.array_equal(a.X, b.X) and a.X.dtype == np.float32
    assert a.X.shape == (1500, 5)
    assert np.allclose(a.X.min(axis=0), 0) and np.allclose(a.X.max(axis=0), 1)
    assert (a.truth >= 0).sum() == 1350 and len(a.subspaces) == 4
    c = datagen.config(0, repeat=1)
    assert not np.array_equal(a.X, c.X)


def test_real_data_loader_follows_run_real():
    # experiments/run_real.py: glass drops its id and class columns,
    # pendigits its class column, vowel keeps all; values normalised to [0, 1]
    path = os.path.join(ROOT, "tests", "golden", "real", "real_inputs.npz")
    shapes = {"glass": (214, 9), "vowel": (990, 10), "pendigits": (7494, 16), "sky_0.5x0.5": (7253, 17),
              "sky_1x1": (7253 + 7598 + 7329 + 7447, 17)}
    for name, shape in shapes.items():
        w = datagen.real(name, path)
        assert w.X.shape == shape and w.X.dtype == np.float32
        assert w.X.min() == 0.0 and w.X.max() == 1.0
        assert w.min_size == int(0.05 * shape[0]) and w.cells == 4 and w.num_obj == 8 and w.eps == 0.01


def test_normalize_matches_reference_formula():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(50, 3)).astype(np.float32)
    want = (x - x.min(0, keepdims=True)) / (x.max(0, keepdims=True) - x.min(0, keepdims=True))
    assert np.array_equal(datagen.normalize(x), want.astype(np.float32))
    torch = pytest.importorskip("torch")
    t = INSCY.normalize(torch.from_numpy(x))
    assert np.array_equal(t.numpy(), want.astype(np.float32))


def test_drop_in_surface_mirrors_reference_names():
    for name in ["CPU", "GPU", "GPU5", "GPU3", "GPU_star", "GPU_memory", "GPU_memory2", "run_cpu", "run_cpu_weak",
                 "run_gpu", "run_gpu_weak", "run_gpu_reduced", "run_gpu_multi", "run_gpu_multi2",
                 "run_gpu_multi2_cl_all", "run_gpu_multi2_cl_re_all", "run_gpu_multi2_cl_multi",
                 "run_gpu_multi2_cl_multi_mem", "run_cpu_gpu_mix", "run_cpu_gpu_mix_cl_steam", "run_gpu_stream",
                 "normalize", "clean_up", "count_number_of_clusters", "load_synt"]:
        assert callable(getattr(INSCY, name)), name
    import inspect
    sig = inspect.signature(INSCY.CPU)
    assert list(sig.parameters) == ["X", "neighborhood_size", "F", "num_obj", "min_size", "r", "number_of_cells",
                                    "rectangular", "entropy_order"]
    assert sig.parameters["r"].default == 1.0 and sig.parameters["number_of_cells"].default == 3
    assert sig.parameters["rectangular"].default is False and sig.parameters["entropy_order"].default == 0


def test_count_number_of_clusters_and_clean_up():
    subs = [[0], [1], [0, 1]]
    cl = [[0, 0, 1, -1, -2], [-1, -1, -1, -1, -1], [3, 3, 3, 3, 5]]
    assert INSCY.count_number_of_clusters(subs, cl) == 4
    s2, c2 = INSCY.clean_up(subs, cl, 2)
    assert [list(x) for x in s2] == [[0], [0, 1]]
    assert c2[0].tolist() == [0, 0, -1, -1, -2]
    assert c2[1].tolist() == [3, 3, 3, 3, -1]


def test_experiment_grids_follow_the_reference_harness():
    from gpu_inscy_amd import experiments
    name, p = experiments.grid("n", 4)
    assert name == "_n" and p["n"] == [500, 1000, 2000, 4000, 8000] and p["d"] == [15] and p["real_cls"] == [4]
    name, p = experiments.grid("d", 10, None, True)
    assert name == "_d_large" and p["n"] == [25000] and p["d"] == [5 * i for i in range(1, 21)]
    name, p = experiments.grid("F", 4, "c")
    assert name == "_F_c" and p["F"] == [.5, 1., 1.5, 2., 2.5] and p["c"] == [2, 4, 6, 8, 10]
    assert experiments.run_file("GPU_INSCY", 1500, 15, 4, 0.01, 1.0, 1.0, 8, 0.05, 0, 4) == \
        "GPU_INSCYn1500d15c4N_size0.01F1.0r1.0num_obj8min_size0.05order0.npz"
    assert experiments.run_file("INSCY", 1500, 15, 4, 0.01, 1.0, 1.0, 8, 0.05, 0, 10).endswith("order0cl10.npz")


def test_tuning_rejects_bad_arguments_and_applies_saved_files(tmp_path):
    # inscy_set_tuning touches no device: arguments are checked, 0 restores the defaults
    import json
    from gpu_inscy_amd import autotune
    for mb, wk in ((-1, 0), (0, -1), (65537, 0), (0, 65)):
        with pytest.raises(_lib.InscyError) as e:
            _lib.set_tuning(mb, wk)
        assert e.value.code == _lib.INSCY_ERR_ARGUMENT
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"best": {"max_batch": 512, "workers": 2}, "results": []}))
    try:
        assert autotune.apply(str(path)) == {"max_batch": 512, "workers": 2}
    finally:
        _lib.set_tuning(0, 0)
