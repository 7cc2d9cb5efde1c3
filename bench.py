"""Benchmark: end-to-end INSCY on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode exact|batched|learned] [--config 2]

One step is one complete INSCY run (all subspaces, all clusters) over the
synthetic data of the chosen workload, already resident in HBM; the timed
region spans exactly K steps between device synchronisations.  The default
workload is n=100,000 d=20 (20 Zipf-sized clusters, eps=0.02, c=8, mu=4,
min_size=1000, rectangular).  The reported value is restricted (subspace,
region) trees evaluated per second, the throughput unit of the reference's
depth-first search; ms_per_step is the end-to-end time of one run (labels of
every result subspace back in host memory).  The default path is the
level-synchronous driver (`batched`, exact results); the depth-first driver
(`exact`, one restricted tree at a time) is timed beside it outside the timed
region and reported as depth_first_ms.

Also reported:
  parity        the last timed run equals the committed golden digests of the
                oracle (tests/golden/config<N>.json), when present.
  roofline      the neighbour-count kernel (the dominant one; k_b_count in the
                level-synchronous driver, k_grid_count in the depth-first one):
                its algorithmic bytes (per counted point 4k+4, per neighbour
                established 4+4k) over its device time measured with HIP events
                on the engine's stream in a separate pass of up to 3 runs after
                the timed steps (ms_per_step_with_events shows their overhead),
                against 8 TB/s HBM; `traffic` comes from the committed rocprofv3
                counter summary.
  cpu_baseline  the oracle (a single-threaded CPU restatement of the
                reference's CPU INSCY) on the same data for a bounded time
                budget, as trees per second, with the host's CPU model.
  dropin        the drop-in call GPU5() (host tensor in, Python lists out, as
                the reference's op returns) once, in seconds.
  learned       the learned-prune mode on the n=200,000 d=32 evaluation set
                against the exact run of the same data: seconds, speed-up and
                cluster F1.
Progress lines go to stderr; the one JSON line to stdout.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["exact", "batched", "learned"], default="batched",
                    help="learned: the learned pruner (gpu_inscy_amd/pruner_weights.npz) against the exact run")
    ap.add_argument("--depth-first-steps", type=int, default=1,
                    help="untimed runs of the depth-first driver for the comparison (0: skip)")
    ap.add_argument("--config", type=int, default=2,
                    help="workload index (2: n=100000 d=20, the default; 1: n=25000 d=15; 4: n=200000 d=32; "
                         "3: 64 data sets)")
    ap.add_argument("--no-dropin", dest="dropin", action="store_false",
                    help="skip the drop-in GPU5() call (host tensor in, Python lists out)")
    ap.add_argument("--learned-config", type=int, default=4,
                    help="workload of the learned-prune leg (0: skip)")
    ap.add_argument("--real", choices=["glass", "vowel", "pendigits", "sky_0.5x0.5", "sky_1x1", "sky_1x2"],
                    help="one of the reference's real-data workloads instead of --config")
    ap.add_argument("--threshold", type=float, default=None,
                    help="learned mode: pruner threshold instead of the trained network's own")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 3 times (calibrates the HBM byte counters)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--phi", action="store_true",
                    help="rectangular=False (the phi density, the wrappers' default) on the same workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def golden_parity(cfg, res, phi=False):
    path = os.path.join(ROOT, "tests", "golden", f"config{cfg}{'_phi' if phi else ''}.json")
    if not os.path.exists(path):
        return None
    import numpy as np
    g = json.load(open(path))
    if [s for s, _, _, _ in g["subspaces"]] != res.subspaces:
        return False
    for (_, _, _, dig), lab in zip(g["subspaces"], res.labels):
        if hashlib.sha256(np.ascontiguousarray(lab, dtype=np.int32).tobytes()).hexdigest() != dig:
            return False
    return res.stats["trees"] == g["trees"] and res.stats["calls"] == g["calls"]


def traffic_from_profile(kernel, cfg):
    path = os.path.join(ROOT, "profiles", f"{kernel}_config{cfg}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("hbm_bytes_per_dispatch")
    except (OSError, ValueError):
        return None


def host_cpu():
    """CPU model and hardware threads of this host (the CPU-baseline leg's machine)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: the engine measures "
              "one GPU's share of the work (independent subspace branches per GPU), with no "
              "collectives", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gpu_inscy_amd import _lib, datagen
    from gpu_inscy_amd import inscy as INSCY

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe > 0:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatches
        # give the scale from the counters to HBM bytes
        nb = args.copy_probe << 20
        a = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb, "copies": 3}), flush=True)
        return
    if args.mode == "learned":
        return bench_learned(args)
    mode = _lib.MODE_EXACT if args.mode == "exact" else _lib.MODE_BATCHED
    if args.config == 3:
        return bench_many(args, mode)
    if args.real:
        return bench_real(args, mode)
    w = datagen.config(args.config)
    if args.phi:
        import dataclasses
        w = dataclasses.replace(w, rectangular=False)
    X = torch.from_numpy(w.X).to("cuda:0")
    torch.cuda.synchronize()

    def step(m=mode):
        return INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order,
                         mode=m)

    res = None
    for _ in range(max(args.warmup, 0)):
        res = None
        res = step()
    say(f"warm-up done ({args.warmup} runs)")

    # timed region: K complete runs; only the last result is held, as a
    # caller that consumes each run would
    _lib.profile(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        res = None
        res = step()
        stats.append(res.stats)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    say(f"timed: {args.steps} runs, {ms_step:.3f} ms/run")
    parity = golden_parity(args.config, res, args.phi)  # the last timed run against the oracle's digests
    say(f"parity of the last timed run: {parity}")

    # the roofline pass: the same runs with the count kernel's launches
    # bracketed by HIP events on the engine's stream (inscy_profile)
    nprof = max(1, min(3, args.steps))
    pstats = []
    _lib.profile(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(nprof):
        r2 = step()
        pstats.append(r2.stats)
        del r2
    torch.cuda.synchronize()
    ms_prof = (time.perf_counter() - t1) / nprof * 1e3
    _lib.profile(False)

    trees = sum(st["trees"] for st in stats)
    kernel_ms = sum(st["count_kernel_ms"] for st in pstats)
    launches = sum(st["count_launches"] for st in pstats)
    kbytes = sum(st["count_bytes"] for st in pstats)
    ebytes = sum(st["count_examined_bytes"] for st in pstats)
    avg_launch_ms = kernel_ms / max(launches, 1)
    achieved = (kbytes / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    kernel = "k_b_count" if args.mode == "batched" else "k_grid_count"
    roofline = {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6),
        "traffic": None if args.phi else traffic_from_profile(kernel, args.config),
        "algorithmic_bytes_per_launch": round(kbytes / max(launches, 1), 1),
        "avg_launch_us": round(avg_launch_ms * 1e3, 3),
        "launches_per_step": launches // nprof,
        # the candidates' packed cells and coordinates the walks read (mostly
        # L2 hits): what the kernel moves through the caches, beside the
        # algorithmic bytes of the neighbours it establishes
        "examined_bytes_per_launch": round(ebytes / max(launches, 1), 1),
        "examined_GBs": round((ebytes / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e9, 2) if avg_launch_ms > 0 else 0.0,
        "kernel_share_of_step": round(kernel_ms / nprof / ms_step, 4),
    }

    # the depth-first driver (one restricted tree at a time) on the same data
    other = same = None
    if args.depth_first_steps > 0:
        omode = _lib.MODE_EXACT if mode == _lib.MODE_BATCHED else _lib.MODE_BATCHED
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.depth_first_steps):
            ores = None
            ores = step(omode)
        torch.cuda.synchronize()
        other = (time.perf_counter() - t1) / args.depth_first_steps * 1e3
        same = ores.subspaces == res.subspaces and np.array_equal(ores.labels, res.labels)
        del ores
        say(f"other driver: {other:.1f} ms/run, same result: {same}")

    # the drop-in call (python/inscy.py GPU5: host tensor in, Python lists out)
    dropin = None
    if args.dropin:
        Xh = torch.from_numpy(w.X)
        t1 = time.perf_counter()
        subs, cls = INSCY.GPU5(Xh, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order)
        dropin = round(time.perf_counter() - t1, 4)
        del subs, cls
        say(f"drop-in GPU5 with lists: {dropin} s")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        model, threads = host_cpu()
        ctrees, csecs, finished = O.ref_rate(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular,
                                             max_seconds=args.cpu_seconds)
        cpu = {
            "value": round(ctrees / csecs, 4),
            "unit": "subspaces/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"first {ctrees} restricted trees of the depth-first run on the same {w.name} data "
                       f"({csecs:.1f} s budget{'; complete run' if finished else ''}); the full run evaluates "
                       f"{res.stats['trees']} trees"),
            "cpu_model": model,
            "host_threads": threads,
        }
        say(f"cpu baseline: {cpu['value']} trees/s")

    learned = bench_learned_leg(args) if args.learned_config else None

    out = {
        "metric": "end-to-end INSCY seconds + subspaces/s; exact vs batched vs learned-prune (F1)",
        "value": round(trees / elapsed, 2),
        "unit": "subspaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: float32 n={w.n} d={w.d}, eps={w.eps} c={w.cells} F={w.F} num_obj={w.num_obj} "
                        f"min_size={w.min_size} r={w.r} rectangular={w.rectangular}",
            "mode": args.mode,
            "depth_first_ms" if args.mode == "batched" else "batched_ms": round(other, 3) if other else None,
            "speedup_vs_depth_first": (round(other / ms_step, 3) if args.mode == "batched" else
                                       round(ms_step / other, 3)) if other else None,
            "modes_agree": same,
            "ms_per_step_with_events": round(ms_prof, 3),
            "dropin_gpu5_lists_seconds": dropin,
            "phase_seconds": dict({k: round(stats[-1][k + "_seconds"], 4) for k in ("expand", "replay", "emit")},
                                  library=round(stats[-1]["seconds"], 4)),
            "library_ms_per_step": round(sum(st["seconds"] for st in stats) / args.steps * 1e3, 3),
            "trees_per_step": trees // max(args.steps, 1),
            "subspaces": res.stats["subspaces"],
            "clusters": res.stats["clusters"],
            "learned": learned,
            "parallelism": "single GPU",
        },
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_learned_leg(args):
    """Learned-prune mode on its evaluation workload against the exact
    level-synchronous run of the same data: seconds of each, speed-up and the
    cluster F1 of the learned result (gpu_inscy_amd.pruner.cluster_f1)."""
    import torch

    from gpu_inscy_amd import _lib, datagen, pruner
    from gpu_inscy_amd import inscy as INSCY

    w = datagen.config(args.learned_config)
    X = torch.from_numpy(w.X).to("cuda:0")
    _, thr, min_depth, _ = pruner.load()
    pruner.use(threshold=thr)

    def run(m):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order, mode=m)
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    run(_lib.MODE_LEARNED)  # warm-up
    exact, exact_s = run(_lib.MODE_BATCHED)
    learned, learned_s = run(_lib.MODE_LEARNED)
    f1 = pruner.cluster_f1(exact, learned)
    out = {"workload": f"{w.name}: n={w.n} d={w.d} eps={w.eps} c={w.cells} min_size={w.min_size}",
           "exact_seconds": round(exact_s, 4), "learned_seconds": round(learned_s, 4),
           "speedup_vs_exact": round(exact_s / learned_s, 3), "cluster_f1": round(f1["f1"], 5),
           "threshold": thr, "trees_exact": exact.stats["trees"], "trees_learned": learned.stats["trees"]}
    say(f"learned leg: {out}")
    return out


# The reference's published end-to-end seconds on its real-data workloads
# (experiments/plot_real.py:13-14): INSCY CPU and GPU-INSCY (INSCY.GPU_memory),
# hardware not stated.
PUBLISHED_REAL = {
    "glass": (2.9117519855499268, 0.061475515365600586),
    "vowel": (36.15232849121094, 0.03729987144470215),
    "pendigits": (31439.810033162434, 3.0369489987691245),
    "sky_0.5x0.5": (77354.21990505855, 4.891563653945923),
    "sky_1x1": (None, 62.529128074645996),
    "sky_1x2": (None, 283.4412166277568),
}


def bench_real(args, mode):
    """One of the reference's real-data workloads: end-to-end seconds per run
    (lower is better) next to the reference's published seconds."""
    import numpy as np
    import torch

    from gpu_inscy_amd import _lib, datagen
    from gpu_inscy_amd import inscy as INSCY

    w = datagen.real(args.real, os.path.join(ROOT, "tests", "golden", "real", "real_inputs.npz"))
    X = torch.from_numpy(w.X).to("cuda:0")

    def step():
        return INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order,
                         mode=mode)

    for _ in range(max(args.warmup, 1)):
        res = step()
    parity = None
    gpath = os.path.join(ROOT, "tests", "golden", "real", "real_golden.json")
    if os.path.exists(gpath):
        g = json.load(open(gpath)).get(args.real)
        if g:
            dig = [hashlib.sha256(np.ascontiguousarray(l, dtype=np.int32).tobytes()).hexdigest() for l in res.labels]
            parity = ([s for s, _, _, _ in g["subspaces"]] == res.subspaces and
                      dig == [x[3] for x in g["subspaces"]] and res.stats["trees"] == g["trees"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    cpu_pub, gpu_pub = PUBLISHED_REAL[args.real]
    out = {
        "metric": "end-to-end INSCY seconds (reference real-data workload)",
        "value": round(sec, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(sec / gpu_pub, 6),
        "dtype": "f32",
        "data": "real (reference's data/real inputs)",
        "config": {
            "workload": f"{w.name}: n={w.n} d={w.d}, eps={w.eps} c={w.cells} F={w.F} num_obj={w.num_obj} "
                        f"min_size={w.min_size} r={w.r} rectangular=True",
            "mode": args.mode,
            "published_gpu_inscy_seconds": gpu_pub,
            "published_cpu_inscy_seconds": cpu_pub,
            "speedup_vs_published_gpu": round(gpu_pub / sec, 2),
            "speedup_vs_published_cpu": round(cpu_pub / sec, 1) if cpu_pub else None,
            "trees_per_step": runs[-1].stats["trees"],
            "subspaces": runs[-1].stats["subspaces"],
            "clusters": runs[-1].stats["clusters"],
        },
        "parity": parity,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned-prune mode against the exact (level-batched) run of the same
    workload: end-to-end seconds per learned run (lower is better), the exact
    run's seconds, and the cluster F1 of the learned result against the exact
    one (gpu_inscy_amd.pruner.cluster_f1)."""
    import torch

    from gpu_inscy_amd import _lib, datagen, pruner
    from gpu_inscy_amd import inscy as INSCY

    w = datagen.config(args.config)
    X = torch.from_numpy(w.X).to("cuda:0")
    _, thr, min_depth, meta = pruner.load()
    if args.threshold is not None:
        thr = args.threshold
    pruner.use(threshold=thr)

    def run(mode):
        return INSCY.run(X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular, w.entropy_order,
                         mode=mode)

    for _ in range(max(args.warmup, 1)):
        learned = run(_lib.MODE_LEARNED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    exact = run(_lib.MODE_BATCHED)
    torch.cuda.synchronize()
    exact_sec = time.perf_counter() - t0
    t0 = time.perf_counter()
    runs = [run(_lib.MODE_LEARNED) for _ in range(args.steps)]
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    learned = runs[-1]
    f1 = pruner.cluster_f1(exact, learned)
    st = learned.stats
    out = {
        "metric": "end-to-end INSCY seconds + subspaces/s; exact vs batched vs learned-prune (F1)",
        "value": round(sec, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (network: bf16 weights, f32 accumulation)",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: n={w.n} d={w.d}, eps={w.eps} c={w.cells} F={w.F} num_obj={w.num_obj} "
                        f"min_size={w.min_size} r={w.r} rectangular=True",
            "mode": "learned",
            "exact_batched_seconds": round(exact_sec, 4),
            "speedup_vs_exact_batched": round(exact_sec / sec, 3),
            "cluster_f1": round(f1["f1"], 5),
            "cluster_recall": round(f1["recall"], 5),
            "cluster_precision": round(f1["precision"], 5),
            "clusters_exact": f1["clusters_exact"],
            "clusters_learned": f1["clusters_approx"],
            "subspaces_exact": len(exact.subspaces),
            "subspaces_learned": len(learned.subspaces),
            "trees_exact": exact.stats["trees"],
            "trees_learned": st["trees"],
            "pairs_scored": st["pairs_scored"],
            "pairs_skipped": st["pairs_skipped"],
            "threshold": thr,
            "min_depth": min_depth,
            "trained_on": meta.get("generator"),
        },
        "parity": None,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_many(args, mode):
    """Config 3: 64 independent data sets through inscy_run_many (one step =
    all of them); value = restricted trees per second over the 64 sets."""
    import torch

    from gpu_inscy_amd import _lib, datagen
    from gpu_inscy_amd import inscy as INSCY

    sets = datagen.multi_instance()
    jobs = [(torch.from_numpy(w.X).to("cuda:0"), w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, w.rectangular,
             w.entropy_order) for w in sets]
    torch.cuda.synchronize()
    for _ in range(max(args.warmup, 1)):
        first = INSCY.run_many(jobs, mode=mode)
    _lib.profile(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = None
    trees = 0
    for _ in range(args.steps):
        last = None
        last = INSCY.run_many(jobs, mode=mode)
        trees += sum(r.stats["trees"] for r in last)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    runs = [last]
    # the last timed call against the oracle's digests of the sets it has finished
    parity, checked = None, 0
    gpath = os.path.join(ROOT, "tests", "golden", "config3.json")
    if os.path.exists(gpath):
        import numpy as np
        gold = {g["workload"]: g for g in json.load(open(gpath))["sets"]}
        parity = True
        for w, r in zip(sets, last):
            g = gold.get(w.name)
            if g is None:
                continue
            checked += 1
            digs = [hashlib.sha256(np.ascontiguousarray(l, dtype=np.int32).tobytes()).hexdigest() for l in r.labels]
            parity = parity and r.subspaces == [x[0] for x in g["subspaces"]] and digs == [x[3] for x in g["subspaces"]]
    # the roofline pass: one call with the count kernel's launches bracketed by HIP events
    _lib.profile(True)
    prof = INSCY.run_many(jobs, mode=mode)
    torch.cuda.synchronize()
    _lib.profile(False)
    kernel_ms = sum(r.stats["count_kernel_ms"] for r in prof)
    launches = sum(r.stats["count_launches"] for r in prof)
    kbytes = sum(r.stats["count_bytes"] for r in prof)
    avg_launch_ms = kernel_ms / max(launches, 1)
    achieved = (kbytes / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    same = all(a.subspaces == b.subspaces for a, b in zip(first, runs[-1]))
    out = {
        "metric": "end-to-end INSCY seconds + subspaces/s; exact vs batched vs learned-prune (F1)",
        "value": round(trees / elapsed, 2),
        "unit": "subspaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(sets)} independent data sets (n {min(w.n for w in sets)}..{max(w.n for w in sets)}, "
                        f"d {min(w.d for w in sets)}..{max(w.d for w in sets)}), eps=0.015 c=6 mu=4 min_size=0.02n, "
                        "inscy_run_many",
            "mode": args.mode,
            "trees_per_step": trees // max(args.steps, 1),
            "subspaces": sum(r.stats["subspaces"] for r in runs[-1]),
            "deterministic": same,
            "parity_sets": checked,
            "parallelism": "single GPU, concurrent streams",
        },
        "parity": parity,
        "roofline": {"bound": "hbm", "kernel": "k_b_count" if args.mode == "batched" else "k_grid_count",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
