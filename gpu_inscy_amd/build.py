"""Builds libinscy_hip.so (the HIP engine + C ABI) in-tree for gfx950.

    python -m gpu_inscy_amd.build          # or gpu_inscy_amd.build.build()

hipcc cross-compiles without a GPU.  Flags:
  -ffp-contract=off      the distance test must round like the reference's host
                         code (separate multiply and add, ClusteringCpu.cu:40-50);
  -packed-fp32-ops off   packed FP32 instructions are disabled for every kernel.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libinscy_hip.so")
SOURCES = ["inscy_kernels.hip", "inscy_level_kernels.hip", "inscy_pruner.hip", "inscy_mlp.hip", "inscy_level.hip",
           "inscy_engine.hip", "inscy_noalloc.hip",
           "inscy_scan.hip", "inscy_perm.hip"]
ARCH = os.environ.get("INSCY_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
    ]


def build(verbose: bool = False, csrc: str = CSRC, out: str = LIB) -> str:
    """Compiles every HIP source (from csrc) and links the shared library
    (out); returns its path."""
    cc = hipcc()
    objs = []
    build_dir = os.path.join(os.path.dirname(out), "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:  # the translation units compile concurrently
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [cc, *flags(), "-c", os.path.join(csrc, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, proc in procs:
        if proc.wait() != 0:
            raise subprocess.CalledProcessError(proc.returncode, cmd)
    tmp = out + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
