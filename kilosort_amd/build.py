"""Build libksm.so (the gfx950 HIP library) in-tree with hipcc.

    python -m kilosort_amd.build

Every source is compiled to an object of its own (in parallel, under
kilosort_amd/_build/), then the objects are linked into kilosort_amd/libksm.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libksm.so")
OBJDIR = os.path.join(HERE, "_build")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # products are rounded before they are added, as in the reference's
    # separate torch multiply and subtract; FMAs are written explicitly
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                  glob.glob(os.path.join(CSRC, "*.cpp")))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}: {cmd[-1]}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP source into kilosort_amd/libksm.so; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    srcs = sources()
    objs = [os.path.join(OBJDIR, os.path.basename(s) + ".o") for s in srcs]
    cmds = [[cc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", s, "-o", o]
            for s, o in zip(srcs, objs)]
    n = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 16)))
    with cf.ThreadPoolExecutor(n) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
