"""Build recipe of the HIP library (gfx950) -- ``python -m plugin_navierstokes_amd.build``.

Produces ``plugin_navierstokes_amd/lib/libnsfv1.so`` in-tree with hipcc.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnsfv1.so")

SOURCES = ["nsfv1_api.cpp", "nsfv1_assemble.hip", "nsfv1_march.hip", "nsfv1_solver.hip", "nsfv1_mlp.hip", "nsfv1_boundary.hip", "nsfv1_export.hip", "nsfv1_coupled.hip", "nsfv1_scratch.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NSFV1_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = True, jobs: int | None = None) -> str:
    """Compile every source (in parallel, one hipcc per source) and link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIBDIR, exist_ok=True)
    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(subprocess.check_call, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *COMMON, "-shared", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
