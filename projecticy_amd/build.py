"""Build the gfx950 shared library ``libicy3d.so`` in-tree with hipcc.

One fat library holds every HIP kernel, the C ABI of include/icy3d.h and
the host FP64 rigid-body solver.  Called by ``__graft_entry__.build()``; the
Python package only loads the result and never builds at import time.
"""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libicy3d.so"
SOURCES = ["icy3d_flow.hip", "icy3d_geometry.hip", "icy3d_thermal.hip", "rigid3d.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 VALU forms are disabled: ordinary float code only.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False, jobs=4):
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    objects, procs = [], []
    for name in SOURCES:
        src = CSRC / name
        obj = build_dir / (name + ".o")
        objects.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        if name.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"hipcc failed for {name}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB.with_suffix(f".so.tmp{os.getpid()}")
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
