"""Compiles the HIP sources under csrc/ into libddg.so (gfx950) in place.

Each source is compiled to an object on its own (in parallel), then the
objects are linked into one shared library.

    python -m distributed_diskann_amd.build
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libddg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",  # the FP32 distances keep the reference's summation order
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    inc = ["-I", os.path.join(HERE, "..", "include")]
    with tempfile.TemporaryDirectory(prefix="ddg_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]
        jobs = min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
        with ThreadPoolExecutor(jobs) as ex:
            rcs = list(ex.map(lambda so: _run([HIPCC, *FLAGS, *inc, "-c", so[0], "-o", so[1]], verbose),
                              zip(srcs, objs)))
        if any(rcs):
            raise RuntimeError("hipcc failed compiling " +
                               ", ".join(os.path.basename(s) for s, rc in zip(srcs, rcs) if rc))
        tmp_lib = LIB + ".tmp"
        if _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib, *objs], verbose):
            raise RuntimeError("hipcc failed linking libddg.so")
    os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
