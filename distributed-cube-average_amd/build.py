"""Build libdc_hip.so (and the dc_hip command-line program) in-tree.

    python distributed-cube-average_amd/build.py

HIP sources are compiled for gfx950 only, with packed-FP32 instruction
selection disabled; the host-side model setup is compiled with gcc (OpenMP,
no FP contraction, so its float results match the reference's host code).
Objects go to distributed-cube-average_amd/_build/, the library and the
program next to this file.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdc_hip.so")
BIN = os.path.join(HERE, "dc_hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + CSRC]
HOST_FLAGS = ["-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c++17",
              "-Wall", "-I" + INC, "-I" + CSRC]

HIP_SOURCES = ["stencil.hip", "halo.hip", "device_data.hip", "program.hip", "batch.hip", "learned.hip",
               "storage.hip", "noalloc.hip", "model.hip", "wall.hip"]
HOST_SOURCES = ["host_model.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o",
                     obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lgomp"])
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "dc_main.cpp")
    if os.path.exists(main_src):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INC, main_src, "-o",
              BIN, "-L" + HERE, "-ldc_hip", "-Wl,-rpath,$ORIGIN",
              "-Wl,-rpath-link,/opt/rocm/lib"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
