"""Build recipe for libmpm_hip.so (gfx950) and the CPU oracle.

    python mpm-engine_amd/build.py          # both
    python mpm-engine_amd/build.py --hip    # only the HIP library

The HIP library is compiled with hipcc for gfx950 only, hardware float atomics
(-munsafe-fp-atomics) and packed-FP32 instructions disabled. Objects go to
mpm-engine_amd/_build, the library to mpm-engine_amd/lib (both git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmpm_hip.so")
SOURCES = ["mpm_kernels.hip", "mpm_particles.hip", "mpm_sort.hip", "mpm_adjoint.hip", "mpm_cdf.hip",
           "mpm_coupling.hip", "mpm_geometry.hip", "mpm_rigid.hip",
           "mpm_modifiers.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics", "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(REPO, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")


def build_hip(jobs: int | None = None):
    """Compile every source (in parallel: each hipcc is one process) and link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objs = [os.path.join(BUILD, src.replace(".hip", ".o")) for src in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)), 16)
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
                  for src, obj in zip(SOURCES, objs)]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle")])


if __name__ == "__main__":
    args = set(sys.argv[1:])
    if "--oracle" not in args:
        print("built", build_hip())
    if "--hip" not in args:
        build_oracle()
        print("built oracle")
