"""Builds libm2.so (HIP kernels + C ABI) and the `mersenne2` command line for gfx950.

    python -m mersenne2gpu_amd.build          # or: __graft_entry__.build()

Outputs go next to this file: libm2.so and mersenne2 (objects in _build/).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("M2_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["m2_kernels.hip", "m2_host.cpp", "m2_api.cpp", "m2_gerbicz.cpp", "m2_noalloc.hip"]
CLI_SOURCES = ["mersenne2_cli.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC, "-Wall"]
    # No packed FP32 instructions anywhere (the kernels use none; the flag
    # keeps the compiler from forming them).
    hipflags = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([cc] + hipflags + common + ["-c", path, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(HERE, "libm2.so")
    tmp = lib + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    cli = os.path.join(HERE, "mersenne2")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE] + [os.path.join(CSRC, s) for s in CLI_SOURCES]
         + ["-L", HERE, "-lm2", "-Wl,-rpath,$ORIGIN", "-o", cli + ".tmp"])
    os.replace(cli + ".tmp", cli)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
