"""Builds libefm.so in-tree for gfx950.

Each .hip file is compiled with hipcc to an object; the objects are linked with
g++ into a shared library that does NOT record a dependency on a HIP runtime.
Its HIP symbols bind, at load time, to the runtime already loaded in the
process — torch's when used from Python (see _capi.py), so that torch's
streams and device pointers are the library's own.

    python -m elastic_fm_vp_vs_rho_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libefm.so")
HOST_EXAMPLE = os.path.join(HERE, "efm_host_example")
SOURCES = ("forward.hip", "adjoint.hip", "reference.hip", "capi.hip", "initnet.hip", "conditioning.hip", "plan.hip")
ARCH = os.environ.get("EFM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 arithmetic stays off (v_pk_*_f32 are not generated)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, warn in results:
        if verbose and warn.strip():
            print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp, *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_example()
    return LIB


def build_host_example():
    """The C host example (csrc/host_example.c): a program without Python that
    links ROCm's HIP runtime and libefm.so (found next to it)."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tmp = HOST_EXAMPLE + ".tmp"
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE,
           "-I", os.path.join(rocm, "include"), os.path.join(CSRC, "host_example.c"), "-o", tmp,
           "-L", HERE, "-l:libefm.so", "-Wl,-rpath,$ORIGIN",
           "-L", os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host example build failed:\n{r.stderr}")
    os.replace(tmp, HOST_EXAMPLE)
    return HOST_EXAMPLE


if __name__ == "__main__":
    print(build(verbose=True))
