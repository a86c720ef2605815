"""Build the gfx950 shared library (orion_amd/_build/liborion_hip.so) in-tree.

hipcc compiles each HIP translation unit for gfx950 and links one shared
library exposing the C ABI in include/orion_hip.h.  Flags that matter for
results:
  -ffp-contract=off      a*b+c stays two rounded operations (the
                         reference's x86-64 build has no FMA);
  no fast-math           f64 division and sqrt stay correctly rounded;
  -packed-fp32-ops off   no packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liborion_hip.so")
SOURCES = ["context.hip", "io.hip", "k_flux.hip", "k_point.hip", "k_bc.hip", "k_lusgs.hip", "k_metrics.hip",
           "noalloc.hip"]
HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["orion_host.cpp", "orion_run.cpp"]
RUNNER = os.path.join(BUILD, "orion_run")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("ORION_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
            "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include"), "-Wno-unused-result"]


def build(verbose: bool = False, jobs: int = 6) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [cc] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr[-20000:]))
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """The C++ host mirror (orion_amd/host) and the orion_run executable,
    compiled by the host C++ compiler against the C ABI only."""
    cxx = os.environ.get("CXX", "g++")
    tmp = RUNNER + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(os.path.dirname(HERE), "include")]
    cmd += [os.path.join(HOST, s) for s in HOST_SOURCES]
    cmd += ["-o", tmp, "-L" + BUILD, "-lorion_hip", "-Wl,-rpath,$ORIGIN", "-L" + os.path.join(ROCM, "lib"),
            "-Wl,-rpath-link," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n%s%s" % (r.stdout, r.stderr[-20000:]))
    os.replace(tmp, RUNNER)
    return RUNNER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
