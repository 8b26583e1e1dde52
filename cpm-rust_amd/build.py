"""Builds the engine library in-tree: cpm-rust_amd/cpm/libcpm_engine.so.

HIP translation units are compiled by hipcc for gfx950 only; host C++ by
g++.  FP64 energy arithmetic must match the reference's operation order, so
no FP contraction and no fast-math anywhere.  Usage: python build.py
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "cpm")
LIB = os.path.join(OUT_DIR, "libcpm_engine.so")
LDS_BW = os.path.join(OUT_DIR, "lds_bandwidth")
OBJ_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-pthread"]

HIP_SOURCES = ["engine.hip"]
CXX_SOURCES = ["host.cpp", "capi.cpp", "rng_jump.cpp", "serve.cpp"]
HEADERS = ["core.hpp", "binomial.hpp", "metropolis.hpp", "host.hpp", "device.hpp", "rng_jump.hpp", "driver.hpp", "serve.hpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for j in jobs]
    failed = False
    for j, p in zip(jobs, procs):
        out, err = p.communicate()
        if p.returncode != 0 or verbose:
            sys.stderr.write(out + err)
        if p.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("engine compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", "-pthread", "-Wl,--no-undefined",
          *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # the LDS read-bandwidth microbenchmark (the LDS roofline's peak)
    _run([HIPCC, "-O3", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", os.path.join(CSRC, "lds_bandwidth.hip"), "-o", LDS_BW])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
