"""Builds libemopic.so (HIP, gfx950) in-tree with hipcc.

    python -m apc_524_project_amd.build

Every source is compiled with -ffp-contract=off where FP64 parity with the
CPU oracle is required, and all HIP code with packed-FP32 instructions
disabled.  Objects go to apc_524_project_amd/_build/, the library to
apc_524_project_amd/libemopic.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libemopic.so")
HOST = os.path.join(PKG, "host")
RUN_MAIN = os.path.join(PKG, "emo_run_main")
CXX = os.environ.get("CXX", "g++")

ARCH = os.environ.get("EMO_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]
# (source, extra flags)
SOURCES = [
    ("plain_kernels.hip", ["-ffp-contract=off"]),
    ("field_kernels.hip", ["-ffp-contract=off"]),
    ("emo_api.cpp", ["-ffp-contract=off"]),
    ("loader.cpp", ["-ffp-contract=off"]),
    ("boundary.cpp", ["-ffp-contract=off"]),
    ("chosen.hip", ["-ffp-contract=off"]),
    ("adjoint.hip", ["-ffp-contract=off"]),
]


def _run(cmd):
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, _run(cmd)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    shutil.move(tmp, LIB)
    build_host()
    return LIB


def build_host() -> str:
    """The C++ facade (host/emopic.cpp) and the driver that runs the
    reference's main loop through it (host/run_main.cpp), linked against
    libemopic.so; host code only, so the host compiler builds it."""
    tmp = RUN_MAIN + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I", os.path.join(ROOT, "include"),
           os.path.join(HOST, "emopic.cpp"), os.path.join(HOST, "readinfo.cpp"), os.path.join(HOST, "run_main.cpp"),
           "-L", PKG, "-lemopic", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host facade build failed:\n{r.stdout}")
    shutil.move(tmp, RUN_MAIN)
    return RUN_MAIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
