"""Builds the HIP library (librgs_hip.so, gfx950) and the oracle.

The library is compiled in-tree so that the built .so travels with the
repository to the GPU host.  Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(PKG_DIR, "librgs_hip.so")
ORACLE_DIR = os.path.join(REPO_DIR, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "librgs_oracle.so")

SOURCES = ["k_project.hip", "k_bin.hip", "k_raster.hip", "k_misc.hip", "rgs_api.hip", "rgs_noalloc.hip", "k_loss.hip", "k_topology.hip", "rgs_topology.hip", "k_eval.hip", "rgs_eval.hip", "rgs_io.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RGS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the binning decisions must be reproducible bit for bit,
#   so no multiply-add is fused unless the source asks for it.
# -fhip-fp32-correctly-rounded-divide-sqrt: IEEE division and square root.
# -munsafe-fp-atomics: hardware float atomics for the gradient accumulators.
# -packed-fp32-ops disabled: no packed FP32 instructions.
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(REPO_DIR, "include"),
]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    return proc.stdout


def build_hip(jobs: int | None = None) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE_DIR])
    return ORACLE_LIB


def build_all() -> None:
    build_oracle()
    build_hip()


if __name__ == "__main__":
    build_all()
    print(LIB_PATH)
    print(ORACLE_LIB)
