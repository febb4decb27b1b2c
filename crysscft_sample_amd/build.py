"""Build recipe of libcrys_solver.so (gfx950) and of the CPU oracle.

``python -m crysscft_sample_amd.build`` compiles in-tree; ``__graft_entry__.build()``
calls :func:`build_all`.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libcrys_solver.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

HIP_SOURCES = ["crys_solver.hip", "batch.hip", "grid_sync.hip"]
CXX_SOURCES = ["operators.cpp", "field_io.cpp"]

ARCH = os.environ.get("CRYS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose: bool = True) -> str:
    """Compile the HIP C-ABI library in-tree (always rebuilds)."""
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC, "-fvisibility=hidden",
              "-Wall"]
    for src in CXX_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        _run(["g++", *common, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        # Packed FP32 instructions are disabled for every device compile.
        _run([hipcc(), f"--offload-arch={ARCH}", *common,
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


# The reference's solver plugin header (solver_base.h and utilities/c_array.h),
# read where the reference lies; only this container has it.
REFERENCE_SRC = "/root/reference/src"
ADAPTER_DIR = os.path.join(PKG, "integration")
ADAPTER_CHECK = os.path.join(ADAPTER_DIR, "adapter_check")


def build_adapter_check(out: str = ADAPTER_CHECK) -> str | None:
    """Compile integration/adapter_check.cpp, the C++ adapter solver_gpu.h
    driven through the reference's own CrysPSSolverBase, against
    solver_base.h from the reference sources and libcrys_solver.so.  Returns
    the binary, or None when the reference sources are absent."""
    if not os.path.exists(os.path.join(REFERENCE_SRC, "solver_base.h")):
        return None
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", REFERENCE_SRC,
          "-I", INCLUDE, "-I", ADAPTER_DIR, os.path.join(ADAPTER_DIR, "adapter_check.cpp"),
          "-L", PKG, "-lcrys_solver", "-Wl,-rpath,$ORIGIN/..", "-o", out])
    return out


def build_oracle() -> str:
    oracle_dir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle_dir, "liboracle.so"])
    return os.path.join(oracle_dir, "liboracle.so")


def build_all() -> None:
    build_library()
    build_oracle()
    build_adapter_check()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
