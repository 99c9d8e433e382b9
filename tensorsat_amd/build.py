"""
Build recipes for the native libraries, in-tree:

* ``tensorsat_amd/libtsplan.so`` — host C++ planner (``csrc/plan.cpp``), built with g++.
* ``tensorsat_amd/libtensorsat_amd.so`` — HIP kernels + C-ABI executor for gfx950
  (``csrc/*.hip``), built with hipcc.

Run ``python -m tensorsat_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

PLAN_LIB = PKG / "libtsplan.so"
HIP_LIB = PKG / "libtensorsat_amd.so"

PLAN_SOURCES = ["plan.cpp", "prog.cpp", "cubes.cpp", "packets.cpp", "searchpack.cpp"]
HIP_SOURCES = ["executor.hip", "dense.hip", "batch.hip", "gnn.hip", "search.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_plan(verbose: bool = True) -> Path:
    cxx = shutil.which("g++") or "g++"
    tmp = PLAN_LIB.with_suffix(".so.tmp")
    srcs = [str(CSRC / s) for s in PLAN_SOURCES]
    _run([cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          f"-I{INCLUDE}", *srcs, "-o", str(tmp)])
    os.replace(tmp, PLAN_LIB)
    return PLAN_LIB


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP executor cannot be built")


def build_hip() -> Path:
    srcs = [str(CSRC / s) for s in HIP_SOURCES if (CSRC / s).exists()]
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", *srcs, "-o", str(tmp)])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_all() -> None:
    build_plan()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "plan" in what:
        build_plan()
    if "hip" in what:
        build_hip()
    if "all" in what:
        build_all()
