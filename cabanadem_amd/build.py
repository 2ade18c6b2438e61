"""Build the HIP library (libcabanadem_hip.so) and the CPU oracle in-tree.

The library is compiled for gfx950 with hipcc.  FMA contraction is off
(-ffp-contract=off) so the neighbour distance test and the overlap test use
the reference's exact expressions; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib" / "libcabanadem_hip.so"
ORACLE_DIR = ROOT / "oracle"

# the force kernels are instantiated per tile size in three translation units
# (dem_force*.hip), compiled in parallel with the others
HIP_SOURCES = ["dem_api.hip", "dem_build.hip", "dem_force.hip", "dem_force_b64.hip",
               "dem_force_b256.hip", "dem_gen.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], cwd=cwd, check=True)


def build_library(extra_flags=None) -> Path:
    """Each source to an object in parallel (hipcc -c), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    LIB.parent.mkdir(parents=True, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = LIB.parent / "obj"
    obj_dir.mkdir(exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra_flags or [])
    objs = [obj_dir / (src + ".o") for src in HIP_SOURCES]
    workers = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(workers) as pool:
        jobs = [pool.submit(_run, [hipcc] + flags + ["-c", CSRC / src, "-o", o], CSRC)
                for src, o in zip(HIP_SOURCES, objs)]
        for j in jobs:
            j.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


GNN_LIB = LIB.parent / "libcabanadem_gnn.so"


def build_gnn_library(extra_flags=None) -> Path:
    """The surrogate's kernels (gnn.hip) as their own shared library."""
    LIB.parent.mkdir(parents=True, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = GNN_LIB.with_suffix(".so.tmp")
    _run([hipcc] + HIPCC_FLAGS + list(extra_flags or []) + [CSRC / "gnn.hip", "-o", tmp], cwd=CSRC)
    os.replace(tmp, GNN_LIB)
    return GNN_LIB


def build_copy_probe() -> Path:
    """Streaming-copy probe (known bytes) used to calibrate the HBM counters."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    exe = LIB.parent / "hbm_copy"
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    _run([hipcc] + flags + [CSRC / "hbm_copy.hip", "-o", exe], cwd=CSRC)
    return exe


def build_oracle() -> Path:
    _run(["make", "-C", ORACLE_DIR, "-B"])
    return ORACLE_DIR / "_build" / "libdem_oracle.so"


EXAMPLES = {
    "dem_validation": "examples/dem_validation.cpp",
    "bonded_tensile": "examples/bonded_tensile.cpp",
    # the GPU insertion test's driver (tests/test_gpu_insertion.py)
    "test_insertion_gpu": "tests/cpp/test_insertion_gpu.cpp",
}


def build_examples() -> list:
    """Host programs on the reference-shaped C++ API, linked to the library
    (rpath relative to the binary, so the tree can move)."""
    out_dir = ROOT / "examples" / "_build"
    out_dir.mkdir(parents=True, exist_ok=True)
    built = []
    for name, src in EXAMPLES.items():
        exe = out_dir / name
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", ROOT / "include", ROOT / src,
              "-L", LIB.parent, "-lcabanadem_hip", "-Wl,-rpath,$ORIGIN/../../cabanadem_amd/lib",
              "-o", exe])
        built.append(exe)
    return built


def build_all():
    build_oracle()
    build_library()
    build_gnn_library()
    build_examples()
    build_copy_probe()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "lib":
        build_library()
    elif what == "gnn":
        build_gnn_library()
    elif what == "oracle":
        build_oracle()
    elif what == "examples":
        build_examples()
    else:
        build_all()
