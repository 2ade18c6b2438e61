"""MI355X-native CabanaDEM time-step library.

The product is libcabanadem_hip.so (HIP kernels for gfx950 behind the C ABI
declared in include/cabanadem.h) and the reference-shaped C++ headers in
include/.  This package holds their sources (csrc/), the build script and a
ctypes binding (abi.py) used by the tests and the benchmark.
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_ROOT = PACKAGE_DIR.parent
INCLUDE_DIR = REPO_ROOT / "include"
LIB_PATH = PACKAGE_DIR / "lib" / "libcabanadem_hip.so"

__all__ = ["PACKAGE_DIR", "REPO_ROOT", "INCLUDE_DIR", "LIB_PATH"]
