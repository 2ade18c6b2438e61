"""The reference-shaped C++ headers compile with the host compiler and the
input reader behaves as src/CabanaDEM_Input.hpp:24-82 does (CPU only)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_inputs_reader(tmp_path):
    exe = tmp_path / "test_inputs"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "cpp" / "test_inputs.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(ROOT / "tests" / "golden" / "inputs")], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")


def test_full_api_compiles_and_links(tmp_path):
    exe = tmp_path / "dem_validation"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "examples" / "dem_validation.cpp"), "-L",
                    str(ROOT / "cabanadem_amd" / "lib"), "-lcabanadem_hip", "-o", str(exe)], check=True)
    assert exe.exists()


def test_particle_insertion_resize_flow(tmp_path):
    """The container flow of examples/tst_01_particle_insertion.cpp (grow,
    refill through slices, move, shrink, CSV output), host only."""
    exe = tmp_path / "test_particle_insertion"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "cpp" / "test_particle_insertion.cpp"), "-L",
                    str(ROOT / "cabanadem_amd" / "lib"), "-lcabanadem_hip",
                    "-Wl,-rpath," + str(ROOT / "cabanadem_amd" / "lib"), "-o", str(exe)], check=True)
    out_dir = tmp_path / "out"
    p = subprocess.run([str(exe), str(out_dir)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.strip().endswith("OK")
    assert sorted(f.name for f in out_dir.glob("particles_*.csv")) == [
        "particles_0.csv", "particles_10.csv", "particles_20.csv", "particles_30.csv"]


HDF5_INC = Path("/opt/conda/include/hdf5.h")
HDF5_LIB = Path("/opt/conda/lib/libhdf5.so")


@pytest.mark.skipif(not (HDF5_INC.exists() and HDF5_LIB.exists()), reason="no HDF5 C library")
def test_hdf5_output_layout(tmp_path):
    """Particles::output and ParticlesBondedDEM::output built with
    CABANADEM_ENABLE_HDF5 write the reference's step files: read back through
    the HDF5 C API, every field sits under its slice label with the values
    written through the slices, with the Time attribute (host only)."""
    # libhdf5 through a directory holding it alone: the library's own folder
    # also holds an older libstdc++ that the HIP runtime cannot use
    libdir = tmp_path / "h5lib"
    libdir.mkdir()
    for name in ("libhdf5.so.103", "libz.so.1"):
        (libdir / name).symlink_to(HDF5_LIB.parent / name)
    exe = tmp_path / "test_hdf5_output"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-DCABANADEM_ENABLE_HDF5",
                    "-I", str(ROOT / "include"), "-I", str(HDF5_INC.parent),
                    str(ROOT / "tests" / "cpp" / "test_hdf5_output.cpp"),
                    "-L", str(ROOT / "cabanadem_amd" / "lib"), "-lcabanadem_hip", str(HDF5_LIB),
                    "-Wl,-rpath," + str(ROOT / "cabanadem_amd" / "lib"), "-Wl,-rpath," + str(libdir),
                    "-o", str(exe)], check=True)
    p = subprocess.run([str(exe), str(tmp_path / "out")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.strip().endswith("OK")
    assert (tmp_path / "out" / "particles_7.xmf").exists()
