"""The reference-shaped C++ headers compile with the host compiler and the
input reader behaves as src/CabanaDEM_Input.hpp:24-82 does (CPU only)."""
import subprocess
from pathlib import Path

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
