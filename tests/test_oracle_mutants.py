"""Mutation checks of the oracle's closed-form pins (CPU only).

Each mutant is a copy of oracle/dem_oracle.cpp with one plausible mistake
planted in the particle-particle branch (src/CabanaDEM_Force.hpp):
  no_coulomb_cap       the mu |F_n| limit of :293-296 dropped
  no_q3_normalisation  the history stored as -F_t/k_t instead of its unit
                       vector (:307-319, quirk Q3)
  meff_from_mi_alone   the damping mass 1/(1/m_i + 1/m_j) of :212-214 formed
                       from m_i only
and in the moving-wall terms of the particle-wall branch (corrected mode):
  wall_velocity_dropped  the wall surface's velocity left out of the
                         relative velocity at the contact point
  wall_spin_about_origin the spin's lever arm c instead of c - x_w
  wall_not_translated    the wall point kept where it was uploaded (along
                         the floors' normal, z)
The copy is compiled into a temporary directory and the closed-form tests
(tests/test_oracle_pp_closed_form.py, tests/test_oracle_moving_walls.py) run
against it through ORACLE_LIB; each mutant must fail the set it targets, and
the unmodified source must pass under the same harness."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "dem_oracle.cpp"

MUTANTS = {
    "no_coulomb_cap": (
        "        if (ftms > prm->fr_pp * fnm) ftms = prm->fr_pp * fnm;\n", ""),
    "no_q3_normalisation": (
        "sx[found_at] = m0 / lm; sy[found_at] = m1 / lm; sz[found_at] = m2 / lm;",
        "sx[found_at] = m0; sy[found_at] = m1; sz[found_at] = m2;"),
    "meff_from_mi_alone": (
        "      t2 = 1. / p->m[j];\n", "      t2 = 1. / p->m[i];\n"),
}
WALL_MUTANTS = {
    "wall_velocity_dropped": (
        "      double v0 = vi0 - vw0, v1 = vi1 - vw1, v2 = vi2 - vw2;",
        "      double v0 = vi0, v1 = vi1, v2 = vi2;"),
    "wall_spin_about_origin": (
        "        const double c0 = x[3 * i + 0] - a_i * nx - xw0;",
        "        const double c0 = x[3 * i + 0] - a_i * nx;"),
    "wall_not_translated": (
        "        xw2 += wl->u[3 * j + 2] * prm->wall_time;",
        "        xw2 += 0.;"),
}
PP_TESTS = "test_oracle_pp_closed_form.py"
WALL_TESTS = "test_oracle_moving_walls.py"


def _build_and_run(tmp_path, source, tests=(PP_TESTS, WALL_TESTS)):
    (tmp_path / "dem_oracle.cpp").write_text(source)
    (tmp_path / "dem_oracle.h").write_text((ROOT / "oracle" / "dem_oracle.h").read_text())
    lib = tmp_path / "libdem_oracle_mutant.so"
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c++17",
                    "-shared", "-o", str(lib), str(tmp_path / "dem_oracle.cpp")], check=True)
    env = dict(os.environ, ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider"] +
                          [str(ROOT / "tests" / t) for t in tests],
                          cwd=ROOT, env=env, capture_output=True, text=True)


def test_unmodified_oracle_passes_under_the_harness(tmp_path):
    r = _build_and_run(tmp_path, SRC.read_text())
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name", sorted(MUTANTS) + sorted(WALL_MUTANTS))
def test_mutant_fails_the_closed_forms(tmp_path, name):
    old, new = MUTANTS[name] if name in MUTANTS else WALL_MUTANTS[name]
    tests = (PP_TESTS,) if name in MUTANTS else (WALL_TESTS,)
    src = SRC.read_text()
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    r = _build_and_run(tmp_path, src.replace(old, new), tests)
    assert r.returncode == 1, r.stdout[-2000:]  # 1: tests ran and some failed
    assert " failed" in r.stdout
