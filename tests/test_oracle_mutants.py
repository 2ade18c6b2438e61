"""Mutation checks of the oracle's closed-form pins (CPU only).

Each mutant is a copy of oracle/dem_oracle.cpp with one plausible mistake
planted in the particle-particle branch (src/CabanaDEM_Force.hpp):
  no_coulomb_cap       the mu |F_n| limit of :293-296 dropped
  no_q3_normalisation  the history stored as -F_t/k_t instead of its unit
                       vector (:307-319, quirk Q3)
  meff_from_mi_alone   the damping mass 1/(1/m_i + 1/m_j) of :212-214 formed
                       from m_i only
The copy is compiled into a temporary directory and the closed-form tests
(tests/test_oracle_pp_closed_form.py) run against it through ORACLE_LIB; each
mutant must fail them, and the unmodified source must pass under the same
harness."""
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


def _build_and_run(tmp_path, source):
    (tmp_path / "dem_oracle.cpp").write_text(source)
    (tmp_path / "dem_oracle.h").write_text((ROOT / "oracle" / "dem_oracle.h").read_text())
    lib = tmp_path / "libdem_oracle_mutant.so"
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c++17",
                    "-shared", "-o", str(lib), str(tmp_path / "dem_oracle.cpp")], check=True)
    env = dict(os.environ, ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                           str(ROOT / "tests" / "test_oracle_pp_closed_form.py")],
                          cwd=ROOT, env=env, capture_output=True, text=True)


def test_unmodified_oracle_passes_under_the_harness(tmp_path):
    r = _build_and_run(tmp_path, SRC.read_text())
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_the_closed_forms(tmp_path, name):
    old, new = MUTANTS[name]
    src = SRC.read_text()
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    r = _build_and_run(tmp_path, src.replace(old, new))
    assert r.returncode == 1, r.stdout[-2000:]  # 1: tests ran and some failed
    assert " failed" in r.stdout
