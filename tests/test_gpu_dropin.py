"""The reference's validation set-ups through the drop-in C++ API on the GPU
(examples/dem_validation.cpp, built by build()), checked against the oracle
running the same set-up and against the closed forms."""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import scenarios as S
from gpu_helpers import normwise
from oracle_driver import solver_run

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "examples" / "_build" / "dem_validation"
INPUTS = ROOT / "tests" / "golden" / "inputs"


def run_driver(case, json_name, out_dir, extra=None):
    assert EXE.exists(), "build() compiles examples/_build/dem_validation"
    cmd = [str(EXE), case, str(INPUTS / json_name), str(out_dir)] + ([extra] if extra else [])
    print("driver:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def state_arrays(res):
    ps = res["particles"]
    return {k: np.array([p[k] for p in ps]) for k in ("x", "u", "omega", "force") if k in ps[0]}


def compare_to_oracle(res, st, tol=1e-9):
    g = state_arrays(res)
    assert normwise(g["x"], st.x) <= tol
    assert normwise(g["u"], st.u) <= tol
    if np.abs(st.omega).max() > 0:
        assert normwise(g["omega"], st.omega) <= tol


def test_head_on_matches_oracle_and_hertz(gpu_lib, tmp_path):
    res = run_driver("head_on", "01_elastic_normal_impact_of_two_identical_particles.json", tmp_path)
    st, w, prm, run, inp = S.ex01()
    solver_run(st, w, prm, run)
    compare_to_oracle(res, st)
    # the solver's CSV output every 30 steps: peak contact force near the Hertz value
    peak = 0.0
    for f in tmp_path.glob("particles_*.csv"):
        a = np.loadtxt(f, delimiter=",", skiprows=2, ndmin=2)
        peak = max(peak, a[1, 9])  # forces_0 of particle 1
    assert abs(peak - 10.697e3) / 10.697e3 < 0.02  # sampled every 3 us of a 40 us contact
    assert len(list(tmp_path.glob("particles_*.csv"))) == 800 // 30 + 1


def test_sphere_wall_matches_oracle(gpu_lib, tmp_path):
    res = run_driver("sphere_wall", "02_elastic_normal_impact_of_particle_wall.json", tmp_path)
    st, w, prm, run, inp = S.ex02()
    solver_run(st, w, prm, run)
    compare_to_oracle(res, st)


@pytest.mark.parametrize("cor", [0.01, 0.2, 0.4, 0.6, 0.8, 1.0])  # automate.py:266-290
def test_restitution_matches_oracle(gpu_lib, tmp_path, cor):
    res = run_driver("cor", "03_normal_particle_wall_different_cor.json", tmp_path, f"--cor-pw={cor}")
    st, w, prm, run, inp = S.ex03(cor)
    solver_run(st, w, prm, run)
    compare_to_oracle(res, st)
    assert abs(res["particles"][0]["u"][1] / inp["velocity_p"] - cor) < 0.02 + 0.02 * cor


ANGLES = [2., 5., 10., 15., 20., 25., 30., 35., 40., 50., 55., 60., 65., 70.]  # automate.py:358-414


@pytest.mark.parametrize("angle", ANGLES)
def test_oblique_matches_oracle(gpu_lib, tmp_path, angle):
    res = run_driver("oblique", "04_oblique_particle_wall_different_angles.json", tmp_path, f"--angle={angle}")
    st, w, prm, run, inp = S.ex04(angle)
    solver_run(st, w, prm, run)
    compare_to_oracle(res, st)
    assert abs(res["particles"][0]["omega"][2] - st.omega[0, 2]) <= 1e-9 * abs(st.omega[0, 2])


def test_oblique_spin_curve_against_reference_csvs(gpu_lib, tmp_path):
    """The spin after an oblique wall impact against the incidence angle, from
    the GPU drop-in driver, against the two data sets the reference plots
    (tests/golden/validation_data: Lethe DEM and Kharaz-Gorham-Salman)."""
    vd = ROOT / "tests" / "golden" / "validation_data"
    angles = [10., 20., 30., 40., 50., 60.]
    om = []
    for a in angles:
        d = tmp_path / f"a{int(a)}"
        d.mkdir()
        res = run_driver("oblique", "04_oblique_particle_wall_different_angles.json", d, f"--angle={a}")
        om.append(res["particles"][0]["omega"][2])
    om = np.array(om)
    assert np.all(om > 0)  # the reference's sign (contact point at +a n)
    for name in ("04_oblique_particle_wall_different_angles_omega_vs_incident_angle_Lethe_DEM_data.csv",
                 "04_oblique_particle_wall_different_angles_omega_vs_incident_angle_Kharaz_Gorham_Salman_experimental_data.csv"):
        csv = np.loadtxt(vd / name, delimiter=",")
        ref = np.interp(angles, csv[:, 0], csv[:, 1])
        assert np.max(np.abs(om - ref)) < 0.10 * ref.max(), name
    assert angles[int(np.argmax(om))] == 30.


def test_one_hits_two_matches_oracle(gpu_lib, tmp_path):
    res = run_driver("one_hits_two", "tst_03_multiple_particles_contacts.json", tmp_path)
    st, w, prm, run, inp = S.tst03()
    solver_run(st, w, prm, run)
    compare_to_oracle(res, st)
    g = state_arrays(res)
    p = (st.m[:, None] * g["u"]).sum(0)
    assert abs(p[0] - st.m[0] * inp["velocity_p"]) < 1e-12 * st.m[0] * inp["velocity_p"]


BONDED_EXE = ROOT / "examples" / "_build" / "bonded_tensile"


@pytest.mark.parametrize("mode,final_time", [("staged", 0.5), ("solver", 4.0)])
def test_bonded_tensile_dropin_matches_oracle(gpu_lib, tmp_path, mode, final_time):
    """examples/tst_01_BDEM_two_particles_tensile_3D.cpp through the drop-in
    bonded API (examples/bonded_tensile.cpp): the reference's staged loop with
    host-side force edits, and the fused solver with the same physics set once."""
    import oracle_py as O
    from oracle_bonded import bonded_run

    assert BONDED_EXE.exists(), "build() compiles examples/_build/bonded_tensile"
    cmd = [str(BONDED_EXE), mode, repr(final_time), str(tmp_path)]
    print("driver:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    res = json.loads(p.stdout.strip().splitlines()[-1])
    sc = S.tst01_bonded()
    st = sc["st"]
    nsteps = int(final_time / sc["dt"]) + 1
    assert res["num_steps"] + 1 == nsteps
    ob = O.setup_bonds(st, sc["factor"], sc["cutoff"], sc["lo"], sc["hi"], sc["max_bonds"])
    bonded_run(st, ob, nsteps=nsteps, prm=O.params(dt=sc["dt"]), bp=O.bond_params(beta=0.1),
               dt=sc["dt"], fext=sc["fext"], group=sc["group"], kinds=sc["kinds"])
    g = state_arrays(res)
    assert normwise(g["x"], st.x) <= 1e-10
    assert normwise(g["u"], st.u) <= 1e-10
    assert [q["bonds"] for q in res["particles"]] == [1, 1]
    assert (tmp_path / "particles_0.csv").exists()
