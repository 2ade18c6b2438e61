"""Reference validation set-ups, restated from the reference's example drivers.

Each builder returns (state, walls, params_kw, run_kw) in the reference layout:
state is a cabanadem_amd.abi.ParticleState, walls a cabanadem_amd.abi.Walls,
params_kw the Force/Integrator scalars and run_kw the solver scalars
(cutoff `delta`, mesh box, number of steps).  Paths are relative to the
ExaPhysics/CabanaDEM root; the JSON inputs are copies under tests/golden/inputs.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

from cabanadem_amd.abi import ParticleState, Walls

GOLDEN = Path(__file__).resolve().parent / "golden"


def inputs(name: str) -> dict:
    """CabanaDEM::Inputs (src/CabanaDEM_Input.hpp:44-55): values by key, plus
    num_steps = int(final_time / timestep) (truncating)."""
    raw = json.loads((GOLDEN / "inputs" / name).read_text())
    vals = {k: v["value"] for k, v in raw.items()}
    vals["num_steps"] = int(vals["final_time"] / vals["timestep"])
    return vals


def _sphere(st: ParticleState, pid: int, r: float, rho: float, E: float, G: float, nu: float):
    # examples/01_elastic_normal_impact_of_two_identical_particles.cpp:106-107,117-123
    m = 4. / 3. * math.pi * r * r * r * rho
    st.m[pid] = m
    st.I[pid] = 2. / 5. * m * r * r
    st.rad[pid] = r
    st.E[pid], st.G[pid], st.nu[pid] = E, G, nu


def ex01():
    """examples/01_elastic_normal_impact_of_two_identical_particles.cpp:34-188."""
    inp = inputs("01_elastic_normal_impact_of_two_identical_particles.json")
    r, v = inp["particle_radius"], inp["velocity_p"]
    st = ParticleState(2, max_cnts=6, max_walls=1)
    for pid in range(2):
        _sphere(st, pid, r, inp["particle_density"], inp["particle_youngs_modulus"],
                inp["particle_shear_modulus"], inp["particle_poissons_ratio"])
        st.x[pid] = (-r + pid * (2. * r) + pid * r / 100., 0., 0.)   # :109-111
        st.u[pid] = (v if pid == 0 else -v, 0., 0.)                   # :112-115
    w = Walls([(0., -3. * r, 0.)], [(0., 1., 0.)], inp["wall_youngs_modulus"],
              inp["wall_poissons_ratio"], inp["wall_shear_modulus"])  # :155-168
    prm = dict(dt=inp["timestep"], cor_pp=inp["coefficient_of_restitution_pp"],
               cor_pw=inp["coefficient_of_restitution_pw"], fr_pp=inp["friction_pp"],
               fr_pw=inp["friction_pw"])
    run = dict(rc=3. * r, lo=(-0.04,) * 3, hi=(0.04,) * 3, nsteps=inp["num_steps"] + 1)
    return st, w, prm, run, inp


def _single_on_floor(inp, cor_pw, u0):
    r = inp["particle_radius"]
    st = ParticleState(1, max_cnts=6, max_walls=1)
    _sphere(st, 0, r, inp["particle_density"], inp["particle_youngs_modulus"],
            inp["particle_shear_modulus"], inp["particle_poissons_ratio"])
    st.x[0] = (0., r + r / 1000000., 0.)
    st.u[0] = u0
    w = Walls([(0., 0., 0.)], [(0., 1., 0.)], inp["wall_youngs_modulus"],
              inp["wall_poissons_ratio"], inp["wall_shear_modulus"])
    prm = dict(dt=inp["timestep"], cor_pp=inp["coefficient_of_restitution_pp"], cor_pw=cor_pw,
               fr_pp=inp["friction_pp"], fr_pw=inp["friction_pw"])
    return st, w, prm


def ex02():
    """examples/02_elastic_normal_impact_of_particle_wall.cpp (init :100-117, wall :146-160,
    mesh :165-171, delta 3r :177)."""
    inp = inputs("02_elastic_normal_impact_of_particle_wall.json")
    r, v = inp["particle_radius"], inp["velocity_p"]
    st, w, prm = _single_on_floor(inp, inp["coefficient_of_restitution_pw"], (0., -v, 0.))
    run = dict(rc=3. * r, lo=(-40. * r, -40. * r, -4. * r), hi=(40. * r, 40. * r, 4. * r),
               nsteps=inp["num_steps"] + 1)
    return st, w, prm, run, inp


def ex03(cor_pw: float):
    """examples/03_normal_particle_wall_different_cor.cpp (cor_pw from --cor-pw=, :37-40)."""
    inp = inputs("03_normal_particle_wall_different_cor.json")
    r, v = inp["particle_radius"], inp["velocity_p"]
    st, w, prm = _single_on_floor(inp, cor_pw, (0., -v, 0.))
    run = dict(rc=3. * r, lo=(-10. * r, -10. * r, -4. * r), hi=(10. * r, 10. * r, 4. * r),
               nsteps=inp["num_steps"] + 1)
    return st, w, prm, run, inp


def ex04(angle_deg: float):
    """examples/04_oblique_particle_wall_different_angles.cpp (--angle=, :48-52; init :117-136)."""
    inp = inputs("04_oblique_particle_wall_different_angles.json")
    r, v = inp["particle_radius"], inp["velocity_p"]
    ang = angle_deg / 180. * math.pi
    st, w, prm = _single_on_floor(inp, inp["coefficient_of_restitution_pw"],
                                  (v * math.sin(ang), -v * math.cos(ang), 0.))
    run = dict(rc=3. * r, lo=(-10. * r, -10. * r, -4. * r), hi=(10. * r, 10. * r, 4. * r),
               nsteps=inp["num_steps"] + 1)
    return st, w, prm, run, inp


def tst03():
    """examples/tst_03_multiple_particles_contacts.cpp:103-196: one sphere hits two."""
    inp = inputs("tst_03_multiple_particles_contacts.json")
    r, v = inp["particle_radius"], inp["velocity_p"]
    st = ParticleState(3, max_cnts=6, max_walls=1)
    for pid in range(3):
        _sphere(st, pid, r, inp["particle_density"], inp["particle_youngs_modulus"],
                inp["particle_shear_modulus"], inp["particle_poissons_ratio"])
    st.x[0] = (-r, 0., 0.)
    st.x[1] = (-r + 1 * (1.5 * r) + 1 * r / 100., r + r / 2., 0.)
    st.x[2] = (-r + 1 * (1.5 * r) + 1 * r / 100., -r - r / 2., 0.)
    st.u[0] = (v, 0., 0.)
    # wall normal (0, 0, 0): every particle "overlaps" it with no force (:164-166)
    w = Walls([(0., -3. * r, 0.)], [(0., 0., 0.)], inp["wall_youngs_modulus"],
              inp["wall_poissons_ratio"], inp["wall_shear_modulus"])
    prm = dict(dt=inp["timestep"], cor_pp=inp["coefficient_of_restitution_pp"],
               cor_pw=inp["coefficient_of_restitution_pw"], fr_pp=inp["friction_pp"],
               fr_pw=inp["friction_pw"])
    run = dict(rc=3. * r, lo=(-0.04,) * 3, hi=(0.04,) * 3, nsteps=inp["num_steps"] + 1)
    return st, w, prm, run, inp


# ---- closed forms (Hertz impact of elastic spheres) --------------------------
def hertz_impact(E_star, R_star, m_star, v_rel):
    """Peak overlap, peak force and contact duration of an undamped Hertz impact.

    F = kn d^1.5 with kn = 4/3 E* sqrt(R*); energy balance gives
    d_max = (5 m* v^2 / (4 kn))^(2/5) and t_c = 2 * I * d_max / v with
    I = int_0^1 dx / sqrt(1 - x^(5/2)) = 1.4716...
    """
    kn = 4. / 3. * E_star * math.sqrt(R_star)
    dmax = (5. * m_star * v_rel ** 2 / (4. * kn)) ** 0.4
    fmax = kn * dmax ** 1.5
    # integral by substitution x = 1 - s^2 to remove the end singularity
    s = np.linspace(0., 1., 200001)
    x = 1. - s * s
    f = 2. * s / np.sqrt(np.maximum(1. - x ** 2.5, 1e-300))
    f[0] = 2. / math.sqrt(2.5)  # limit s -> 0
    integral = float(np.trapezoid(f, s))
    tc = 2. * integral * dmax / v_rel
    return dmax, fmax, tc, kn


# ---- synthetic packings (workload generator shared with the benchmark) ------
def box_walls(lo, hi, E, nu, G):
    """Six inward-facing planes on a box."""
    x = [(lo[0], 0, 0), (hi[0], 0, 0), (0, lo[1], 0), (0, hi[1], 0), (0, 0, lo[2]), (0, 0, hi[2])]
    nrm = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    return Walls(x, nrm, E, nu, G)


def tst01_bonded(max_bonds: int = 27):
    """Two bonded spheres of examples/tst_01_BDEM_two_particles_tensile_3D.cpp:128-232:
    r = 0.5, rho = 2000, E = 1e4, G = 1e3, nu = 0.25, at x = -r and +r,
    binning box +-4 r, dt = 1e-4, 4 s; external force -+1e3 N along x on
    particles 0 and 1 (:287-291); particle 0's force zeroed (:57-59)."""
    r = 0.5
    st = ParticleState(2, max_cnts=6, max_walls=1)
    for i in range(2):
        _sphere(st, i, r, 2000., 1e4, 1e3, 0.25)
        st.x[i] = (-r + i * 2. * r, 0., 0.)
    fext = np.zeros((2, 3))
    fext[0, 0], fext[1, 0] = -1e3, 1e3
    return dict(st=st, r=r, dt=1e-4, final_time=4., lo=(-4. * r,) * 3, hi=(4. * r,) * 3,
                fext=fext, group=np.array([1, 0], dtype=np.int32), kinds=[1], vel=None,
                max_bonds=max_bonds, factor=3.0, cutoff=4. * r)


def bonded_block(nx: int, ny: int, nz: int, *, spacing=1e-3, r=0.5e-3, E=1e9, nu=0.25, rho=2600.,
                 jitter=0.0, seed=0, max_cnts=8):
    """Simple-cubic block of touching spheres (the bonded tensile configuration's
    shape: spacing 2r, bonds to the six face neighbours at factor 2.1)."""
    n = nx * ny * nz
    st = ParticleState(n, max_cnts=max_cnts, max_walls=1)
    rng = np.random.default_rng(seed)
    G = E / (2. * (1. + nu))
    k = 0
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                _sphere(st, k, r, rho, E, G, nu)
                st.x[k] = (ix * spacing, iy * spacing, iz * spacing)
                k += 1
    if jitter > 0:
        st.x += rng.uniform(-jitter, jitter, size=st.x.shape)
    lo = np.array([-2. * spacing] * 3)
    hi = np.array([(nx + 1) * spacing, (ny + 1) * spacing, (nz + 1) * spacing])
    return st, lo, hi
