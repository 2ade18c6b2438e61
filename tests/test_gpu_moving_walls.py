"""Moving walls on the GPU against the oracle (corrected mode; the reference
reads wall velocity and spin and leaves them unused,
src/CabanaDEM_Force.hpp:468-469,529).

A single sphere on a belt, on a turntable and hit by a rising floor (the
scenarios whose closed forms tests/test_oracle_moving_walls.py checks on
the oracle) match the oracle within 1e-10; a box of spheres with a sinking
lid and a spinning floor matches it within the run bar; in reference mode
moving walls equal static ones bitwise; captured graphs fall back to eager
steps while walls translate."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from cabanadem_amd import abi
from cabanadem_amd.abi import ParticleState, Walls
from gpu_helpers import granular_box, make_system, normwise
from scenarios import _sphere

pytestmark = pytest.mark.gpu

R, RHO, E, NU = 1e-3, 2600., 1e7, 0.3
G_MOD = E / (2. * (1. + NU))
DT = 1e-6


def _one(z, u=(0., 0., 0.)):
    st = ParticleState(1, max_cnts=4, max_walls=1)
    _sphere(st, 0, R, RHO, E, G_MOD, NU)
    st.x[0] = (0., 0., z)
    st.u[0] = u
    return st


def _pair_run(st, w, steps, mode, gravity, cor=0.5, mu=0.3, rc=3e-3, lo=(-5e-3,) * 3,
              hi=(5e-3,) * 3, chunks=(1, 37, 400)):
    """The same run on the GPU (dem_step in chunks) and in the oracle."""
    ref = st.copy()
    prm = O.params(dt=DT, cor_pp=cor, cor_pw=cor, fr_pp=mu, fr_pw=mu, gravity=gravity, mode=mode)
    O.run(ref, w, prm, nsteps=steps, rc=rc, lo=lo, hi=hi, max_n=64)
    sysm = make_system(st, w, dt=DT, rc=rc, lo=lo, hi=hi, cor_pp=cor, cor_pw=cor, fr_pp=mu,
                       fr_pw=mu, gravity=gravity, mode=mode)
    left = steps
    k = 0
    while left > 0:
        c = min(chunks[k % len(chunks)], left)
        sysm.step(c)
        left -= c
        k += 1
    got = sysm.download()
    sysm.close()
    return got, ref


@pytest.mark.parametrize("case", ["belt", "turntable", "rising_floor"])
def test_single_sphere_matches_oracle(gpu_lib, case):
    if case == "rising_floor":
        st = _one(R + 0.1e-3)
        w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, u=(0., 0., 0.2))
        g = (0., 0., 0.)
    else:
        st = _one(R - 3e-6)
        g = (0., 0., -9.81)
        if case == "belt":
            w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, u=(0.1, 0.02, 0.))
        else:
            w = Walls([(-1., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, omega=(0.02, 0., 0.1))
    got, ref = _pair_run(st, w, 3000, abi.MODE_CORRECTED, g)
    assert np.abs(ref.u).max() > 1e-3  # the wall moved the sphere
    for f in ("x", "u", "omega"):
        assert normwise(getattr(got, f), getattr(ref, f)) <= 1e-10, f


def _lid_box(seed=4):
    st, w, lo, hi = granular_box(8, 8, 8, spacing_r=2.05, jitter_r=0.02, vsigma=0.1, seed=seed,
                                 wall_offset=0.99e-3)
    u = np.zeros((6, 3))
    om = np.zeros((6, 3))
    u[5] = (0., 0., -0.3)      # the lid (top wall, normal -z) sinks
    om[4] = (0., 0., 20.)      # the floor spins about its point
    w = Walls(w.x, w.normal, w.E, w.nu, w.G, u=u, omega=om)
    return st, w, lo - 2e-3, hi + 2e-3


def test_box_with_moving_lid_and_spinning_floor(gpu_lib):
    st, w, lo, hi = _lid_box()
    g = (0., 0., -9.81)
    # 300 steps: later, stick-slip switches make a one-ulp restart of the
    # oracle itself move by 1e-3 and more
    got, ref = _pair_run(st, w, 300, abi.MODE_CORRECTED, g, cor=0.8, mu=0.4, rc=2.4e-3, lo=lo,
                         hi=hi, chunks=(100, 200))
    assert ref.ss_cnt.sum() > 50 and np.abs(ref.sw_x).max() > 0
    # the oracle's own sensitivity to a one-ulp restart bounds the per-particle gap
    pert = st.copy()
    pert.u[:] = np.nextafter(pert.u, np.inf)
    prm = O.params(dt=DT, cor_pp=0.8, cor_pw=0.8, fr_pp=0.4, fr_pw=0.4, gravity=g, mode=1)
    O.run(pert, w, prm, nsteps=300, rc=2.4e-3, lo=lo, hi=hi, max_n=64)
    bar = max(1e-8, 10. * normwise(pert.u, ref.u))
    assert normwise(got.u, ref.u) <= bar
    assert normwise(got.omega, ref.omega) <= max(1e-8, 10. * normwise(pert.omega, ref.omega))


def test_reference_mode_ignores_wall_motion_and_graphs_fall_back(gpu_lib):
    st, w, lo, hi = _lid_box(seed=5)
    static = Walls(w.x, w.normal, w.E, w.nu, w.G)
    kw = dict(dt=DT, rc=2.4e-3, lo=lo, hi=hi, gravity=(0., 0., -9.81), rebuild_every=10,
              skin=0.4e-3)
    a = make_system(st, w, mode=abi.MODE_REFERENCE, **kw)
    b = make_system(st, static, mode=abi.MODE_REFERENCE, **kw)
    a.step(300)
    b.step(300)
    ga, gb = a.download(), b.download()
    for f in ("x", "u", "omega", "ss_idx", "sw_x"):
        assert np.array_equal(getattr(ga, f), getattr(gb, f)), f
    # corrected mode with translating walls: graphs on or off, the same steps
    c = make_system(st, w, mode=abi.MODE_CORRECTED, **kw)
    d = make_system(st, w, mode=abi.MODE_CORRECTED, **kw)
    d.set_graph(True)
    for n in (15, 60):
        c.step(n)
        d.step(n)
    gc, gd = c.download(), d.download()
    for f in ("x", "u", "omega", "ss_idx", "sw_x"):
        assert np.array_equal(getattr(gc, f), getattr(gd, f)), f
    for s_ in (a, b, c, d):
        s_.close()
