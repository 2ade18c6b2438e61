"""Closed-form pins of the particle-wall tangential model
(src/CabanaDEM_Force.hpp:588-655) on the CPU oracle.

A sphere resting on a floor at its static overlap and sliding at u0 feels
kinetic friction mu m g once its tangential spring has loaded (a few steps),
so u(t) = u0 - mu g t and omega(t) = 5 mu g t / (2 r) about the axis n x u;
a sphere spinning in place drifts forward and spins down, and in both the
kinetic energy only falls.  The reference's contact point sits on the wrong
side (quirk Q5), which flips the sign of the spin it builds up; its normal
damping of the limited force is inert there (eta_t = 0, quirk Q2), while in
corrected mode damping is on and the spring keeps only the elastic part
(quirk Q16: the reference adds +eta_t v_t to the applied force).
"""
import math

import numpy as np
import pytest

import oracle_py as O
from cabanadem_amd.abi import ParticleState, Walls
from scenarios import _sphere

R, RHO, E, NU = 1e-3, 2600., 1e7, 0.3
G_MOD = E / (2. * (1. + NU))
MU, G = 0.3, 9.81


def _floor_sphere(u0, w0):
    st = ParticleState(1, max_cnts=4, max_walls=1)
    _sphere(st, 0, R, RHO, E, G_MOD, NU)
    # static overlap: m g = 4/3 E* sqrt(R) d^1.5 with E* of sphere and wall
    e_star = 1. / (2. * (1. - NU * NU) / E)
    kn = 4. / 3. * e_star * math.sqrt(R)
    d = (st.m[0] * G / kn) ** (2. / 3.)
    st.x[0] = (0., 0., R - d)
    st.u[0] = u0
    st.omega[0] = w0
    w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD)
    return st, w


def _run(st, w, mode, steps, dt=1e-6):
    prm = O.params(dt=dt, cor_pp=0.8, cor_pw=0.8, fr_pp=MU, fr_pw=MU, gravity=(0., 0., -G),
                   mode=mode)
    c = np.zeros(1, dtype=np.int32)
    ids = np.zeros((1, 4), dtype=np.int32)
    ke = []
    for _ in range(steps):
        O.stage(st, 1, dt)
        O.stage(st, 2, dt)
        O.compute_forces(st, w, c, ids, prm)
        O.stage(st, 3, dt)
        ke.append(0.5 * st.m[0] * (st.u[0] ** 2).sum() + 0.5 * st.I[0] * (st.omega[0] ** 2).sum())
    return np.array(ke)


@pytest.mark.parametrize("mode,spin_sign", [(0, -1.), (1, 1.)])
def test_sliding_sphere_kinetic_friction(mode, spin_sign):
    st, w = _floor_sphere((0.1, 0., 0.), (0., 0., 0.))
    steps = 3000
    ke = _run(st, w, mode, steps)
    t = steps * 1e-6
    assert st.u[0, 0] == pytest.approx(0.1 - MU * G * t, rel=2e-3)
    assert st.omega[0, 1] == pytest.approx(spin_sign * 5. * MU * G * t / (2. * R), rel=0.06)
    assert abs(st.u[0, 1]) < 1e-12 and abs(st.omega[0, 0]) < 1e-9
    assert np.all(np.diff(ke) <= 1e-18)  # kinetic energy never grows


@pytest.mark.parametrize("mode", [0, 1])
def test_spinning_sphere_spins_down(mode):
    st, w = _floor_sphere((0., 0., 0.), (0., 100., 0.))
    ke = _run(st, w, mode, 3000)
    assert np.all(np.diff(ke) <= 1e-18)
    assert 0. < abs(st.omega[0, 1]) < 100.
    if mode == 1:
        # the bottom point moves -x, friction pushes the sphere +x
        assert st.u[0, 0] > 0.
