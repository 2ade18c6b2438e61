"""Closed-form pins of moving walls on the CPU oracle (corrected mode).

The reference reads each wall's velocity and spin and leaves them unused
(src/CabanaDEM_Force.hpp:468-469,529): its walls never move.  In corrected
mode the plane translates with u and its surface moves with u + omega x
(c - x_w) at the contact point c, so:

* a sphere set on a belt (a floor moving along itself at V) slides, is
  accelerated by kinetic friction mu g and rolls without slipping once its
  contact point moves with the belt: u = 2V/7, omega = 5V/(7a) (linear
  momentum and angular momentum about the contact line);
* a turntable (a floor spinning about an axis at distance D from the
  sphere) is a belt of speed Omega D there, so the same 2/7 follows with
  V = Omega D: the distance is measured from the wall's point x_w;
* a floor rising at V_n hits a sphere at rest: in the floor's frame this
  is a sphere arriving at V_n, so it leaves at V_n (1 + e) with e the
  restitution (the reference's own restitution sweep bounds e_out within
  0.02 + 2 %), and contact starts when the floor has closed the gap, at
  t = gap / V_n;
* in reference mode none of this happens: the sphere never moves.
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
DT = 1e-6


def _sphere_at(z, u=(0., 0., 0.)):
    st = ParticleState(1, max_cnts=4, max_walls=1)
    _sphere(st, 0, R, RHO, E, G_MOD, NU)
    st.x[0] = (0., 0., z)
    st.u[0] = u
    return st


def _static_overlap(m):
    e_star = 1. / (2. * (1. - NU * NU) / E)
    kn = 4. / 3. * e_star * math.sqrt(R)
    return (m * G / kn) ** (2. / 3.)


def _run(st, w, steps, mode=1, gravity=(0., 0., -G), cor=0.5, record=None):
    """Single sphere over a wall, in chunks so the velocity can be sampled;
    wall_time continues from chunk to chunk."""
    prm = O.params(dt=DT, cor_pp=cor, cor_pw=cor, fr_pp=MU, fr_pw=MU, gravity=gravity, mode=mode)
    out = []
    done = 0
    chunk = 100
    while done < steps:
        k = min(chunk, steps - done)
        O.run(st, w, prm, nsteps=k, rc=3 * R, lo=(-5e-3,) * 3, hi=(5e-3,) * 3, max_n=4)
        done += k
        prm.wall_time = done * DT
        out.append((done * DT, st.u[0].copy(), st.omega[0].copy(), st.x[0].copy()))
    return out


def _settled_on_floor():
    st = _sphere_at(0.)
    st.x[0, 2] = R - _static_overlap(st.m[0])
    return st


@pytest.mark.parametrize("V", [0.1, -0.05])
def test_belt_drags_the_sphere_to_two_sevenths(V):
    st = _settled_on_floor()
    w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, u=(V, 0., 0.))
    t_roll = 2. * abs(V) / (7. * MU * G)
    rec = _run(st, w, int(3 * t_roll / DT))
    t = np.array([r[0] for r in rec])
    ux = np.array([r[1][0] for r in rec])
    wy = np.array([r[2][1] for r in rec])
    # sliding phase: u = mu g t toward the belt's motion
    k = np.searchsorted(t, 0.4 * t_roll)
    assert ux[k] == pytest.approx(math.copysign(MU * G * t[k], V), rel=0.05)
    # rolling: u = 2V/7 and the contact point moves with the belt,
    # u - a omega_y = V with a ~ R (corrected mode: contact point at -a n)
    late = t > 2. * t_roll
    assert np.mean(ux[late]) == pytest.approx(2. * V / 7., rel=0.02)
    assert np.mean(wy[late]) == pytest.approx(-5. * V / (7. * R), rel=0.02)


def test_turntable_speed_is_measured_from_the_wall_point():
    st = _settled_on_floor()
    D, Om = 1.0, 0.1  # the axis through x_w = (-D, 0, 0): surface speed Om D along +y
    w = Walls([(-D, 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, omega=(0., 0., Om))
    V = Om * D
    t_roll = 2. * V / (7. * MU * G)
    rec = _run(st, w, int(3 * t_roll / DT))
    t = np.array([r[0] for r in rec])
    uy = np.array([r[1][1] for r in rec])
    ux = np.array([r[1][0] for r in rec])
    late = t > 2. * t_roll
    assert np.mean(uy[late]) == pytest.approx(2. * V / 7., rel=0.02)
    assert abs(np.mean(ux[late])) < 1e-3 * V


@pytest.mark.parametrize("cor", [0.5, 0.8])
def test_rising_floor_launches_the_sphere(cor):
    gap, Vn = 0.1e-3, 0.2
    st = _sphere_at(R + gap)
    w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, u=(0., 0., Vn))
    rec = _run(st, w, 2000, gravity=(0., 0., 0.), cor=cor)
    t = np.array([r[0] for r in rec])
    uz = np.array([r[1][2] for r in rec])
    # contact starts once the floor has moved the gap (sampled every 100 steps)
    first = t[np.argmax(uz > 0.)]
    assert gap / Vn < first <= gap / Vn + 100 * DT + 1e-12
    e_out = uz[-1] / Vn - 1.
    assert abs(e_out - cor) <= 0.02 + 0.02 * cor
    assert st.x[0, 2] > R + Vn * t[-1]  # it flies off ahead of the floor


def test_reference_mode_walls_stay_put():
    gap, Vn = 0.1e-3, 0.2
    st = _sphere_at(R + gap)
    w = Walls([(0., 0., 0.)], [(0., 0., 1.)], E, NU, G_MOD, u=(0.3, 0., Vn), omega=(0., 0., 5.))
    _run(st, w, 2000, mode=0, gravity=(0., 0., 0.))
    assert not st.u.any() and not st.omega.any()
    st = _settled_on_floor()
    _run(st, w, 1000, mode=0)
    assert abs(st.u[0, 0]) == 0. and abs(st.u[0, 1]) == 0.
