"""Closed-form pins of the oracle's particle-particle damping and tangential
branch (src/CabanaDEM_Force.hpp:206-215 and :238-319).

* Head-on impact of two spheres of unequal mass with e < 1.  Along the line
  of centres the pair is one body of reduced mass m* = m_i m_j / (m_i + m_j)
  against a rigid plane with E*, R*; the same damping then gives the same
  restitution as the sphere-wall sweep (e_out = e within 0.02 + 2 %), and the
  two contact forces are equal and opposite, so the total momentum is kept to
  rounding.  A damping mass other than m* breaks one or the other.
* A sustained sliding contact at fixed overlap delta and fixed sliding speed
  v (no normal motion, e = 1 so no damping).  Hertz F_n = 4/3 E* sqrt(R*)
  delta^1.5 and the Mindlin stiffness k_t = 8 G* sqrt(R* delta) fix the
  tangential force of every step:
    - reference mode: k_t v dt on the first step; the history is then stored
      as a unit vector (quirk Q3), so from the second step the trial force is
      about k_t x 1 m and the Coulomb limit mu |F_n| holds it;
    - corrected mode: the stick spring k k_t v dt on step k until it reaches
      mu |F_n|, then mu |F_n|.
  A missing Coulomb cap or a missing Q3 normalisation fails these.
These tests use only the oracle (tests/test_oracle_mutants.py re-runs them
against broken copies of it and expects them to fail)."""
import math

import numpy as np
import pytest

import oracle_py as O
from cabanadem_amd.abi import ParticleState


def _sphere(st, i, r, rho, E, nu):
    m = 4. / 3. * math.pi * r ** 3 * rho
    st.rad[i], st.m[i], st.I[i] = r, m, 0.4 * m * r * r
    st.E[i], st.nu[i], st.G[i] = E, nu, E / (2. * (1. + nu))


def _stage_loop(st, prm, nsteps, rc, lo, hi):
    dt = prm.dt
    for _ in range(nsteps):
        O.stage(st, 1, dt)
        O.stage(st, 2, dt)
        c, i = O.neighbors_brute(st.x, rc)
        assert O.compute_forces(st, None, c, i, prm) == 0
        O.stage(st, 3, dt)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("cor", [0.5, 0.9])
def test_pp_head_on_unequal_masses(mode, cor):
    r0, r1 = 1.0e-3, 1.4e-3
    st = ParticleState(2, max_cnts=4, max_walls=1)
    _sphere(st, 0, r0, 2600., 1e7, 0.3)
    _sphere(st, 1, r1, 7800., 1e7, 0.3)
    assert st.m[1] / st.m[0] > 5.  # clearly unequal masses
    st.x[:] = [[0., 0., 0.], [r0 + r1 + 1e-5, 0., 0.]]
    st.u[:] = [[0.6, 0., 0.], [-0.2, 0., 0.]]
    p0 = (st.m[:, None] * st.u).sum(0)
    v_in = st.u[0, 0] - st.u[1, 0]
    prm = O.params(dt=1e-6, cor_pp=cor, fr_pp=0.3, mode=mode)
    _stage_loop(st, prm, 700, 3e-3, (-5e-3,) * 3, (5e-3,) * 3)
    gap = st.x[1, 0] - st.x[0, 0] - (r0 + r1)
    assert gap > 0, "still in contact"
    v_out = st.u[1, 0] - st.u[0, 0]
    e_out = v_out / v_in
    assert abs(e_out - cor) < 0.02 + 0.02 * cor, e_out
    # equal and opposite contact forces: momentum kept to rounding
    m_red = st.m[0] * st.m[1] / (st.m[0] + st.m[1])
    p1 = (st.m[:, None] * st.u).sum(0)
    assert np.abs(p1 - p0).max() < 1e-12 * m_red * v_in
    assert np.all(st.u[:, 1:] == 0.)


def _sliding_pair(delta, v):
    r = 1e-3
    st = ParticleState(2, max_cnts=4, max_walls=1)
    for i in range(2):
        _sphere(st, i, r, 2600., 1e7, 0.3)
    st.x[:] = [[0., 0., 0.], [0., 2. * r - delta, 0.]]  # line of centres along y
    st.u[:] = [[v, 0., 0.], [0., 0., 0.]]              # sliding along x
    return st


def _hertz_mindlin(st, delta):
    E, nu, G, r = st.E[0], st.nu[0], st.G[0], st.rad[0]
    E_star = 1. / (2. * (1. - nu * nu) / E)
    R_star = r / 2.
    G_star = 1. / (2. * (2. - nu) / G)
    Fn = 4. / 3. * E_star * math.sqrt(R_star) * delta ** 1.5
    kt = 8. * G_star * math.sqrt(R_star * delta)
    return Fn, kt


def _forces_at_fixed_state(st, prm, nsteps):
    """Force passes at a frozen state (positions and velocities fixed), so
    only the tangential history evolves."""
    c, i = O.neighbors_brute(st.x, 3e-3)
    out = []
    for _ in range(nsteps):
        assert O.compute_forces(st, None, c, i, prm) == 0
        out.append(st.force.copy())
    return np.array(out)


@pytest.mark.parametrize("mode", [0, 1])
def test_pp_sliding_contact_coulomb_limit_and_spring(mode):
    delta, v, dt, mu = 2e-5, 1.0, 1e-6, 0.4
    st = _sliding_pair(delta, v)
    Fn, kt = _hertz_mindlin(st, delta)
    lim = mu * Fn
    k_sat = math.ceil(lim / (kt * v * dt))
    assert 4 <= k_sat <= 10  # the spring needs a few steps to reach the limit
    f = _forces_at_fixed_state(st, O.params(dt=dt, cor_pp=1.0, fr_pp=mu, mode=mode), 14)
    for k in range(1, 15):
        if mode == 0:
            ft = kt * v * dt if k == 1 else lim  # Q3: unit-vector history from step 2
        else:
            ft = min(k * kt * v * dt, lim)
        F0, F1 = f[k - 1, 0], f[k - 1, 1]
        # normal: Hertz, pushing 0 away from 1 (-y); tangential: against the slip
        assert F0[1] == pytest.approx(-Fn, rel=1e-12)
        assert F0[0] == pytest.approx(-ft, rel=1e-12), (k, F0[0], -ft)
        assert F0[2] == 0.
        assert F1[0] == pytest.approx(ft, rel=1e-12)
        assert F1[1] == pytest.approx(Fn, rel=1e-12)
    # the tracked history: a unit vector (reference) or the stick/slip spring
    h = np.array([st.ss_x[0, 0], st.ss_y[0, 0], st.ss_z[0, 0]])
    if mode == 0:
        assert np.linalg.norm(h) == pytest.approx(1.0, rel=1e-14)
    else:
        assert h[0] == pytest.approx(lim / kt, rel=1e-12)


def test_pp_corrected_spring_keeps_only_the_elastic_part():
    # With damping on (e < 1) the stick force is k k_t v dt + |eta_t| v: the
    # damping part is the same every step, so consecutive forces differ by
    # exactly the spring increment k_t v dt (the history does not absorb it).
    delta, v, dt = 2e-5, 0.2, 1e-6
    st = _sliding_pair(delta, v)
    _, kt = _hertz_mindlin(st, delta)
    f = _forces_at_fixed_state(st, O.params(dt=dt, cor_pp=0.5, fr_pp=10.0, mode=1), 8)
    ft = -f[:, 0, 0]
    assert ft[0] > kt * v * dt * 1.01  # the damping adds to the first step's force
    np.testing.assert_allclose(np.diff(ft), kt * v * dt, rtol=1e-9)
