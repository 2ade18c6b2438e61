"""Pins of the bonded-DEM oracle (src/CabanaBondedDEM_*.hpp) against brute
force, closed forms and invariants.  CPU only.

The reference's own bonded test (examples/tst_01_BDEM_two_particles_tensile_3D.cpp)
holds no expected values, and its trajectory depends on quirk Q9 (rest
length read at index j instead of slot k: particle 0 reads init_len(0, 1) = 0),
so it cannot pin the slot-indexed oracle.  The checks below are chosen so
that a dropped term, a wrong stiffness, a wrong sign or a wrong index fails
one of them:
  - set-up equals the brute-force pair set r_ij < factor r_i;
  - undamped axial bond = discrete harmonic oscillator (exact for velocity
    Verlet);
  - torsion = discrete torsional oscillator with k_tor = G J / L;
  - first-pass shear force and bending moment of a spinning partner from the
    beam stiffnesses k_t = G A / L and k_ben = E I / L, Newton's third law;
  - non-viscous normal damping: the first peak of a suddenly loaded bond is
    2 F / (k (1 -+ beta)) (reference sign, quirk Q15, vs. corrected sign);
  - tensile breakage at the closed-form step under prescribed velocity.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_py as O
from oracle_bonded import GROUP_FIXED_FORCE, GROUP_VELOCITY, bonded_run
from cabanadem_amd.abi import ParticleState
from scenarios import _sphere, bonded_block, tst01_bonded


def _pair(r=1e-3, gap=0.0, rho=2600., E=1e7, G=4e6, nu=0.25):
    st = ParticleState(2, max_cnts=6, max_walls=1)
    for i in range(2):
        _sphere(st, i, r, rho, E, G, nu)
    st.x[0] = (0., 0., 0.)
    st.x[1] = (2. * r + gap, 0., 0.)
    return st


def _bond_pair(st, r, max_bonds=4, factor=3.0):
    lo, hi = (-8 * r,) * 3, (8 * r,) * 3
    b = O.setup_bonds(st, factor, 4. * r, lo, hi, max_bonds)
    assert list(b.count) == [1, 1] and b.idx[0, 0] == 1 and b.idx[1, 0] == 0
    return b


def _beam(r_b, L, E, G):
    A = math.pi * r_b ** 2
    I = math.pi * r_b ** 4 / 4.
    return dict(k_n=E * A / L, k_t=G * A / L, k_tor=G * 2. * I / L, k_ben=E * I / L)


@pytest.mark.parametrize("seed", [0, 1])
def test_setup_bonds_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    st, lo, hi = bonded_block(6, 5, 4, jitter=0.08e-3, seed=seed)
    st.rad[:] = rng.uniform(0.45e-3, 0.55e-3, size=st.n)
    factor = 2.3
    b = O.setup_bonds(st, factor, factor * st.rad.max() * 1.01, lo, hi, 16)
    got = {}
    for i in range(st.n):
        for k in range(b.count[i]):
            got[(i, int(b.idx[i, k]))] = b.init_len[i, k]
    want = {}
    for i in range(st.n):
        for j in range(st.n):
            if i == j:
                continue
            p = st.x[i] - st.x[j]
            rij = math.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
            if rij < factor * st.rad[i]:
                want[(i, j)] = rij
    assert got.keys() == want.keys()
    assert all(got[k] == want[k] for k in want)  # bit-exact rest lengths
    # every row is filled in slot order, unused slots empty
    for i in range(st.n):
        assert (b.idx[i, b.count[i]:] == -1).all()


def test_setup_bonds_reports_overflow():
    st, lo, hi = bonded_block(3, 3, 3)
    b = O.Bonds(st.n, 2)
    x = np.ascontiguousarray(st.x)
    rc = O.lib().oracle_setup_bonds(st.n, O._p(x), O._p(st.rad), 2.1, 1.2e-3,
                                    O._p(np.asarray(lo, dtype=np.float64)),
                                    O._p(np.asarray(hi, dtype=np.float64)), O.C.byref(b.struct()))
    assert rc == -3  # the centre sphere has 6 face neighbours > 2 slots
    assert (b.count <= 2).all()


@pytest.mark.parametrize("mode", [0, 1])
def test_axial_bond_is_discrete_harmonic_oscillator(mode):
    r, dt, Fe = 1e-3, 1e-6, 1e-2
    st = _pair(r)
    b = _bond_pair(st, r)
    k = _beam(r, 2. * r, st.E[1], st.G[1])["k_n"]
    m = st.m[1]
    delta = Fe / k
    theta = math.acos(1. - (dt * dt * k / m) / 2.)
    st.force[1, 0] = Fe  # consistent initial force: no half-kick lag
    fext = np.zeros((2, 3))
    fext[1, 0] = Fe
    x0 = st.x[1, 0]
    xs = []
    bonded_run(st, b, nsteps=400, prm=O.params(dt=dt, mode=mode),
               bp=O.bond_params(beta=0.0, mode=mode), dt=dt, fext=fext,
               group=np.array([1, 0]), kinds=[GROUP_FIXED_FORCE],
               record=lambda s, p, bb: xs.append(p.x[1, 0] - x0))
    n = np.arange(1, 401)
    want = delta * (1. - np.cos(n * theta))
    assert np.max(np.abs(np.array(xs) - want)) < 1e-8 * delta
    assert st.x[0, 0] == 0. and np.all(st.x[:, 1:] == 0.)


@pytest.mark.parametrize("mode", [0, 1])
def test_torsion_is_discrete_torsional_oscillator(mode):
    r, dt, w0 = 1e-3, 1e-6, 3.0
    st = _pair(r)
    b = _bond_pair(st, r)
    st.omega[1] = (w0, 0., 0.)
    k_tor = _beam(r, 2. * r, st.E[1], st.G[1])["k_tor"]
    Omega2 = k_tor / st.I[1]
    phi = math.acos(1. - dt * dt * Omega2 / 2.)
    amp = w0 * dt / math.sin(phi)
    mx = []
    bonded_run(st, b, nsteps=300, prm=O.params(dt=dt, mode=mode),
               bp=O.bond_params(beta=0.0, mode=mode), dt=dt,
               group=np.array([1, 0]), kinds=[GROUP_VELOCITY], vel=[(0., 0., 0.)],
               record=lambda s, p, bb: mx.append(bb.mn[1, 0, 0]))
    n = np.arange(1, 301)
    want = -k_tor * amp * np.sin(n * phi)
    assert np.max(np.abs(np.array(mx) - want)) < 1e-9 * k_tor * amp
    # pure torsion: no bending moment, no shear force, no motion of the centres
    assert np.all(b.mt == 0.) and np.all(b.ft == 0.)
    assert np.all(st.x[1] == (2. * r, 0., 0.))


@pytest.mark.parametrize("mode", [0, 1])
def test_first_pass_shear_and_bending(mode):
    r, gap, dt, w = 1e-3, 0.1e-3, 1e-6, 2.0
    st = _pair(r, gap=gap, E=1e7, G=3e6)
    b = _bond_pair(st, r)
    st.omega[1] = (0., w, 0.)
    prm = O.params(dt=dt, mode=mode)
    O.zero_forces(st, prm)
    O.bond_force(st, b, O.bond_params(beta=0.0, mode=mode), dt)
    kk = _beam(r, 2. * r, 1e7, 3e6)
    a = r + gap / 2.  # lever arm: radius plus half the gap
    # i = 1: contact-point velocity (w y) x (-x) a = +w a z... relative to a
    # resting partner; shear spring -k_t v_t dt, bending -k_ben w_t dt
    np.testing.assert_allclose(st.force[1], (0., 0., w * a * kk["k_t"] * dt), rtol=1e-14, atol=0)
    np.testing.assert_allclose(st.torque[1], (0., -w * kk["k_ben"] * dt, 0.), rtol=1e-14, atol=0)
    # third law on the pair's forces and bending moments
    np.testing.assert_allclose(st.force[0], -st.force[1], rtol=1e-14, atol=0)
    np.testing.assert_allclose(st.torque[0], -st.torque[1], rtol=1e-14, atol=0)


@pytest.mark.parametrize("mode,sign", [(0, -1.0), (1, +1.0)])
def test_normal_damping_first_peak(mode, sign):
    """Suddenly applied load on a resting bond: while the bond stretches the
    damped spring acts as k (1 + sign beta), so the first peak is
    2 F / (k (1 + sign beta)).  The reference adds +beta |f_n| sign(v_n)
    (src/CabanaBondedDEM_Force.hpp:220-222, quirk Q15: sign = -1)."""
    r, dt, Fe, beta = 1e-3, 2e-7, 1e-2, 0.1
    st = _pair(r)
    b = _bond_pair(st, r)
    k = _beam(r, 2. * r, st.E[1], st.G[1])["k_n"]
    st.force[1, 0] = Fe
    fext = np.zeros((2, 3))
    fext[1, 0] = Fe
    x0 = st.x[1, 0]
    xs = []
    period = 2. * math.pi * math.sqrt(st.m[1] / k)
    bonded_run(st, b, nsteps=int(0.75 * period / dt), prm=O.params(dt=dt, mode=mode),
               bp=O.bond_params(beta=beta, mode=mode), dt=dt, fext=fext,
               group=np.array([1, 0]), kinds=[GROUP_FIXED_FORCE],
               record=lambda s, p, bb: xs.append(p.x[1, 0] - x0))
    peak = max(xs)
    want = 2. * Fe / (k * (1. + sign * beta))
    assert abs(peak - want) < 2e-3 * want
    other = 2. * Fe / (k * (1. - sign * beta))
    assert abs(peak - other) > 0.1 * want


@pytest.mark.parametrize("mode", [0, 1])
def test_breakage_step_under_prescribed_velocity(mode):
    r, dt, v, E = 1e-3, 1e-6, 0.05, 1e7
    st = _pair(r, E=E)
    b = _bond_pair(st, r)
    # break once E (r_ij - L0) / (2 r) > S with r_ij - L0 = k v dt after k drifts
    kstar = 100.5
    S = E * kstar * v * dt / (2. * r)
    bp = O.bond_params(beta=0.1, strength=S, mode=mode)
    broke = []
    bonded_run(st, b, nsteps=110, prm=O.params(dt=dt, mode=mode), bp=bp, dt=dt,
               group=np.array([1, 2]), kinds=[GROUP_VELOCITY, GROUP_VELOCITY],
               vel=[(0., 0., 0.), (v, 0., 0.)],
               record=lambda s, p, bb: broke.append(int((bb.idx[:, 0] < 0).sum())))
    first = broke.index(2)
    assert first == 100  # step index 100 = the 101st drift
    assert broke[99] == 0
    # prescribed velocity: exact drift, no spin
    np.testing.assert_allclose(st.x[1, 0], 2. * r + 110 * v * dt, rtol=1e-13)
    assert np.all(st.omega == 0.)


def test_tst01_scenario_runs_and_stays_axial():
    sc = tst01_bonded()
    st = sc["st"]
    lo, hi = sc["lo"], sc["hi"]
    b = O.setup_bonds(st, sc["factor"], sc["cutoff"], lo, hi, sc["max_bonds"])
    assert list(b.count) == [1, 1]
    bonded_run(st, b, nsteps=2000, prm=O.params(dt=sc["dt"]), bp=O.bond_params(beta=0.1),
               dt=sc["dt"], fext=sc["fext"], group=sc["group"], kinds=sc["kinds"])
    # particle 0 is held by the zeroed force; the pair stays on the x axis
    assert np.all(st.x[0] == (-0.5, 0., 0.))
    assert np.all(st.x[:, 1:] == 0.)
    assert st.x[1, 0] > 0.5
