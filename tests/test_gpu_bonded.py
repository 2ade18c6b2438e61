"""Bonded DEM on the GPU against the oracle (src/CabanaBondedDEM_*.hpp).

Bond sets and rest lengths bit-exact; forces, torques and bond histories
within 1e-12 normwise on identical states; runs (pulled faces, breakage,
contacts between unbonded pairs, the reference's two-particle tensile
driver) within 1e-10 normwise with identical broken-bond sets.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_py as O
from oracle_bonded import GROUP_FIXED_FORCE, GROUP_VELOCITY, bonded_run
from cabanadem_amd import abi
from gpu_helpers import make_system, normwise, oracle_params
from scenarios import bonded_block, tst01_bonded

pytestmark = pytest.mark.gpu


def oracle_bond_map(b: O.Bonds):
    out = {}
    for i in range(b.n):
        for k in range(b.count[i]):
            h = np.concatenate([b.ft[i, k], b.mn[i, k], b.mt[i, k]])
            out[(i, int(b.idx[i, k]))] = (b.init_len[i, k], h)
    return out


def gpu_bond_map(b: abi.BondState):
    hist = b.history()
    out = {}
    for i in range(b.n):
        for k in range(b.count[i]):
            out[(i, int(b.idx[i, k]))] = (b.init_len[i, k], hist[i, k])
    return out


def oracle_to_view(b: O.Bonds) -> abi.BondState:
    v = abi.BondState(b.n, b.max_bonds)
    v.idx[:] = b.idx
    v.count[:] = b.count
    v.init_len[:] = b.init_len
    for q, (arr, c) in enumerate([(b.ft, 0), (b.ft, 1), (b.ft, 2), (b.mn, 0), (b.mn, 1), (b.mn, 2),
                                  (b.mt, 0), (b.mt, 1), (b.mt, 2)]):
        getattr(v, abi.BondState.HIST[q])[:] = arr[:, :, c]
    return v


def _block(seed, jitter=0.05e-3):
    rng = np.random.default_rng(seed)
    st, lo, hi = bonded_block(10, 9, 8, jitter=jitter, seed=seed, E=1e7)
    st.rad[:] = rng.uniform(0.46e-3, 0.54e-3, size=st.n)
    st.E[:] = rng.uniform(0.8e7, 1.2e7, size=st.n)
    st.G[:] = st.E / 2.5
    return st, lo, hi


@pytest.mark.parametrize("seed", [0, 3])
def test_setup_bonds_bit_exact(gpu_lib, seed):
    st, lo, hi = _block(seed)
    factor = 2.3
    rc = factor * st.rad.max() * 1.01
    ob = O.setup_bonds(st, factor, rc, lo, hi, 16)
    sysm = make_system(st, None, dt=1e-6, rc=rc, lo=lo, hi=hi)
    sysm.enable_bonds(16)
    sysm.setup_bonds(factor)
    gb = sysm.download_bonds()
    om, gm = oracle_bond_map(ob), gpu_bond_map(gb)
    assert om.keys() == gm.keys()
    assert all(om[k][0] == gm[k][0] for k in om)
    assert all(not gm[k][1].any() for k in gm)
    assert (gb.count == ob.count).all()
    sysm.close()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("strength", [0.0, 6e3])
def test_bond_forces_identical_state(gpu_lib, mode, strength):
    st, lo, hi = _block(7)
    rng = np.random.default_rng(11)
    factor = 2.3
    rc = factor * st.rad.max() * 1.01
    ob = O.setup_bonds(st, factor, rc, lo, hi, 16)
    # live state: moved positions (stretched and compressed bonds), spins,
    # velocities and non-zero bond histories
    st.x += rng.uniform(-0.02e-3, 0.02e-3, size=st.x.shape)
    st.u[:] = rng.normal(0., 0.05, size=st.u.shape)
    st.omega[:] = rng.normal(0., 50., size=st.omega.shape)
    for a in (ob.ft, ob.mn, ob.mt):
        a[:] = rng.normal(0., 1e-3, size=a.shape)
    bp = O.bond_params(beta=0.1, strength=strength, mode=mode)
    sysm = make_system(st, None, dt=1e-6, rc=rc, lo=lo, hi=hi, mode=mode)
    sysm.enable_bonds(16, beta_bond=0.1, tensile_strength=strength)
    sysm.upload_bonds(oracle_to_view(ob))
    sysm.zero_forces()
    sysm.bond_forces()
    g = sysm.download()
    gb = sysm.download_bonds()
    O.zero_forces(st, O.params(dt=1e-6, mode=mode))
    nbroken = O.bond_force(st, ob, bp, 1e-6)
    assert normwise(g.force, st.force) < 1e-12
    assert normwise(g.torque, st.torque) < 1e-12
    om, gm = oracle_bond_map(ob), gpu_bond_map(gb)
    assert om.keys() == gm.keys()  # same broken bonds (partner -1 at the same slots)
    live = [k for k in om if k[1] >= 0]
    H_o = np.array([om[k][1] for k in live])
    H_g = np.array([gm[k][1] for k in live])
    assert normwise(H_g, H_o) < 1e-12
    assert sysm.counters().bonds_broken == nbroken
    if strength > 0:
        assert 0 < nbroken < sum(ob.count)
    sysm.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_pulled_block_run(gpu_lib, mode):
    """Simple-cubic block with face bonds, x-end faces pulled apart at
    +-v, breakage on, Hertz contacts between unbonded pairs."""
    nx, ny, nz = 8, 6, 6
    st, lo, hi = bonded_block(nx, ny, nz, E=1e7, jitter=0.01e-3, seed=5)
    dt, v, S, nsteps = 1e-6, 0.05, 2.5e4, 400
    factor, rc = 2.1, 1.2e-3
    lo = lo - 1e-3
    hi = hi + 1e-3
    xs = st.x[:, 0]
    group = np.where(xs < 0.5e-3, 1, np.where(xs > (nx - 1.5) * 1e-3, 2, 0)).astype(np.int32)
    kinds, vel = [GROUP_VELOCITY, GROUP_VELOCITY], [(-v, 0., 0.), (v, 0., 0.)]
    ob = O.setup_bonds(st, factor, rc, lo, hi, 8)
    sysm = make_system(st, None, dt=dt, rc=rc, lo=lo, hi=hi, mode=mode, max_contacts=8)
    sysm.enable_bonds(8, beta_bond=0.1, tensile_strength=S, unbonded_contacts=True)
    sysm.setup_bonds(factor)
    sysm.set_groups(group, kinds, vel)
    nb_o = bonded_run(st, ob, nsteps=nsteps, prm=oracle_params(dt, mode=mode),
                      bp=O.bond_params(beta=0.1, strength=S, mode=mode), dt=dt,
                      group=group, kinds=kinds, vel=vel, contacts=True, rc=rc, lo=lo, hi=hi)
    sysm.step(nsteps)
    g = sysm.download()
    gb = sysm.download_bonds()
    assert nb_o > 0, "the run must exercise breakage"
    assert sysm.counters().bonds_broken == nb_o
    # broken slots hold partner -1, so compare the sets of intact bonds
    live_o = {k for k in oracle_bond_map(ob) if k[1] >= 0}
    live_g = {k for k in gpu_bond_map(gb) if k[1] >= 0}
    assert live_o == live_g
    assert normwise(g.x, st.x) < 1e-10
    assert normwise(g.u, st.u) < 1e-10
    assert normwise(g.omega, st.omega) < 1e-10
    sysm.close()


def test_tst01_two_particle_tensile_driver(gpu_lib):
    """examples/tst_01_BDEM_two_particles_tensile_3D.cpp end to end: setupBonds,
    then stage1, stage2, reset, +external force, bond forces, F_0 = 0, stage3,
    for num_steps + 1 = int(4 / 1e-4) + 1 steps (src/CabanaDEM_Input.hpp:53 style
    truncation).  The reference's own trajectory reads the rest length of
    particle 0's bond at index j = 1 (quirk Q9); both sides here use the slot."""
    sc = tst01_bonded()
    st = sc["st"]
    nsteps = int(sc["final_time"] / sc["dt"]) + 1
    ob = O.setup_bonds(st, sc["factor"], sc["cutoff"], sc["lo"], sc["hi"], sc["max_bonds"])
    sysm = make_system(st, None, dt=sc["dt"], rc=sc["cutoff"], lo=sc["lo"], hi=sc["hi"],
                       rebuild_every=1 << 30)
    sysm.enable_bonds(sc["max_bonds"], beta_bond=0.1)
    sysm.setup_bonds(sc["factor"])
    sysm.set_external_force(sc["fext"])
    sysm.set_groups(sc["group"], sc["kinds"])
    xs_o = []
    bonded_run(st, ob, nsteps=nsteps, prm=O.params(dt=sc["dt"]), bp=O.bond_params(beta=0.1),
               dt=sc["dt"], fext=sc["fext"], group=sc["group"], kinds=sc["kinds"],
               record=lambda s, p, bb: xs_o.append(p.x[1, 0]) if (s + 1) % 1000 == 0 else None)
    xs_g = []
    done = 0
    while done < nsteps:
        k = min(1000, nsteps - done)
        sysm.step(k)
        done += k
        if done % 1000 == 0:
            xs_g.append(sysm.download().x[1, 0])
    g = sysm.download()
    assert normwise(np.array(xs_g), np.array(xs_o)) < 1e-10
    assert normwise(g.x, st.x) < 1e-10
    assert normwise(g.u, st.u) < 1e-10
    assert np.all(g.x[0] == (-0.5, 0., 0.))
    sysm.close()


def test_bond_api_errors(gpu_lib):
    st, lo, hi = bonded_block(3, 3, 3, E=1e7)
    sysm = make_system(st, None, dt=1e-6, rc=1.2e-3, lo=lo, hi=hi)
    with pytest.raises(abi.DemError) as e:
        sysm.setup_bonds(2.1)
    assert e.value.code == -8
    sysm.enable_bonds(2)
    with pytest.raises(abi.DemError) as e:
        sysm.setup_bonds(3.0)  # 3 r > cutoff
    assert e.value.code == -1
    with pytest.raises(abi.DemError) as e:
        sysm.setup_bonds(2.1)  # the centre sphere has 6 face neighbours
    assert e.value.code == -7
    with pytest.raises(abi.DemError) as e:
        sysm.enable_bonds(4)
    assert e.value.code == -1
    with pytest.raises(abi.DemError) as e:
        sysm.set_groups(np.full(st.n, 3, dtype=np.int32), [0, 0], [(0, 0, 0), (0, 0, 0)])
    assert e.value.code == -1
    # bonded systems run force tiles of 128; cell tiles may change
    with pytest.raises(abi.DemError) as e:
        sysm.set_shape(64, 0)
    assert e.value.code == -1
    sysm.set_shape(0, 16)
    assert sysm.shape() == (128, 16)
    sysm.close()
    # a context with 64-slot tiles falls back to 128 once bonds are enabled
    sysm = make_system(st, None, dt=1e-6, rc=1.2e-3, lo=lo, hi=hi)
    sysm.set_shape(64, 0)
    assert sysm.shape() == (64, 8)
    sysm.enable_bonds(8)
    assert sysm.shape()[0] == 128
    sysm.close()
