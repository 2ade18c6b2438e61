"""Bonded-DEM step loop over the CPU oracle.

TEST INFRASTRUCTURE ONLY (tests/, smoke and the benchmark's CPU baseline):
the product path never imports it.

Order of the reference's bonded driver
(examples/tst_01_BDEM_two_particles_tensile_3D.cpp:297-309): stage1, stage2,
reset forces, + external force (:69-117), bond forces
(src/CabanaBondedDEM_Force.hpp:52-348), boundary condition (:18-66), stage3.
Extensions, in the order the library applies them: prescribed-velocity
groups after each half kick, Hertz contacts between unbonded pairs over a
fresh neighbour list and wall contacts after the bond forces.
"""
from __future__ import annotations

import numpy as np

import oracle_py as O

GROUP_VELOCITY = 0
GROUP_FIXED_FORCE = 1


def group_masks(group, kinds, n):
    """(velocity-group array for oracle_velocity_groups, fixed-force mask)."""
    if group is None:
        return None, None
    group = np.asarray(group, dtype=np.int32)
    kinds = np.asarray(kinds, dtype=np.int32)
    vg = np.where((group > 0) & (kinds[np.maximum(group, 1) - 1] == GROUP_VELOCITY), group, 0)
    ff = (group > 0) & (kinds[np.maximum(group, 1) - 1] == GROUP_FIXED_FORCE)
    return vg.astype(np.int32), ff


class _Bound:
    """ctypes views of the state, built once per run (the arrays are only
    updated in place, so the pointers stay valid)."""

    def __init__(self, st, bonds, prm, bp):
        self.L = O.lib()
        self.p = O.C.byref(O.particles(st))
        self.b = O.C.byref(bonds.struct())
        self.prm = O.C.byref(prm)
        self.bp = O.C.byref(bp)
        self.st, self.bonds = st, bonds


def bonded_step(st, bonds, *, prm, bp, dt, fext=None, vg=None, vel=None, ff=None,
                contacts=False, rc=None, lo=None, hi=None, walls=None, bound=None, nbrs=None):
    """One step; returns the number of bonds broken in it.  nbrs: a neighbour
    list (counts, ids) to reuse instead of building one for this step."""
    B = bound if bound is not None else _Bound(st, bonds, prm, bp)
    L = B.L
    L.oracle_stage1(B.p, dt)
    if vg is not None and vel is not None:
        L.oracle_velocity_groups(B.p, O._p(vg, O.C.c_int), O._p(vel), vel.shape[0])
    L.oracle_stage2(B.p, dt)
    L.oracle_zero_forces(B.p, B.prm)
    if fext is not None:
        st.force += fext
    nb = O.C.c_int(0)
    L.oracle_bond_force(B.p, B.b, B.bp, dt, O.C.byref(nb))
    if contacts:
        counts, ids = nbrs if nbrs is not None else O.neighbors_cell(st.x, rc, lo, hi)
        B.last_nbrs = (counts, ids)
        L.oracle_update_tangential(B.p)
        ov = L.oracle_force_pp_unbonded(B.p, O._p(counts, O.C.c_int), O._p(ids, O.C.c_int),
                                        ids.shape[1], B.prm, B.b)
        assert ov == 0, "tracked-contact capacity exceeded"
    if walls is not None and walls.x.shape[0] > 0:
        O.force_pw(st, walls, prm)
    if ff is not None and ff.any():
        st.force[ff] = 0.
    L.oracle_stage3(B.p, dt)
    if vg is not None and vel is not None:
        L.oracle_velocity_groups(B.p, O._p(vg, O.C.c_int), O._p(vel), vel.shape[0])
    return nb.value


def bonded_run(st, bonds, *, nsteps, prm, bp, dt, fext=None, group=None, kinds=None, vel=None,
               contacts=False, rc=None, lo=None, hi=None, walls=None, record=None,
               rebuild_every=1):
    vg, ff = group_masks(group, kinds, st.x.shape[0])
    if vel is not None:
        vel = np.ascontiguousarray(np.asarray(vel, dtype=np.float64).reshape(-1, 3))
    if fext is not None:
        fext = np.asarray(fext, dtype=np.float64)
    B = _Bound(st, bonds, prm, bp)
    broken = 0
    nbrs = None
    for step in range(nsteps):
        if contacts and rebuild_every > 1 and step % rebuild_every == 0:
            nbrs = None
        broken += bonded_step(st, bonds, prm=prm, bp=bp, dt=dt, fext=fext, vg=vg, vel=vel, ff=ff,
                              contacts=contacts, rc=rc, lo=lo, hi=hi, walls=walls, bound=B,
                              nbrs=nbrs)
        if contacts and rebuild_every > 1 and nbrs is None:
            nbrs = B.last_nbrs
        if record is not None:
            record(step, st, bonds)
    return broken
