"""ctypes binding of the CPU oracle (oracle/dem_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker.  The product package
(cabanadem_amd) never imports it.

State objects are duck-typed: anything with the reference-layout numpy
attributes x, u, omega, force, torque [n,3]; m, rad, E, nu, G, I [n];
sw_x/y/z [n, max_walls]; ss_x/y/z [n, max_cnts] (float64); ss_idx
[n, max_cnts] and ss_cnt [n] (int32).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# ORACLE_LIB points the binding at another build of dem_oracle.cpp; the
# mutation tests (tests/test_oracle_mutants.py) use it to run the closed-form
# checks against deliberately broken copies of the oracle.
LIB = Path(os.environ["ORACLE_LIB"]) if os.environ.get("ORACLE_LIB") else HERE / "_build" / "libdem_oracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OracleParticles(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("max_cnts", C.c_int), ("max_walls", C.c_int),
        ("x", _dp), ("u", _dp), ("omega", _dp), ("force", _dp), ("torque", _dp),
        ("m", _dp), ("rad", _dp), ("E", _dp), ("nu", _dp), ("G", _dp), ("I", _dp),
        ("sw_x", _dp), ("sw_y", _dp), ("sw_z", _dp),
        ("ss_x", _dp), ("ss_y", _dp), ("ss_z", _dp),
        ("ss_idx", _ip), ("ss_cnt", _ip),
    ]


class OracleWalls(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _dp), ("normal", _dp), ("E", _dp), ("nu", _dp), ("G", _dp),
                ("u", _dp), ("omega", _dp)]


class OracleParams(C.Structure):
    _fields_ = [("cor_pp", C.c_double), ("cor_pw", C.c_double), ("fr_pp", C.c_double),
                ("fr_pw", C.c_double), ("dt", C.c_double), ("gravity", C.c_double * 3),
                ("mode", C.c_int), ("wall_time", C.c_double)]


class OracleRun(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("rc", C.c_double), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("rebuild_every", C.c_int), ("max_n", C.c_int),
                ("diag_every", C.c_int), ("threads", C.c_int)]


class OracleBonds(C.Structure):
    _fields_ = [("max_bonds", C.c_int), ("idx", _ip), ("init_len", _dp), ("ft", _dp),
                ("mn", _dp), ("mt", _dp), ("count", _ip)]


class OracleBondParams(C.Structure):
    _fields_ = [("beta_bond", C.c_double), ("radius_factor", C.c_double),
                ("tensile_strength", C.c_double), ("contacts", C.c_int), ("mode", C.c_int)]


def build():
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)


_L = None


def lib():
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.oracle_neighbors_cell.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp, C.c_int, _ip, _ip]
        L.oracle_neighbors_cell.restype = C.c_int
        L.oracle_neighbors_brute.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _ip, _ip]
        L.oracle_neighbors_brute.restype = C.c_int
        L.oracle_zero_forces.argtypes = [C.POINTER(OracleParticles), C.POINTER(OracleParams)]
        L.oracle_update_tangential.argtypes = [C.POINTER(OracleParticles)]
        L.oracle_force_pp.argtypes = [C.POINTER(OracleParticles), _ip, _ip, C.c_int, C.POINTER(OracleParams)]
        L.oracle_force_pp.restype = C.c_int
        L.oracle_force_pw.argtypes = [C.POINTER(OracleParticles), C.POINTER(OracleWalls), C.POINTER(OracleParams)]
        for s in ("oracle_stage1", "oracle_stage2", "oracle_stage3"):
            getattr(L, s).argtypes = [C.POINTER(OracleParticles), C.c_double]
        L.oracle_run.argtypes = [C.POINTER(OracleParticles), C.POINTER(OracleWalls), C.POINTER(OracleParams),
                                 C.POINTER(OracleRun), _dp, C.c_int, _ip, _dp]
        L.oracle_run.restype = C.c_int
        L.oracle_diagnostics.argtypes = [C.POINTER(OracleParticles), _dp]
        L.oracle_setup_bonds.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                         C.POINTER(OracleBonds)]
        L.oracle_setup_bonds.restype = C.c_int
        L.oracle_bond_force.argtypes = [C.POINTER(OracleParticles), C.POINTER(OracleBonds),
                                        C.POINTER(OracleBondParams), C.c_double, _ip]
        L.oracle_force_pp_unbonded.argtypes = [C.POINTER(OracleParticles), _ip, _ip, C.c_int,
                                               C.POINTER(OracleParams), C.POINTER(OracleBonds)]
        L.oracle_force_pp_unbonded.restype = C.c_int
        L.oracle_velocity_groups.argtypes = [C.POINTER(OracleParticles), _ip, _dp, C.c_int]
        _L = L
    return _L


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def _prep(st):
    for f in ("x", "u", "omega", "force", "torque", "m", "rad", "E", "nu", "G", "I",
              "sw_x", "sw_y", "sw_z", "ss_x", "ss_y", "ss_z"):
        a = getattr(st, f)
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            setattr(st, f, np.ascontiguousarray(a, dtype=np.float64))
    for f in ("ss_idx", "ss_cnt"):
        a = getattr(st, f)
        if a.dtype != np.int32 or not a.flags.c_contiguous:
            setattr(st, f, np.ascontiguousarray(a, dtype=np.int32))


def particles(st) -> OracleParticles:
    _prep(st)
    p = OracleParticles()
    p.n = st.x.shape[0]
    p.max_cnts = st.ss_x.shape[1]
    p.max_walls = st.sw_x.shape[1]
    for f in ("x", "u", "omega", "force", "torque", "m", "rad", "E", "nu", "G", "I",
              "sw_x", "sw_y", "sw_z", "ss_x", "ss_y", "ss_z"):
        setattr(p, f, _p(getattr(st, f)))
    p.ss_idx = _p(st.ss_idx, C.c_int)
    p.ss_cnt = _p(st.ss_cnt, C.c_int)
    p._keep = st
    return p


def walls(w) -> OracleWalls:
    o = OracleWalls()
    if w is None:
        o.n = 0
        return o
    o.n = w.x.shape[0]
    o.x, o.normal, o.E, o.nu, o.G = (_p(np.ascontiguousarray(a, dtype=np.float64))
                                     for a in (w.x, w.normal, w.E, w.nu, w.G))
    for f in ("u", "omega"):
        a = getattr(w, f, None)
        setattr(o, f, None if a is None else _p(np.ascontiguousarray(a, dtype=np.float64)))
    o._keep = w
    return o


def params(*, dt, cor_pp=1.0, cor_pw=1.0, fr_pp=0.0, fr_pw=0.0, gravity=(0., 0., 0.), mode=0):
    p = OracleParams()
    p.cor_pp, p.cor_pw, p.fr_pp, p.fr_pw, p.dt = cor_pp, cor_pw, fr_pp, fr_pw, dt
    p.gravity[:] = list(gravity)
    p.mode = mode
    return p


def neighbors_cell(x, rc, lo, hi, max_n=64):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    ids = np.zeros((n, max_n), dtype=np.int32)
    lo3 = np.asarray(lo, dtype=np.float64)
    hi3 = np.asarray(hi, dtype=np.float64)
    mc = lib().oracle_neighbors_cell(n, _p(x), rc, _p(lo3), _p(hi3), max_n, _p(counts, C.c_int), _p(ids, C.c_int))
    if mc > max_n:
        return neighbors_cell(x, rc, lo, hi, mc)
    if mc < 0:
        raise ValueError("non-finite positions")
    return counts, ids


def neighbors_brute(x, rc, max_n=64):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    ids = np.zeros((n, max_n), dtype=np.int32)
    mc = lib().oracle_neighbors_brute(n, _p(x), rc, max_n, _p(counts, C.c_int), _p(ids, C.c_int))
    if mc > max_n:
        return neighbors_brute(x, rc, mc)
    return counts, ids


def zero_forces(st, prm):
    lib().oracle_zero_forces(C.byref(particles(st)), C.byref(prm))


def update_tangential(st):
    lib().oracle_update_tangential(C.byref(particles(st)))


def force_pp(st, counts, ids, prm):
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    return lib().oracle_force_pp(C.byref(particles(st)), _p(counts, C.c_int), _p(ids, C.c_int),
                                 ids.shape[1], C.byref(prm))


def force_pw(st, w, prm):
    lib().oracle_force_pw(C.byref(particles(st)), C.byref(walls(w)), C.byref(prm))


def stage(st, which, dt):
    getattr(lib(), f"oracle_stage{which}")(C.byref(particles(st)), dt)


def compute_forces(st, w, counts, ids, prm):
    """zero, update tangential, pp, pw -- src/CabanaDEM_Force.hpp:698-710."""
    zero_forces(st, prm)
    update_tangential(st)
    ov = force_pp(st, counts, ids, prm)
    if w is not None and w.x.shape[0] > 0:
        force_pw(st, w, prm)
    return ov


def diagnostics(st):
    out = np.zeros(5)
    lib().oracle_diagnostics(C.byref(particles(st)), _p(out))
    return out


def run(st, w, prm, *, nsteps, rc, lo, hi, rebuild_every=1, max_n=64, diag_every=0, threads=0):
    r = OracleRun()
    r.nsteps, r.rc = nsteps, rc
    r.lo[:] = list(lo)
    r.hi[:] = list(hi)
    r.rebuild_every, r.max_n, r.diag_every, r.threads = rebuild_every, max_n, diag_every, threads
    cap = (nsteps // diag_every + 1) if diag_every > 0 else 1
    diag = np.zeros((cap, 6))
    nd = C.c_int(0)
    secs = C.c_double(0.)
    rc_ = lib().oracle_run(C.byref(particles(st)), C.byref(walls(w)), C.byref(prm), C.byref(r),
                           _p(diag), cap, C.byref(nd), C.byref(secs))
    if rc_ != 0:
        raise RuntimeError(f"oracle_run failed with {rc_}")
    return diag[: nd.value], secs.value


class Bonds:
    """Bond state in the reference layout: idx/init_len [n][max_bonds], ft/mn/mt
    [n][max_bonds][3], count [n] (src/CabanaBondedDEM_Particles.hpp:40-47,187-201)."""

    def __init__(self, n, max_bonds):
        self.n, self.max_bonds = n, max_bonds
        self.idx = -np.ones((n, max_bonds), dtype=np.int32)
        self.init_len = np.zeros((n, max_bonds))
        self.ft = np.zeros((n, max_bonds, 3))
        self.mn = np.zeros((n, max_bonds, 3))
        self.mt = np.zeros((n, max_bonds, 3))
        self.count = np.zeros(n, dtype=np.int32)

    def copy(self):
        c = Bonds(self.n, self.max_bonds)
        for f in ("idx", "init_len", "ft", "mn", "mt", "count"):
            setattr(c, f, getattr(self, f).copy())
        return c

    def struct(self):
        b = OracleBonds()
        b.max_bonds = self.max_bonds
        b.idx = _p(self.idx, C.c_int)
        b.init_len = _p(self.init_len)
        b.ft, b.mn, b.mt = _p(self.ft), _p(self.mn), _p(self.mt)
        b.count = _p(self.count, C.c_int)
        b._keep = self
        return b


def bond_params(beta=0.1, radius_factor=1.0, strength=0.0, contacts=1, mode=0):
    bp = OracleBondParams()
    bp.beta_bond, bp.radius_factor, bp.tensile_strength = beta, radius_factor, strength
    bp.contacts, bp.mode = contacts, mode
    return bp


def setup_bonds(st, factor, rc, lo, hi, max_bonds):
    b = Bonds(st.n, max_bonds)
    x = np.ascontiguousarray(st.x, dtype=np.float64)
    rad = np.ascontiguousarray(st.rad, dtype=np.float64)
    lo3, hi3 = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    rc_ = lib().oracle_setup_bonds(st.n, _p(x), _p(rad), factor, rc, _p(lo3), _p(hi3),
                                   C.byref(b.struct()))
    if rc_ != 0:
        raise RuntimeError(f"oracle_setup_bonds failed with {rc_}")
    return b


def bond_force(st, bonds, bp, dt):
    nb = C.c_int(0)
    lib().oracle_bond_force(C.byref(particles(st)), C.byref(bonds.struct()), C.byref(bp), dt,
                            C.byref(nb))
    return nb.value


def force_pp_unbonded(st, counts, ids, prm, bonds):
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    return lib().oracle_force_pp_unbonded(C.byref(particles(st)), _p(counts, C.c_int),
                                          _p(ids, C.c_int), ids.shape[1], C.byref(prm),
                                          C.byref(bonds.struct()))


def velocity_groups(st, group, vel):
    group = np.ascontiguousarray(group, dtype=np.int32)
    vel = np.ascontiguousarray(np.asarray(vel, dtype=np.float64).reshape(-1, 3))
    lib().oracle_velocity_groups(C.byref(particles(st)), _p(group, C.c_int), _p(vel), vel.shape[0])
