"""ctypes binding of include/cabanadem.h (libcabanadem_hip.so).

This is plumbing for the Python tests and the benchmark: every compute call
goes through the C ABI into the HIP kernels.  There is no Python or CPU
fallback; if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("CABANADEM_LIB") or
                Path(__file__).resolve().parent / "lib" / "libcabanadem_hip.so")

DEM_ABI_VERSION = 5  # include/cabanadem.h DEM_ABI_VERSION
DEM_MAX_WALLS = 8
DEM_MAX_GROUPS = 8
DEM_MAX_BONDS = 32
GROUP_VELOCITY = 0
GROUP_FIXED_FORCE = 1
DEM_OK = 0
ERRORS = {
    -1: "DEM_ERR_ARG",
    -2: "DEM_ERR_HIP",
    -3: "DEM_ERR_NEIGH_OVERFLOW",
    -4: "DEM_ERR_CONTACT_OVERFLOW",
    -5: "DEM_ERR_NONFINITE",
    -6: "DEM_ERR_SKIN",
    -7: "DEM_ERR_BOND_OVERFLOW",
    -8: "DEM_ERR_STATE",
}
MODE_REFERENCE = 0
MODE_CORRECTED = 1
REBUILD_AUTO = -1  # include/cabanadem.h DEM_REBUILD_AUTO
FORCE_PASS_FUSED, FORCE_PASS_SPLIT, FORCE_PASS_WAVE = 0, 1, 2  # DEM_FORCE_PASS_*

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DemConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("n_walls", C.c_int), ("max_neighbors", C.c_int), ("max_contacts", C.c_int),
        ("dt", C.c_double), ("cor_pp", C.c_double), ("cor_pw", C.c_double),
        ("friction_pp", C.c_double), ("friction_pw", C.c_double),
        ("gravity", C.c_double * 3), ("cutoff", C.c_double), ("skin", C.c_double),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("rebuild_every", C.c_int), ("physics_mode", C.c_int), ("block_size", C.c_int),
        ("reserved", C.c_int),
    ]


class DemParticlesView(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("max_cnts", C.c_int), ("max_walls", C.c_int), ("reserved", C.c_int),
        ("x", _dp), ("u", _dp), ("omega", _dp), ("force", _dp), ("torque", _dp),
        ("m", _dp), ("rad", _dp), ("E", _dp), ("nu", _dp), ("G", _dp), ("I", _dp),
        ("sw_x", _dp), ("sw_y", _dp), ("sw_z", _dp),
        ("ss_x", _dp), ("ss_y", _dp), ("ss_z", _dp),
        ("ss_idx", _ip), ("ss_cnt", _ip),
    ]


class DemWallsView(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("reserved", C.c_int),
        ("x", _dp), ("normal", _dp), ("E", _dp), ("nu", _dp), ("G", _dp),
        ("u", _dp), ("omega", _dp),
    ]


class DemCounters(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("rebuilds", C.c_int64),
        ("max_neighbors_seen", C.c_int), ("neighbor_capacity", C.c_int),
        ("mean_neighbors", C.c_double), ("mean_contacts", C.c_double),
        ("mean_wall_contacts", C.c_double), ("bonds_broken", C.c_int64),
        ("displacement_checks", C.c_int64),
    ]


class DemSystemDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("n_walls", C.c_int),
        ("cor_pp", C.c_double), ("cor_pw", C.c_double),
        ("friction_pp", C.c_double), ("friction_pw", C.c_double),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
    ]


DEM_TUNE_MAX_TRIALS = 128


class DemTuneResult(C.Structure):
    _fields_ = [
        ("rebuild_every", C.c_int), ("grid_blocks", C.c_int), ("use_graph", C.c_int),
        ("force_pass", C.c_int), ("block", C.c_int), ("tile_rows", C.c_int),
        ("n_trials", C.c_int), ("reserved", C.c_int),
        ("ms_per_step", C.c_double),
        ("trial_rebuild_every", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_grid_blocks", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_graph", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_force_pass", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_block", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_tile_rows", C.c_int * DEM_TUNE_MAX_TRIALS),
        ("trial_ms_per_step", C.c_double * DEM_TUNE_MAX_TRIALS),
    ]

    def trials(self):
        return [dict(rebuild_every=self.trial_rebuild_every[k], grid_blocks=self.trial_grid_blocks[k],
                     graph=self.trial_graph[k], force_pass=self.trial_force_pass[k],
                     block=self.trial_block[k], tile_rows=self.trial_tile_rows[k],
                     ms_per_step=self.trial_ms_per_step[k])
                for k in range(self.n_trials)]


class DemBondConfig(C.Structure):
    _fields_ = [
        ("max_bonds", C.c_int), ("unbonded_contacts", C.c_int),
        ("beta_bond", C.c_double), ("radius_factor", C.c_double),
        ("tensile_strength", C.c_double),
    ]


class DemBondsView(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("max_bonds", C.c_int), ("idx", _ip), ("count", _ip),
        ("init_len", _dp),
        ("ft_x", _dp), ("ft_y", _dp), ("ft_z", _dp),
        ("mn_x", _dp), ("mn_y", _dp), ("mn_z", _dp),
        ("mt_x", _dp), ("mt_y", _dp), ("mt_z", _dp),
    ]


SYMBOLS = {
    "dem_create": (C.c_int, [C.POINTER(DemConfig), C.POINTER(C.c_void_p)]),
    "dem_destroy": (None, [C.c_void_p]),
    "dem_last_error": (C.c_char_p, [C.c_void_p]),
    "dem_abi_version": (C.c_int, []),
    "dem_set_params": (C.c_int, [C.c_void_p, C.POINTER(DemConfig)]),
    "dem_device_count": (C.c_int, []),
    "dem_upload_particles": (C.c_int, [C.c_void_p, C.POINTER(DemParticlesView)]),
    "dem_download_particles": (C.c_int, [C.c_void_p, C.POINTER(DemParticlesView)]),
    "dem_upload_walls": (C.c_int, [C.c_void_p, C.POINTER(DemWallsView)]),
    "dem_build_neighbors": (C.c_int, [C.c_void_p]),
    "dem_get_neighbors": (C.c_int, [C.c_void_p, _ip, _ip, C.c_int]),
    "dem_zero_forces": (C.c_int, [C.c_void_p]),
    "dem_set_kinematics_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "dem_get_kinematics_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dem_update_tangential": (C.c_int, [C.c_void_p]),
    "dem_force_pp": (C.c_int, [C.c_void_p]),
    "dem_force_pw": (C.c_int, [C.c_void_p]),
    "dem_compute_forces": (C.c_int, [C.c_void_p]),
    "dem_stage1": (C.c_int, [C.c_void_p]),
    "dem_stage2": (C.c_int, [C.c_void_p]),
    "dem_stage3": (C.c_int, [C.c_void_p]),
    "dem_step": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_set_diag_every": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_get_diag_history": (C.c_int, [C.c_void_p, _dp, C.c_int, _ip]),
    "dem_diagnostics": (C.c_int, [C.c_void_p, _dp]),
    "dem_get_counters": (C.c_int, [C.c_void_p, C.POINTER(DemCounters)]),
    "dem_synchronize": (C.c_int, [C.c_void_p]),
    "dem_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_get_timing": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int64)]),
    "dem_create_ensemble": (C.c_int, [C.POINTER(DemConfig), C.c_int, C.POINTER(DemSystemDesc),
                                      C.POINTER(C.c_void_p)]),
    "dem_get_systems": (C.c_int, [C.c_void_p, _ip, _ip]),
    "dem_upload_system_walls": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(DemWallsView)]),
    "dem_set_system_params": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double]),
    "dem_system_diagnostics": (C.c_int, [C.c_void_p, _dp]),
    "dem_set_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_set_force_pass": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_set_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "dem_get_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "dem_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "dem_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(DemTuneResult)]),
    "dem_enable_bonds": (C.c_int, [C.c_void_p, C.POINTER(DemBondConfig)]),
    "dem_setup_bonds": (C.c_int, [C.c_void_p, C.c_double]),
    "dem_upload_bonds": (C.c_int, [C.c_void_p, C.POINTER(DemBondsView)]),
    "dem_download_bonds": (C.c_int, [C.c_void_p, C.POINTER(DemBondsView)]),
    "dem_bond_forces": (C.c_int, [C.c_void_p]),
    "dem_set_external_force": (C.c_int, [C.c_void_p, _dp]),
    "dem_set_groups": (C.c_int, [C.c_void_p, _ip, C.c_int, _ip, _dp]),
    "dem_generate_lattice": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_uint64, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                       _dp]),
}


def load_library(path: Path | str | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python cabanadem_amd/build.py lib` "
            "(there is no CPU fallback for the DEM kernels)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.dem_abi_version() != DEM_ABI_VERSION:
        raise RuntimeError(f"{p} has ABI version {lib.dem_abi_version()}, this binding expects "
                           f"{DEM_ABI_VERSION}: rebuild the library")
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


class DemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _ptr(a, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def generate_lattice(nx, ny, nz, spacing, lo, jitter, rmin, rmax, vsigma, rho, E, nu, seed):
    """Jittered lattice packing from the library's host generator (no GPU needed)."""
    n = nx * ny * nz
    out = {k: np.zeros(3 * n if k in ("x", "u") else n) for k in
           ("x", "u", "rad", "m", "I", "E", "nu", "G")}
    lo3 = np.asarray(lo, dtype=np.float64)
    rc = lib().dem_generate_lattice(nx, ny, nz, spacing, _ptr(lo3), jitter, rmin, rmax, vsigma,
                                    rho, E, nu, seed, _ptr(out["x"]), _ptr(out["u"]),
                                    _ptr(out["rad"]), _ptr(out["m"]), _ptr(out["I"]),
                                    _ptr(out["E"]), _ptr(out["nu"]), _ptr(out["G"]))
    if rc != 0:
        raise DemError(rc, "dem_generate_lattice failed")
    out["x"] = out["x"].reshape(n, 3)
    out["u"] = out["u"].reshape(n, 3)
    return out


class ParticleState:
    """Host particle state in the reference layout (xyz interleaved, row-major histories)."""

    def __init__(self, n: int, max_cnts: int = 16, max_walls: int = 6):
        self.n, self.max_cnts, self.max_walls = n, max_cnts, max_walls
        z3 = lambda: np.zeros((n, 3))
        self.x, self.u, self.omega, self.force, self.torque = z3(), z3(), z3(), z3(), z3()
        self.m, self.rad, self.E, self.nu, self.G, self.I = (np.ones(n) for _ in range(6))
        self.sw_x, self.sw_y, self.sw_z = (np.zeros((n, max_walls)) for _ in range(3))
        self.ss_x, self.ss_y, self.ss_z = (np.zeros((n, max_cnts)) for _ in range(3))
        self.ss_idx = -np.ones((n, max_cnts), dtype=np.int32)
        self.ss_cnt = np.zeros(n, dtype=np.int32)

    FIELDS = ("x", "u", "omega", "force", "torque", "m", "rad", "E", "nu", "G", "I",
              "sw_x", "sw_y", "sw_z", "ss_x", "ss_y", "ss_z", "ss_idx", "ss_cnt")

    def copy(self) -> "ParticleState":
        c = ParticleState(self.n, self.max_cnts, self.max_walls)
        for f in self.FIELDS:
            setattr(c, f, getattr(self, f).copy())
        return c

    def view(self, fields=None) -> DemParticlesView:
        for f in self.FIELDS:
            a = getattr(self, f)
            want = np.int32 if f in ("ss_idx", "ss_cnt") else np.float64
            if a.dtype != want or not a.flags.c_contiguous:
                setattr(self, f, np.ascontiguousarray(a, dtype=want))
        v = DemParticlesView()
        v.n, v.max_cnts, v.max_walls = self.n, self.max_cnts, self.max_walls
        for f in self.FIELDS:
            if fields is not None and f not in fields:
                continue
            ct = C.c_int if f in ("ss_idx", "ss_cnt") else C.c_double
            setattr(v, f, _ptr(getattr(self, f), ct))
        return v


class Walls:
    """Infinite planes: point x, unit normal, material; optional velocity u and
    spin omega ([n][3], used in corrected mode only)."""

    def __init__(self, x, normal, E, nu, G, u=None, omega=None):
        self.x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 3))
        self.normal = np.ascontiguousarray(np.asarray(normal, dtype=np.float64).reshape(-1, 3))
        n = self.x.shape[0]
        f = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
        self.E, self.nu, self.G = f(E), f(nu), f(G)
        v3 = lambda v: None if v is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(v, dtype=np.float64), (n, 3)))
        self.u, self.omega = v3(u), v3(omega)

    @property
    def n(self):
        return self.x.shape[0]

    def view(self) -> DemWallsView:
        v = DemWallsView()
        v.n = self.n
        v.x, v.normal, v.E, v.nu, v.G = (_ptr(a) for a in (self.x, self.normal, self.E, self.nu, self.G))
        v.u = None if self.u is None else _ptr(self.u)
        v.omega = None if self.omega is None else _ptr(self.omega)
        return v


class BondState:
    """Host bond state in the reference layout: rows [n][max_bonds] per field."""

    HIST = ("ft_x", "ft_y", "ft_z", "mn_x", "mn_y", "mn_z", "mt_x", "mt_y", "mt_z")

    def __init__(self, n: int, max_bonds: int):
        self.n, self.max_bonds = n, max_bonds
        self.idx = -np.ones((n, max_bonds), dtype=np.int32)
        self.count = np.zeros(n, dtype=np.int32)
        self.init_len = np.zeros((n, max_bonds))
        for f in self.HIST:
            setattr(self, f, np.zeros((n, max_bonds)))

    def history(self) -> np.ndarray:
        """(n, max_bonds, 9): ft, mn, mt."""
        return np.stack([getattr(self, f) for f in self.HIST], axis=-1)

    def view(self) -> DemBondsView:
        for f in ("idx", "count", "init_len") + self.HIST:
            a = getattr(self, f)
            want = np.int32 if f in ("idx", "count") else np.float64
            if a.dtype != want or not a.flags.c_contiguous:
                setattr(self, f, np.ascontiguousarray(a, dtype=want))
        v = DemBondsView()
        v.n, v.max_bonds = self.n, self.max_bonds
        v.idx, v.count = _ptr(self.idx, C.c_int), _ptr(self.count, C.c_int)
        v.init_len = _ptr(self.init_len)
        for f in self.HIST:
            setattr(v, f, _ptr(getattr(self, f)))
        return v


class System:
    """One DEM context on the GPU (a thin owner of a dem_ctx*)."""

    def __init__(self, n, *, dt, cutoff, lo, hi, n_walls=0, max_neighbors=0, max_contacts=16,
                 cor_pp=1.0, cor_pw=1.0, friction_pp=0.0, friction_pw=0.0, gravity=(0., 0., 0.),
                 skin=0.0, rebuild_every=1, physics_mode=MODE_REFERENCE, systems=None):
        """systems: None for one system, or a list of dicts (n, n_walls, cor_pp,
        cor_pw, friction_pp, friction_pw, lo, hi) for an ensemble, whose
        particles are concatenated in system order."""
        self._lib = lib()
        cfg = DemConfig()
        cfg.n, cfg.n_walls, cfg.max_neighbors, cfg.max_contacts = n, n_walls, max_neighbors, max_contacts
        cfg.dt, cfg.cor_pp, cfg.cor_pw = dt, cor_pp, cor_pw
        cfg.friction_pp, cfg.friction_pw = friction_pp, friction_pw
        cfg.gravity[:] = list(gravity)
        cfg.cutoff, cfg.skin = cutoff, skin
        cfg.lo[:] = list(lo)
        cfg.hi[:] = list(hi)
        cfg.rebuild_every, cfg.physics_mode, cfg.block_size = rebuild_every, physics_mode, 0
        self.cfg = cfg
        self.n = n
        self.n_systems = 1
        self.offsets = np.array([0, n])
        h = C.c_void_p()
        if systems is None:
            rc = self._lib.dem_create(C.byref(cfg), C.byref(h))
        else:
            arr = (DemSystemDesc * len(systems))()
            for k, sd in enumerate(systems):
                arr[k].n, arr[k].n_walls = sd["n"], sd.get("n_walls", 0)
                arr[k].cor_pp, arr[k].cor_pw = sd["cor_pp"], sd.get("cor_pw", sd["cor_pp"])
                arr[k].friction_pp = sd["friction_pp"]
                arr[k].friction_pw = sd.get("friction_pw", sd["friction_pp"])
                arr[k].lo[:] = list(sd["lo"])
                arr[k].hi[:] = list(sd["hi"])
            rc = self._lib.dem_create_ensemble(C.byref(cfg), len(systems), arr, C.byref(h))
            self.n_systems = len(systems)
            self.offsets = np.concatenate([[0], np.cumsum([sd["n"] for sd in systems])])
            self.n = int(self.offsets[-1])
        if rc != 0:
            raise DemError(rc, self._lib.dem_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise DemError(rc, self._lib.dem_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, st: ParticleState, fields=None):
        v = st.view(fields)
        self._check(self._lib.dem_upload_particles(self._h, C.byref(v)))

    def download(self, st: ParticleState | None = None, fields=None) -> ParticleState:
        if st is None:
            st = ParticleState(self.n, self.cfg.max_contacts, DEM_MAX_WALLS)
        v = st.view(fields)
        self._check(self._lib.dem_download_particles(self._h, C.byref(v)))
        return st

    def upload_walls(self, w: Walls):
        self._check(self._lib.dem_upload_walls(self._h, C.byref(w.view())))

    def build_neighbors(self):
        self._check(self._lib.dem_build_neighbors(self._h))

    def neighbors(self, max_n=64):
        counts = np.zeros(self.n, dtype=np.int32)
        ids = np.zeros((self.n, max_n), dtype=np.int32)
        self._check(self._lib.dem_get_neighbors(self._h, _ptr(counts, C.c_int), _ptr(ids, C.c_int), max_n))
        return counts, ids

    def zero_forces(self): self._check(self._lib.dem_zero_forces(self._h))

    def set_kinematics_device(self, x_ptr, u_ptr, clear_history=True):
        """Positions / velocities from device arrays (double [n][3], reference
        order; addresses as ints or None), on the context's stream."""
        self._check(self._lib.dem_set_kinematics_device(self._h, x_ptr, u_ptr, 1 if clear_history else 0))

    def get_kinematics_device(self, x_ptr, u_ptr):
        self._check(self._lib.dem_get_kinematics_device(self._h, x_ptr, u_ptr))
    def update_tangential(self): self._check(self._lib.dem_update_tangential(self._h))
    def force_pp(self): self._check(self._lib.dem_force_pp(self._h))
    def force_pw(self): self._check(self._lib.dem_force_pw(self._h))
    def compute_forces(self): self._check(self._lib.dem_compute_forces(self._h))
    def stage1(self): self._check(self._lib.dem_stage1(self._h))
    def stage2(self): self._check(self._lib.dem_stage2(self._h))
    def stage3(self): self._check(self._lib.dem_stage3(self._h))
    def step(self, nsteps=1): self._check(self._lib.dem_step(self._h, nsteps))
    def sync(self): self._check(self._lib.dem_synchronize(self._h))

    def set_diag_every(self, every):
        self._check(self._lib.dem_set_diag_every(self._h, every))

    def diag_history(self) -> np.ndarray:
        nrows = C.c_int(0)
        self._check(self._lib.dem_get_diag_history(self._h, None, 0, C.byref(nrows)))
        out = np.zeros((max(nrows.value, 1), 9))
        self._check(self._lib.dem_get_diag_history(self._h, _ptr(out), nrows.value, C.byref(nrows)))
        return out[: nrows.value]

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(5)
        self._check(self._lib.dem_diagnostics(self._h, _ptr(out)))
        return out

    def counters(self) -> DemCounters:
        c = DemCounters()
        self._check(self._lib.dem_get_counters(self._h, C.byref(c)))
        return c

    # ---- ensembles ----
    def upload_system_walls(self, system: int, w: Walls):
        self._check(self._lib.dem_upload_system_walls(self._h, system, C.byref(w.view())))

    def set_system_params(self, system, cor_pp, cor_pw, friction_pp, friction_pw):
        self._check(self._lib.dem_set_system_params(self._h, system, cor_pp, cor_pw, friction_pp,
                                                    friction_pw))

    def system_diagnostics(self) -> np.ndarray:
        out = np.zeros((self.n_systems, 5))
        self._check(self._lib.dem_system_diagnostics(self._h, _ptr(out)))
        return out

    def set_launch(self, grid_blocks: int):
        self._check(self._lib.dem_set_launch(self._h, grid_blocks))

    def set_force_pass(self, mode):
        """FORCE_PASS_FUSED (one kernel, each lane walks its own contacts),
        FORCE_PASS_SPLIT (detection kernel + force kernel) or FORCE_PASS_WAVE
        (one kernel, contact arithmetic spread over each wave); identical
        results.  True / False stand for split / fused."""
        if mode is True:
            mode = FORCE_PASS_SPLIT
        elif mode is False:
            mode = FORCE_PASS_FUSED
        self._check(self._lib.dem_set_force_pass(self._h, int(mode)))

    def set_shape(self, block: int = 0, tile_rows: int = 0):
        """Force-kernel tile size (64, 128, 256; 0: 128) and cell-tile rows
        (1-32; 0: 8): dem_set_shape."""
        self._check(self._lib.dem_set_shape(self._h, int(block), int(tile_rows)))

    def shape(self):
        b, r = C.c_int(), C.c_int()
        self._check(self._lib.dem_get_shape(self._h, C.byref(b), C.byref(r)))
        return b.value, r.value

    def set_params(self, **changes):
        """Non-structural dem_config fields (dt, cor_pp, ..., rebuild_every, physics_mode)."""
        for k, v in changes.items():
            if k in ("gravity", "lo", "hi"):
                getattr(self.cfg, k)[:] = list(v)
            else:
                setattr(self.cfg, k, v)
        self._check(self._lib.dem_set_params(self._h, C.byref(self.cfg)))

    def set_graph(self, on: bool):
        self._check(self._lib.dem_set_graph(self._h, 1 if on else 0))

    def autotune(self, trial_steps: int = 100) -> DemTuneResult:
        r = DemTuneResult()
        self._check(self._lib.dem_autotune(self._h, trial_steps, C.byref(r)))
        return r

    # ---- bonded DEM, groups, external force ----
    def enable_bonds(self, max_bonds, *, beta_bond=0.1, radius_factor=1.0, tensile_strength=0.0,
                     unbonded_contacts=False):
        b = DemBondConfig()
        b.max_bonds, b.unbonded_contacts = max_bonds, 1 if unbonded_contacts else 0
        b.beta_bond, b.radius_factor, b.tensile_strength = beta_bond, radius_factor, tensile_strength
        self._check(self._lib.dem_enable_bonds(self._h, C.byref(b)))
        self.max_bonds = max_bonds

    def setup_bonds(self, factor=3.0):
        self._check(self._lib.dem_setup_bonds(self._h, factor))

    def upload_bonds(self, b: BondState):
        self._check(self._lib.dem_upload_bonds(self._h, C.byref(b.view())))

    def download_bonds(self, b: BondState | None = None) -> BondState:
        if b is None:
            b = BondState(self.n, self.max_bonds)
        self._check(self._lib.dem_download_bonds(self._h, C.byref(b.view())))
        return b

    def bond_forces(self):
        self._check(self._lib.dem_bond_forces(self._h))

    def set_external_force(self, f):
        if f is None:
            self._check(self._lib.dem_set_external_force(self._h, None))
            return
        a = np.ascontiguousarray(np.asarray(f, dtype=np.float64).reshape(self.n, 3))
        self._check(self._lib.dem_set_external_force(self._h, _ptr(a)))

    def set_groups(self, group, kinds=None, vel=None):
        """group[i] in 0..len(kinds) (0 = none); kinds[g] = GROUP_*; vel[g] = (vx, vy, vz)."""
        if group is None:
            self._check(self._lib.dem_set_groups(self._h, None, 0, None, None))
            return
        g = np.ascontiguousarray(np.asarray(group, dtype=np.int32))
        k = np.ascontiguousarray(np.asarray(kinds, dtype=np.int32))
        v = None if vel is None else np.ascontiguousarray(np.asarray(vel, dtype=np.float64).reshape(-1, 3))
        self._check(self._lib.dem_set_groups(self._h, _ptr(g, C.c_int), len(k), _ptr(k, C.c_int),
                                             _ptr(v)))

    def set_timing(self, on):
        """True / 1: time every force pass; k > 1: every k-th; False / 0: off."""
        every = int(on) if not isinstance(on, bool) else (1 if on else 0)
        self._check(self._lib.dem_set_timing(self._h, every))

    def timing(self):
        ms = np.zeros(2)
        n = (C.c_int64 * 2)()
        self._check(self._lib.dem_get_timing(self._h, _ptr(ms), n))
        return ms, [n[0], n[1]]


def device_count() -> int:
    return int(lib().dem_device_count())
