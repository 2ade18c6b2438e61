"""Synthetic plain-DEM workloads (host side), shared by the benchmark and tests.

plain_small : 10,000 monodisperse spheres (25 x 20 x 20 lattice, spacing 2.05 r),
              r = 1 mm, rho 2600, E 1e7, nu 0.3, e 0.9, mu 0.3, jitter U(+-0.02 r),
              u ~ N(0, 0.05 m/s), gravity -9.81 z, dt 1e-6, cutoff 2.2 mm,
              rebuild every 20 steps, six walls one spacing outside the lattice.
plain_large : 4,000,000 polydisperse spheres (200 x 200 x 100 lattice, spacing
              2.4 mm), r ~ U[0.8, 1.2] mm, rho 2600, E 1e7, nu 0.3, e 0.8,
              mu 0.4, jitter U(+-0.1 mm), u = 0, gravity -9.81 z, dt 1e-6,
              skin 0.2 mm -> cutoff 2 r_max + skin = 2.6 mm, rebuild every 20
              steps, six walls one spacing outside the lattice.
plain_dense : plain_large's lattice at spacing 2.0 mm, r ~ U[0.97, 1.03] mm, no
              jitter (about 1.3 tracked contacts per sphere once settled).
plain_packed: a pre-compressed bed held by its walls: spacing 1.98 mm, r ~
              U[0.995, 1.005] mm, so every axial lattice pair overlaps by 10-30
              um and the walls press the outer layers by about 16 um (about six
              tracked contacts per sphere, one per neighbour in the rows).
All are built with the library's counter-based generator (dem_generate_lattice).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Workload:
    name: str
    state: abi.ParticleState
    walls: abi.Walls
    dt: float
    cutoff: float
    skin: float
    lo: tuple
    hi: tuple
    cor: float
    friction: float
    gravity: tuple
    rebuild_every: int
    max_contacts: int = 16
    extra: dict = field(default_factory=dict)

    def system(self, mode=abi.MODE_REFERENCE, **kw) -> abi.System:
        s = abi.System(self.state.n, dt=self.dt, cutoff=self.cutoff, lo=self.lo, hi=self.hi,
                       n_walls=self.walls.n, max_contacts=self.max_contacts, cor_pp=self.cor,
                       cor_pw=self.cor, friction_pp=self.friction, friction_pw=self.friction,
                       gravity=self.gravity, skin=self.skin, rebuild_every=self.rebuild_every,
                       physics_mode=mode, **kw)
        s.upload_walls(self.walls)
        s.upload(self.state)
        return s


def _box(nx, ny, nz, sp, E, nu, off=None):
    """Six walls at distance off (default one spacing) outside the lattice."""
    off = sp if off is None else off
    lo = (-off, -off, -off)
    hi = ((nx - 1) * sp + off, (ny - 1) * sp + off, (nz - 1) * sp + off)
    G = E / (2. * (1. + nu))
    x = [(lo[0], 0, 0), (hi[0], 0, 0), (0, lo[1], 0), (0, hi[1], 0), (0, 0, lo[2]), (0, 0, hi[2])]
    nrm = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    return lo, hi, abi.Walls(x, nrm, E, nu, G)


def lattice(nx, ny, nz, *, spacing, jitter, rmin, rmax, vsigma, rho, E, nu, seed, max_cnts=16):
    g = abi.generate_lattice(nx, ny, nz, spacing, (0., 0., 0.), jitter, rmin, rmax, vsigma,
                             rho, E, nu, seed)
    st = abi.ParticleState(nx * ny * nz, max_cnts=max_cnts, max_walls=6)
    st.x[:] = g["x"]
    st.u[:] = g["u"]
    st.rad[:], st.m[:], st.I[:] = g["rad"], g["m"], g["I"]
    st.E[:], st.nu[:], st.G[:] = g["E"], g["nu"], g["G"]
    return st


def plain_small(seed: int = 0, dims=(25, 20, 20)) -> Workload:
    r, E, nu = 1e-3, 1e7, 0.3
    sp = 2.05 * r
    st = lattice(*dims, spacing=sp, jitter=0.02 * r, rmin=r, rmax=r, vsigma=0.05, rho=2600.,
                 E=E, nu=nu, seed=seed)
    lo, hi, w = _box(*dims, sp, E, nu)
    skin = 0.2 * r
    return Workload("plain_small", st, w, 1e-6, 2 * r + skin, skin, lo, hi, 0.9, 0.3,
                    (0., 0., -9.81), 20, extra={"dims": dims})


def plain_large(seed: int = 1, dims=(200, 200, 100)) -> Workload:
    E, nu = 1e7, 0.3
    sp = 2.4e-3
    st = lattice(*dims, spacing=sp, jitter=0.1e-3, rmin=0.8e-3, rmax=1.2e-3, vsigma=0.0,
                 rho=2600., E=E, nu=nu, seed=seed)
    lo, hi, w = _box(*dims, sp, E, nu)
    skin = 0.2e-3
    return Workload("plain_large", st, w, 1e-6, 2 * 1.2e-3 + skin, skin, lo, hi, 0.8, 0.4,
                    (0., 0., -9.81), 20, extra={"dims": dims})


def plain_dense(seed: int = 1, dims=(200, 200, 100)) -> Workload:
    """The large workload as a contact-dense bed: spacing 2.0 mm with radii
    U[0.97, 1.03] mm, no jitter, so about half of the lattice neighbour
    pairs start overlapping (by up to 0.06 mm) and the bed holds about three
    tracked contacts per sphere; otherwise plain_large (E 1e7, nu 0.3, e 0.8,
    mu 0.4, gravity, dt 1e-6, skin 0.2 mm, cutoff 2 r_max + skin, rebuild
    every 20 steps, six walls one spacing outside the lattice)."""
    E, nu = 1e7, 0.3
    sp = 2.0e-3
    rmax = 1.03e-3
    st = lattice(*dims, spacing=sp, jitter=0.0, rmin=0.97e-3, rmax=rmax, vsigma=0.0,
                 rho=2600., E=E, nu=nu, seed=seed)
    lo, hi, w = _box(*dims, sp, E, nu)
    skin = 0.2e-3
    return Workload("plain_dense", st, w, 1e-6, 2 * rmax + skin, skin, lo, hi, 0.8, 0.4,
                    (0., 0., -9.81), 20, extra={"dims": dims})


def plain_packed(seed: int = 1, dims=(200, 200, 100)) -> Workload:
    """The large workload as a pre-compressed bed, held together by its walls:
    spacing 1.98 mm with radii U[0.995, 1.005] mm, no jitter, so each of the
    six axial lattice neighbours overlaps by 10-30 um (the diagonal ones,
    2.8 mm apart, lie outside the 2.21 mm cutoff).  The walls stand 0.984 mm
    outside the outer lattice planes, pressing the outer spheres by about
    16 um: the Hertz force of a sphere on a wall of the same material at
    2^(-1/3) of a pair's overlap equals the pair's, so the bed starts close
    to balance and stays packed under gravity (a bottom sphere carries the
    weight of 100 layers, about 0.01 N, against about 0.015 N per contact).
    Otherwise plain_large (E 1e7, nu 0.3, e 0.8, mu 0.4, gravity, dt 1e-6,
    skin 0.2 mm, cutoff 2 r_max + skin, rebuild every 20 steps)."""
    E, nu = 1e7, 0.3
    sp = 1.98e-3
    rmax = 1.005e-3
    st = lattice(*dims, spacing=sp, jitter=0.0, rmin=0.995e-3, rmax=rmax, vsigma=0.0,
                 rho=2600., E=E, nu=nu, seed=seed)
    lo, hi, w = _box(*dims, sp, E, nu, off=0.984e-3)
    skin = 0.2e-3
    return Workload("plain_packed", st, w, 1e-6, 2 * rmax + skin, skin, lo, hi, 0.8, 0.4,
                    (0., 0., -9.81), 20, extra={"dims": dims})


@dataclass
class EnsembleWorkload:
    name: str
    systems: list          # per system: n, n_walls, cor_pp, cor_pw, friction_pp, friction_pw, lo, hi
    walls: list            # abi.Walls per system
    states: list           # abi.ParticleState per system
    dt: float
    cutoff: float
    skin: float
    gravity: tuple
    rebuild_every: int
    max_contacts: int = 16

    @property
    def n(self):
        return sum(sd["n"] for sd in self.systems)

    def concatenated(self) -> abi.ParticleState:
        st = abi.ParticleState(self.n, max_cnts=self.max_contacts, max_walls=6)
        off = 0
        for s in self.states:
            for f in ("x", "u", "omega", "m", "rad", "E", "nu", "G", "I"):
                getattr(st, f)[off:off + s.n] = getattr(s, f)
            off += s.n
        return st

    def _common(self, mode, **kw):
        return dict(dt=self.dt, cutoff=self.cutoff, max_contacts=self.max_contacts,
                    gravity=self.gravity, skin=self.skin, rebuild_every=self.rebuild_every,
                    physics_mode=mode, **kw)

    def system(self, mode=abi.MODE_REFERENCE, **kw) -> abi.System:
        """All systems in one context."""
        sd0 = self.systems[0]
        s = abi.System(self.n, lo=sd0["lo"], hi=sd0["hi"], systems=self.systems,
                       **self._common(mode, **kw))
        for k, w in enumerate(self.walls):
            s.upload_system_walls(k, w)
        s.upload(self.concatenated())
        return s

    def single(self, k, mode=abi.MODE_REFERENCE, **kw) -> abi.System:
        """System k alone in its own context, same configuration."""
        sd = self.systems[k]
        s = abi.System(sd["n"], lo=sd["lo"], hi=sd["hi"], n_walls=sd["n_walls"],
                       cor_pp=sd["cor_pp"], cor_pw=sd["cor_pw"], friction_pp=sd["friction_pp"],
                       friction_pw=sd["friction_pw"], **self._common(mode, **kw))
        s.upload_walls(self.walls[k])
        s.upload(self.states[k])
        return s


def ensemble(n_systems=256, seed=3, nmin=1000, nmax=64000, rmin=0.8e-3, rmax=1.2e-3,
             E=1e7, nu=0.3, rho=2600., spacing_r=2.05, vsigma=0.05,
             wall_offset=None) -> EnsembleWorkload:
    """Batched ensemble: particle counts log-uniform in [nmin, nmax]
    (floor(exp(U[ln nmin, ln nmax]))), restitution U[0.5, 0.95], friction
    U[0.1, 0.6], radii U[rmin, rmax], lattice spacing 2.05 rmax, jitter
    +-0.02 rmax, u ~ N(0, 0.05 m/s), six walls one spacing outside the
    lattice, gravity -9.81 z, dt 1e-6, skin 0.2 rmax over 2 rmax, rebuild
    every 20 steps.  Each system's lattice is the smallest near-cube of
    sites holding its count, filled in x, y, z order; the per-system draws
    come from numpy's PCG64 seeded with `seed`, the lattices from the
    library's counter-based generator seeded per system."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_systems))).astype(int)
    cors = rng.uniform(0.5, 0.95, size=n_systems)
    frs = rng.uniform(0.1, 0.6, size=n_systems)
    sp = spacing_r * rmax
    systems, walls, states = [], [], []
    for k in range(n_systems):
        n = int(counts[k])
        a = int(np.ceil(n ** (1. / 3.)))
        dims = (a, a, int(np.ceil(n / (a * a))))
        st_full = lattice(*dims, spacing=sp, jitter=0.02 * rmax, rmin=rmin, rmax=rmax,
                          vsigma=vsigma, rho=rho, E=E, nu=nu, seed=seed * 100003 + k)
        st = abi.ParticleState(n, max_cnts=16, max_walls=6)
        for f in ("x", "u", "m", "rad", "E", "nu", "G", "I"):
            getattr(st, f)[:] = getattr(st_full, f)[:n]
        lo, hi, w = _box(*dims, sp, E, nu, wall_offset)
        systems.append(dict(n=n, n_walls=6, cor_pp=float(cors[k]), cor_pw=float(cors[k]),
                            friction_pp=float(frs[k]), friction_pw=float(frs[k]), lo=lo, hi=hi))
        walls.append(w)
        states.append(st)
    skin = 0.2 * rmax
    return EnsembleWorkload("ensemble", systems, walls, states, 1e-6, 2 * rmax + skin, skin,
                            (0., 0., -9.81), 20)


def column(dims=(40, 40, 40), seed=0, rmin=0.9e-3, rmax=1.1e-3, E=1e7, nu=0.3, rho=2600.,
           cor=0.8, friction=0.4, runout=3.0) -> Workload:
    """Granular column collapse: a jittered lattice column (spacing 2.05
    rmax, jitter +-0.02 rmax, at rest) on a floor with retaining walls at
    x = 0 and y = 0, W; the +x side is open and the floor runs `runout`
    column widths beyond it.  Material as the large plain workload
    (E 1e7, nu 0.3, e 0.8, mu 0.4); gravity -9.81 z, dt 1e-6, skin 0.2 mm,
    rebuild every 20 steps.  Released at t = 0."""
    nx, ny, nz = dims
    sp = 2.05 * rmax
    g = abi.generate_lattice(nx, ny, nz, sp, (0.5 * sp, 0.5 * sp, 0.5 * sp), 0.02 * rmax, rmin, rmax,
                             0.0, rho, E, nu, seed)
    n = nx * ny * nz
    st = abi.ParticleState(n, max_cnts=16, max_walls=6)
    st.x[:] = g["x"]
    st.rad[:], st.m[:], st.I[:] = g["rad"], g["m"], g["I"]
    st.E[:], st.nu[:], st.G[:] = g["E"], g["nu"], g["G"]
    W_y = ny * sp
    G = E / (2. * (1. + nu))
    walls = abi.Walls([(0., 0., 0.), (0., 0., 0.), (0., 0., 0.), (0., W_y, 0.)],
                      [(0., 0., 1.), (1., 0., 0.), (0., 1., 0.), (0., -1., 0.)], E, nu, G)
    skin = 0.2e-3
    lo = (-sp, -sp, -sp)
    hi = (nx * sp * (1. + runout), W_y + sp, nz * sp + sp)
    wl = Workload("column", st, walls, 1e-6, 2 * rmax + skin, skin, lo, hi, cor, friction,
                  (0., 0., -9.81), 20, extra={"dims": dims, "width_y": W_y, "spacing": sp,
                                              "planes": [(0., 0., 1., 0.), (1., 0., 0., 0.),
                                                         (0., 1., 0., 0.), (0., -1., 0., -W_y)]})
    return wl


@dataclass
class BondedWorkload:
    name: str
    state: abi.ParticleState
    dt: float
    cutoff: float
    skin: float
    lo: tuple
    hi: tuple
    rebuild_every: int
    bond_factor: float
    max_bonds: int
    tensile_strength: float
    group: np.ndarray
    group_vel: list
    cor: float = 0.8
    friction: float = 0.4
    max_contacts: int = 8
    extra: dict = field(default_factory=dict)

    def system(self, mode=abi.MODE_REFERENCE, **kw) -> abi.System:
        s = abi.System(self.state.n, dt=self.dt, cutoff=self.cutoff, lo=self.lo, hi=self.hi,
                       max_contacts=self.max_contacts, cor_pp=self.cor, cor_pw=self.cor,
                       friction_pp=self.friction, friction_pw=self.friction, skin=self.skin,
                       rebuild_every=self.rebuild_every, physics_mode=mode, **kw)
        s.upload(self.state)
        s.enable_bonds(self.max_bonds, beta_bond=0.1, tensile_strength=self.tensile_strength,
                       unbonded_contacts=True)
        s.setup_bonds(self.bond_factor)
        s.set_groups(self.group, [abi.GROUP_VELOCITY, abi.GROUP_VELOCITY], self.group_vel)
        return s


def bonded_block(dims=(100, 100, 100), spacing=1.0e-3, r=0.5e-3, E=1e9, nu=0.25, rho=2600.,
                 v_pull=0.01, strength=1e6) -> BondedWorkload:
    """Bonded tensile block: simple-cubic lattice of touching spheres, bonds
    to the lattice neighbours closer than 1.05 spacing (the six face
    neighbours), x-end faces pulled apart at +-v_pull, dt 2e-8, Hertz
    contacts between unbonded pairs.  Cutoff 1.2 spacing (bond search and
    contacts with 0.2 spacing of skin); the faces move 0.4 um in 20,000
    steps, so the list is rebuilt every 500 steps under the skin check."""
    nx, ny, nz = dims
    n = nx * ny * nz
    st = abi.ParticleState(n, max_cnts=8, max_walls=1)
    idx = np.arange(n)
    ix, iy, iz = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    st.x[:] = np.stack([ix, iy, iz], axis=-1) * spacing
    m = 4. / 3. * np.pi * r ** 3 * rho
    st.rad[:] = r
    st.m[:] = m
    st.I[:] = 2. / 5. * m * r * r
    st.E[:] = E
    st.nu[:] = nu
    st.G[:] = E / (2. * (1. + nu))
    group = np.where(ix == 0, 1, np.where(ix == nx - 1, 2, 0)).astype(np.int32)
    pad = 3. * spacing
    lo = (-pad, -pad, -pad)
    hi = ((nx - 1) * spacing + pad, (ny - 1) * spacing + pad, (nz - 1) * spacing + pad)
    return BondedWorkload("bonded_block", st, 2e-8, 1.2 * spacing, 0.2 * spacing, lo, hi, 500,
                          1.05 * spacing / r, 8, strength, group,
                          [(-v_pull, 0., 0.), (v_pull, 0., 0.)], extra={"dims": dims})


def bytes_bonded_passes(bbar, nbar, cbar):
    """Algorithmic bytes per particle-step of a bonded step, reference data
    model: 352 + 156 b (kick + drift pass, read x,u,omega,r,E,G, bond count
    and ids 4b, rest lengths 8b, bond history ft, mn, mt read and written
    144b, write F,T), plus the unbonded-contact terms of the force pass:
    list count and ids (4 + 4n), contact count read and written (8) and
    contact history read and written (56c)."""
    return 352. + 156. * bbar + 12. + 4. * nbar + 56. * cbar


# Bytes per particle-step (8 B per double, 4 B per id/count), with n = list
# neighbours, c = tracked contacts, w = wall contacts per particle.
def bytes_reference_passes(nbar, cbar, wbar):
    """Algorithmic bytes of the work the fused step kernel does, in the
    reference data model: kick + drift pass 208 B (read x,u,omega,F,T 120 +
    m,I 16, write x,u,omega 72) plus force pass 172 + 4n + 56c + 48w B
    (read x,u,omega 72 + r,m,E,nu,G 40 + list 4 + 4n + history 4 + 28c,
    write F,T 48 + history 4 + 28c, wall history 48w)."""
    return 380. + 4. * nbar + 56. * cbar + 48. * wbar


def bytes_reference_model(nbar, cbar, wbar, K):
    """Whole step in the reference data model: the two passes plus the
    amortised rebuild (read x,r 32, write count and ids 4 + 4n) every K steps."""
    return bytes_reference_passes(nbar, cbar, wbar) + (36. + 4. * nbar) / K


def bytes_fused_step(nbar, cbar, wbar):
    """Compulsory traffic of this library's fused step kernel: reads x,r (32),
    u,omega (48), m,I (16), row count (4), row ids (4n), contact word (4), old
    contacts (28c), wall history (24w); writes x,r (32), u,omega (48),
    contact word (4), new contacts (28c), wall history (24w)."""
    return 188. + 4. * nbar + 56. * cbar + 48. * wbar
