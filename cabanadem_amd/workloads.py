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
Both are built with the library's counter-based generator (dem_generate_lattice).
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


def _box(nx, ny, nz, sp, E, nu):
    lo = (-sp, -sp, -sp)
    hi = (nx * sp, ny * sp, nz * sp)
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
