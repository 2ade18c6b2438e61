"""Shared set-ups for the GPU parity tests (the oracle is the checker)."""
from __future__ import annotations

import numpy as np

import oracle_py as O
from cabanadem_amd import abi
from scenarios import box_walls


def normwise(a, b):
    """||a - b||_2 / ||b||_2 (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return 0.0 if d == 0.0 else d / (nb if nb > 0 else 1.0)


def history_map(st):
    """(i, j) -> (hx, hy, hz) of the tracked pp contacts (order-free)."""
    out = {}
    for i in range(st.n):
        for k in range(int(st.ss_cnt[i])):
            j = int(st.ss_idx[i, k])
            out[(i, j)] = (st.ss_x[i, k], st.ss_y[i, k], st.ss_z[i, k])
    return out


def compare_history(gpu_st, ref_st, tol):
    hg, hr = history_map(gpu_st), history_map(ref_st)
    assert set(hg) == set(hr), (len(set(hg) ^ set(hr)), "tracked contact sets differ")
    keys = sorted(hr)
    if keys:
        a = np.array([hg[k] for k in keys])
        b = np.array([hr[k] for k in keys])
        assert normwise(a, b) <= tol, normwise(a, b)
    return len(keys)


def granular_box(nx, ny, nz, *, spacing_r=2.05, r=1e-3, rmin=None, rmax=None, jitter_r=0.02,
                 vsigma=0.05, omega_sigma=0.0, seed=1, E=1e7, nu=0.3, rho=2600., max_cnts=16,
                 wall_offset=None):
    """Jittered lattice in a six-wall box (the plain-DEM small configuration's shape)."""
    rmin = r if rmin is None else rmin
    rmax = r if rmax is None else rmax
    sp = spacing_r * r
    g = abi.generate_lattice(nx, ny, nz, sp, (0., 0., 0.), jitter_r * r, rmin, rmax, vsigma,
                             rho, E, nu, seed)
    n = nx * ny * nz
    st = abi.ParticleState(n, max_cnts=max_cnts, max_walls=6)
    st.x[:] = g["x"]
    st.u[:] = g["u"]
    st.rad[:], st.m[:], st.I[:] = g["rad"], g["m"], g["I"]
    st.E[:], st.nu[:], st.G[:] = g["E"], g["nu"], g["G"]
    if omega_sigma > 0:
        rng = np.random.default_rng(seed + 100)
        st.omega[:] = rng.normal(0., omega_sigma, size=(n, 3))
    off = sp if wall_offset is None else wall_offset  # wall distance from the outer lattice sites
    lo = np.array([-off] * 3)
    hi = np.array([(nx - 1) * sp + off, (ny - 1) * sp + off, (nz - 1) * sp + off])
    w = box_walls(lo, hi, E, nu, E / (2 * (1 + nu)))
    return st, w, lo, hi


def make_system(st, w, *, dt, rc, lo, hi, cor_pp=0.9, cor_pw=0.9, fr_pp=0.3, fr_pw=0.3,
                gravity=(0., 0., 0.), rebuild_every=1, mode=0, max_contacts=16, skin=None,
                max_neighbors=0):
    if skin is None:
        # rows reused for several steps need a skin: the cover the cutoff
        # leaves beyond the largest contact distance 2 r_max
        skin = max(rc - 2. * float(st.rad.max()), 0.) if rebuild_every > 1 else 0.
    sysm = abi.System(st.n, dt=dt, cutoff=rc, lo=lo, hi=hi, n_walls=0 if w is None else w.n,
                      max_contacts=max_contacts, cor_pp=cor_pp, cor_pw=cor_pw,
                      friction_pp=fr_pp, friction_pw=fr_pw, gravity=gravity, skin=skin,
                      rebuild_every=rebuild_every, physics_mode=mode, max_neighbors=max_neighbors)
    if w is not None:
        sysm.upload_walls(w)
    sysm.upload(st)
    return sysm


def oracle_params(dt, cor_pp=0.9, cor_pw=0.9, fr_pp=0.3, fr_pw=0.3, gravity=(0., 0., 0.), mode=0):
    return O.params(dt=dt, cor_pp=cor_pp, cor_pw=cor_pw, fr_pp=fr_pp, fr_pw=fr_pw,
                    gravity=gravity, mode=mode)


def sorted_rows(counts, ids):
    """Neighbour rows as a sortable array: each row's first counts[i] ids in
    ascending order, the rest of the row filled with INT32_MAX (order-free
    comparison of two lists, vectorised for millions of rows)."""
    counts = np.asarray(counts)
    ids = np.asarray(ids)
    cols = np.arange(ids.shape[1])[None, :]
    filled = np.where(cols < counts[:, None], ids, np.iinfo(np.int32).max)
    return np.sort(filled, axis=1)


def same_neighbor_sets(a, b):
    """(counts, ids) pairs of two lists hold the same sets, row by row."""
    (ca, ia), (cb, ib) = a, b
    if not np.array_equal(np.asarray(ca), np.asarray(cb)):
        return False
    w = int(max(np.max(ca, initial=0), 1))
    return np.array_equal(sorted_rows(ca, ia)[:, :w], sorted_rows(cb, ib)[:, :w])


def history_arrays(st):
    """Tracked pp contacts as (keys i * n + j ascending, histories [k][3]):
    the order-free form of the (i, j) -> (hx, hy, hz) map, vectorised."""
    n = st.n
    cols = np.arange(st.ss_idx.shape[1])[None, :]
    live = cols < st.ss_cnt[:, None]
    i = np.broadcast_to(np.arange(n)[:, None], st.ss_idx.shape)[live].astype(np.int64)
    key = i * n + st.ss_idx[live].astype(np.int64)
    h = np.stack([st.ss_x[live], st.ss_y[live], st.ss_z[live]], axis=-1)
    order = np.argsort(key, kind="stable")
    return key[order], h[order]


def compare_history_arrays(gpu_st, ref_st, tol):
    """compare_history for large states: the same tracked (i, j) pairs and
    their histories within tol normwise.  Returns the number of pairs."""
    kg, hg = history_arrays(gpu_st)
    kr, hr = history_arrays(ref_st)
    assert np.array_equal(kg, kr), "tracked contact sets differ"
    if len(kr):
        assert normwise(hg, hr) <= tol, normwise(hg, hr)
    return len(kr)


def compare_diag(diag, diag_ref, tol):
    """KE_trans, KE_rot and momentum histories (rows of dem_get_diag_history
    and the oracle's run) within tol of each column's largest magnitude."""
    assert diag.shape[0] == diag_ref.shape[0] and diag.shape[0] > 0
    assert np.array_equal(diag[:, 0], diag_ref[:, 0])
    worst = 0.
    for k in range(1, 6):
        scale = np.abs(diag_ref[:, k]).max()
        if scale == 0:
            assert np.abs(diag[:, k]).max() == 0, k
            continue
        err = np.abs(diag[:, k] - diag_ref[:, k]).max() / scale
        assert err <= tol, (k, err)
        worst = max(worst, err)
    return worst
