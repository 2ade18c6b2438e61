"""Ensembles of independent systems in one context (one fused launch per
step for all systems).  Each system must evolve bit for bit as it does alone
in its own context: same kernels, same per-system grid and cell order, same
neighbour rows, so every floating-point operation is the same.  The single
system itself is checked against the oracle in test_gpu_parity.py."""
from __future__ import annotations

import numpy as np
import pytest

from cabanadem_amd import abi
from cabanadem_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "omega", "force", "torque", "ss_idx", "ss_cnt", "ss_x", "ss_y", "ss_z",
          "sw_x", "sw_y", "sw_z")


def small_ensemble(nsys=7, seed=5):
    """Dense and fast, so that particle and wall contacts start at once:
    spacing 1.9 r_max (polydisperse pairs overlap), u ~ N(0, 0.5 m/s),
    walls 0.95 r_max outside the lattice (the largest spheres start on them)."""
    return W.ensemble(n_systems=nsys, seed=seed, nmin=200, nmax=3000, spacing_r=1.9,
                      vsigma=0.5, wall_offset=0.95 * 1.2e-3)


def contact_means(sysm):
    h = sysm.diag_history()
    return h[:, 7].mean(), h[:, 8].mean()


def split(st, offsets):
    """Per-system views; tracked-contact ids made local to the system."""
    out = []
    for k in range(len(offsets) - 1):
        d = {f: getattr(st, f)[offsets[k]:offsets[k + 1]] for f in FIELDS}
        idx = d["ss_idx"].copy()
        idx[idx >= 0] -= offsets[k]
        d["ss_idx"] = idx
        out.append(d)
    return out


def assert_same(a, b, where):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f"{where}: field {f} differs"


@pytest.mark.parametrize("mode,block,rows", [(0, 0, 0), (1, 0, 0), (0, 64, 16), (1, 256, 4)])
def test_ensemble_equals_independent_runs_bitwise(gpu_lib, mode, block, rows):
    """Also at other force tiles (each system's tile table for that size)
    and other cell tiles (each system's grid and key range re-laid)."""
    wl = small_ensemble()
    ens = wl.system(mode=mode)
    ens.set_shape(block, rows)
    ens.set_diag_every(25)
    ens.step(300)
    c, w = contact_means(ens)
    assert c > 0.1 and w > 0.001, (c, w)  # particle and wall contacts occur
    got = split(ens.download(), ens.offsets)
    for k in range(len(wl.systems)):
        one = wl.single(k, mode=mode)
        one.set_shape(block, rows)
        one.step(300)
        want = one.download()
        assert_same(got[k], {f: getattr(want, f) for f in FIELDS}, f"system {k}")
        one.close()
    # the run exercised contacts with walls and between particles
    sd = ens.system_diagnostics()
    assert np.all(sd[:, 0] > 0)
    ens.close()


def test_persistent_launch_is_bitwise_equal(gpu_lib):
    wl = small_ensemble(nsys=5, seed=9)
    a = wl.system()
    b = wl.system()
    b.set_launch(16)  # 16 workgroups striding over all tiles
    a.step(120)
    b.step(120)
    sa, sb = a.download(), b.download()
    for f in FIELDS:
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
    a.close()
    b.close()


def test_system_diagnostics_match_single_contexts(gpu_lib):
    wl = small_ensemble(nsys=4, seed=2)
    ens = wl.system()
    ens.step(50)
    sd = ens.system_diagnostics()
    for k in range(4):
        one = wl.single(k)
        one.step(50)
        d = one.diagnostics()
        # same particle states, sums over the same tiles in the same order
        assert np.array_equal(sd[k], d), (k, sd[k], d)
        one.close()
    ens.close()


def test_ensemble_argument_errors(gpu_lib):
    wl = small_ensemble(nsys=3, seed=1)
    ens = wl.system()
    with pytest.raises(abi.DemError) as e:
        ens.upload_walls(wl.walls[0])
    assert e.value.code == -8
    with pytest.raises(abi.DemError) as e:
        ens.upload_system_walls(3, wl.walls[0])
    assert e.value.code == -1
    with pytest.raises(abi.DemError) as e:
        ens.set_system_params(0, 0.0, 0.5, 0.1, 0.1)
    assert e.value.code == -1
    # per-system materials take effect: one system made elastic diverges
    # from its twin, the others do not
    twin = wl.system()
    twin.set_system_params(1, 1.0, 1.0, 0.0, 0.0)
    ens.step(200)
    twin.step(200)
    a = split(ens.download(), ens.offsets)
    b = split(twin.download(), twin.offsets)
    assert np.array_equal(a[0]["x"], b[0]["x"]) and np.array_equal(a[2]["x"], b[2]["x"])
    assert not np.array_equal(a[1]["u"], b[1]["u"])
    ens.close()
    twin.close()


def test_force_pass_shapes_ensemble_are_bitwise_equal(gpu_lib):
    """The split force pass (detection kernel + force kernel) and the wave
    pass over the tile table of an ensemble equal the default fused pass
    (each lane walking its own contacts), graphs and persistent grid
    included."""
    wl = small_ensemble(nsys=5, seed=9)
    a = wl.system()
    b = wl.system()
    b.set_force_pass(abi.FORCE_PASS_SPLIT)
    b.set_launch(24)
    b.set_graph(True)
    c = wl.system()
    c.set_force_pass(abi.FORCE_PASS_WAVE)
    c.set_launch(20)
    # other force tiles: their own tile tables
    d = wl.system()
    d.set_force_pass(abi.FORCE_PASS_WAVE)
    d.set_shape(64, 0)
    d.set_graph(True)
    e = wl.system()
    e.set_force_pass(abi.FORCE_PASS_SPLIT)
    e.set_shape(256, 0)
    e.set_launch(12)
    for nsteps in (17, 63):
        for s_ in (a, b, c, d, e):
            s_.step(nsteps)
        sa = a.download()
        for s_ in (b, c, d, e):
            got = s_.download()
            for f in FIELDS:
                assert np.array_equal(getattr(sa, f), getattr(got, f)), (nsteps, f)
    assert sa.ss_cnt.sum() > 0
    da = a.system_diagnostics()
    assert np.array_equal(da, b.system_diagnostics())
    for s_ in (d, e):  # per-tile partials of other sizes: rounding only
        dd = s_.system_diagnostics()
        assert np.all(np.abs(dd - da) <= 1e-13 * np.abs(da).max(axis=0)), dd - da
    for s_ in (a, b, c, d, e):
        s_.close()


def test_graph_replay_is_bitwise_equal(gpu_lib):
    """Captured 2K-step cycles replayed several times (with the program's own
    eager steps, downloads and synchronisations in between) equal eager steps."""
    wl = small_ensemble(nsys=4, seed=4)
    a = wl.system()
    b = wl.system()
    b.set_graph(True)
    for nsteps in (13, 47, 40, 91):  # heads, whole cycles (K = 20) and tails
        a.step(nsteps)
        b.step(nsteps)
        sa, sb = a.download(), b.download()
        for f in FIELDS:
            assert np.array_equal(getattr(sa, f), getattr(sb, f)), (nsteps, f)
    assert a.counters().steps == b.counters().steps == 191
    assert a.counters().rebuilds == b.counters().rebuilds
    a.close()
    b.close()


def test_autotune_restores_state_and_applies_choice(gpu_lib):
    wl = small_ensemble(nsys=4, seed=6)
    a = wl.system()
    a.step(30)
    before = a.download()
    r = a.autotune(trial_steps=40)
    after = a.download()
    for f in FIELDS:  # the trials left no trace
        assert np.array_equal(getattr(before, f), getattr(after, f)), f
    assert a.counters().steps == 30
    trials = r.trials()
    assert r.n_trials == len(trials) >= 8
    ok = [t for t in trials if t["ms_per_step"] > 0]
    assert min(t["ms_per_step"] for t in ok) == r.ms_per_step
    # the chosen configuration is now in force: same run as setting it by hand
    b = wl.system()
    b.step(30)
    b.set_params(rebuild_every=r.rebuild_every)
    b.set_launch(r.grid_blocks)
    b.set_graph(bool(r.use_graph))
    b.set_force_pass(r.force_pass)
    b.set_shape(r.block, r.tile_rows)
    assert a.shape() == (r.block, r.tile_rows)
    assert {t["force_pass"] for t in trials} == {0, 1, 2}
    assert {t["block"] for t in trials} == {64, 128, 256}
    assert {t["tile_rows"] for t in trials} == {4, 8, 16}
    a.step(100)
    b.step(100)
    sa, sb = a.download(), b.download()
    for f in FIELDS:
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
    a.close()
    b.close()
