"""GPU parity against the CPU oracle, through the C ABI.

Bar (from the project's contract): neighbour sets bit-exact; per-step
forces, torques and tangential histories within 1e-12 relative (normwise)
on identical states; integrator stages bit-exact; kinetic-energy and
momentum histories within 1e-8 relative over a run."""
import numpy as np
import pytest

import oracle_py as O
from cabanadem_amd import abi
from gpu_helpers import (compare_history, granular_box, make_system, normwise,
                         oracle_params)

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_RUN = 1e-8


def _rows(counts, ids):
    return [tuple(sorted(ids[i, : counts[i]].tolist())) for i in range(len(counts))]


def _dense_state(seed=3, mode=0, steps=40):
    """A packing that is mid-collision: overlapping contacts, non-zero
    tangential histories and wall contacts, prepared by the oracle."""
    st, w, lo, hi = granular_box(6, 6, 6, spacing_r=1.97, jitter_r=0.03, vsigma=0.2,
                                 omega_sigma=50., seed=seed, wall_offset=0.97e-3)
    dt, rc = 1e-6, 2.4e-3
    prm = oracle_params(dt, gravity=(0., 0., -9.81), mode=mode)
    O.run(st, w, prm, nsteps=steps, rc=rc, lo=lo - 1e-3, hi=hi + 1e-3, rebuild_every=1, max_n=64)
    return st, w, lo - 1e-3, hi + 1e-3, dt, rc


def test_neighbor_sets_bit_exact(gpu_lib):
    st, w, lo, hi = granular_box(12, 10, 9, spacing_r=2.05, jitter_r=0.02, seed=5)
    rc = 2.2e-3
    sysm = make_system(st, w, dt=1e-6, rc=rc, lo=lo, hi=hi)
    sysm.build_neighbors()
    c_g, i_g = sysm.neighbors(64)
    c_o, i_o = O.neighbors_cell(st.x, rc, lo, hi)
    c_b, i_b = O.neighbors_brute(st.x, rc)
    assert _rows(c_g, i_g) == _rows(c_o, i_o) == _rows(c_b, i_b)
    assert c_g.sum() > 0


def test_neighbor_sets_ties_clamped_and_after_resort(gpu_lib):
    # exact ties (spacing == cutoff in binary fractions), particles outside
    # the binning box, then a re-sort after motion
    rc = 0.25
    g = np.arange(10) * rc
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = np.concatenate([x, [[4.0, 4.0, 4.0], [4.0, 4.0, 4.25], [-2.0, 0.0, 0.0]]])
    n = len(x)
    st = abi.ParticleState(n, max_cnts=16, max_walls=6)
    st.x[:] = x
    st.rad[:] = 0.1
    sysm = make_system(st, None, dt=1e-3, rc=rc, lo=(0., 0., 0.), hi=(2., 2., 2.))
    sysm.build_neighbors()
    c_g, i_g = sysm.neighbors(64)
    c_b, i_b = O.neighbors_brute(x, rc)
    assert _rows(c_g, i_g) == _rows(c_b, i_b)
    rng = np.random.default_rng(0)
    st.x[:] = x + rng.uniform(-0.05, 0.05, size=x.shape)
    sysm.upload(st, fields=("x",))
    for _ in range(2):
        sysm.build_neighbors()
    c_g, i_g = sysm.neighbors(64)
    c_b, i_b = O.neighbors_brute(st.x, rc)
    assert _rows(c_g, i_g) == _rows(c_b, i_b)


@pytest.mark.parametrize("mode", [0, 1])
def test_forces_torques_history_identical_state(gpu_lib, mode):
    st, w, lo, hi, dt, rc = _dense_state(mode=mode)
    assert st.ss_cnt.sum() > 100  # mid-collision, histories carried
    ref = st.copy()
    prm = oracle_params(dt, gravity=(0., 0., -9.81), mode=mode)
    c, i = O.neighbors_cell(ref.x, rc, lo, hi)
    O.compute_forces(ref, w, c, i, prm)

    sysm = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=(0., 0., -9.81), mode=mode)
    sysm.build_neighbors()
    sysm.compute_forces()
    got = sysm.download(abi.ParticleState(st.n, 16, 6))
    assert normwise(got.force, ref.force) <= TOL_STEP
    assert normwise(got.torque, ref.torque) <= TOL_STEP
    assert np.abs(ref.torque).max() > 0  # wall contacts (and pp in corrected mode) act
    ncont = compare_history(got, ref, TOL_STEP)
    assert ncont > 100
    for comp in ("sw_x", "sw_y", "sw_z"):
        assert normwise(getattr(got, comp), getattr(ref, comp)) <= TOL_STEP
    assert np.abs(ref.sw_x).max() + np.abs(ref.sw_y).max() > 0
    assert np.array_equal(got.ss_cnt, ref.ss_cnt)


def test_stagewise_calls_match_fused_forces(gpu_lib):
    st, w, lo, hi, dt, rc = _dense_state(seed=4)
    a = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=(0., 0., -9.81))
    b = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=(0., 0., -9.81))
    for s in (a, b):
        s.build_neighbors()
    a.compute_forces()
    b.zero_forces()
    b.update_tangential()
    b.force_pp()
    b.force_pw()
    ga, gb = a.download(), b.download()
    assert np.array_equal(ga.force, gb.force) and np.array_equal(ga.torque, gb.torque)
    assert np.array_equal(ga.ss_cnt, gb.ss_cnt)


def test_update_tangential_drops_separated_contacts(gpu_lib):
    st, w, lo, hi, dt, rc = _dense_state(seed=6)
    ref = st.copy()
    st_sep = st.copy()
    st_sep.x[:] *= 1.05  # pull every pair apart a little
    ref.x[:] = st_sep.x
    O.update_tangential(ref)
    sysm = make_system(st_sep, w, dt=dt, rc=rc, lo=lo, hi=hi)
    sysm.update_tangential()
    got = sysm.download()
    compare_history(got, ref, 0.0)


def test_integrator_stages_bit_exact(gpu_lib):
    st, w, lo, hi, dt, rc = _dense_state(seed=7)
    rng = np.random.default_rng(1)
    st.force[:] = rng.normal(0., 1e-2, size=st.force.shape)
    st.torque[:] = rng.normal(0., 1e-6, size=st.torque.shape)
    sysm = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi)
    ref = st.copy()
    for stage in (1, 2, 3):
        O.stage(ref, stage, dt)
        getattr(sysm, f"stage{stage}")()
    got = sysm.download()
    assert np.array_equal(got.x, ref.x)
    assert np.array_equal(got.u, ref.u)
    assert np.array_equal(got.omega, ref.omega)


@pytest.mark.parametrize("mode,K,block,rows", [(0, 20, 0, 0), (1, 20, 0, 0), (0, 1, 0, 0),
                                               (0, 20, 64, 4), (1, 20, 256, 16)])
def test_run_ke_momentum_history(gpu_lib, mode, K, block, rows):
    # the small plain-DEM configuration's physics at 12x12x14 = 2016 spheres,
    # at the default tile shapes and at other force tiles and cell tiles
    st, w, lo, hi = granular_box(12, 12, 14, spacing_r=2.05, jitter_r=0.02, vsigma=0.05,
                                 seed=11)
    dt, rc = 1e-6, 2.2e-3
    nsteps, every = 600, 50
    g = (0., 0., -9.81)
    ref = st.copy()
    diag_ref, _ = O.run(ref, w, oracle_params(dt, gravity=g, mode=mode), nsteps=nsteps, rc=rc,
                        lo=lo, hi=hi, rebuild_every=K, max_n=64, diag_every=every)
    sysm = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=g, rebuild_every=K, mode=mode)
    sysm.set_shape(block, rows)
    assert sysm.shape() == (block or 128, rows or 8)
    sysm.set_diag_every(every)
    sysm.step(nsteps)
    diag = sysm.diag_history()
    assert diag.shape[0] == diag_ref.shape[0] == nsteps // every
    assert np.array_equal(diag[:, 0], diag_ref[:, 0])
    for k in range(1, 6):  # KE_t, KE_r, px, py, pz
        scale = np.abs(diag_ref[:, k]).max()
        if scale == 0:
            continue
        assert np.abs(diag[:, k] - diag_ref[:, k]).max() <= TOL_RUN * scale, k
    got = sysm.download()
    assert ref.ss_cnt.sum() > 0
    # Per-particle states: rounding-order differences grow through the
    # contacts over the run.  The bar is the oracle's own sensitivity: the
    # oracle restarted from a state one ulp away in u (10x margin), or the
    # run-level 1e-8, whichever is larger.
    pert = st.copy()
    pert.u[:] = np.nextafter(pert.u, np.inf)
    O.run(pert, w, oracle_params(dt, gravity=g, mode=mode), nsteps=nsteps, rc=rc, lo=lo, hi=hi,
          rebuild_every=K, max_n=64)
    self_u = normwise(pert.u, ref.u)
    self_x = normwise(pert.x - st.x, ref.x - st.x)
    assert normwise(got.u, ref.u) <= max(TOL_RUN, 10. * self_u)
    assert normwise(got.x - st.x, ref.x - st.x) <= max(TOL_RUN, 10. * self_x)


@pytest.mark.parametrize("mode", [0, 1])
def test_force_pass_shapes_are_bitwise_equal(gpu_lib, mode):
    """The three force-pass shapes -- fused with each lane walking its own
    contacts (the default), split (a detection kernel, then the force kernel
    over its contact lists) and one kernel with the contact arithmetic
    spread over each wave -- at force tiles of 64, 128 and 256 slots.  A
    dense run with rebuilds every 5 steps and sampled diagnostics, then a
    stage-wise force pass: every particle field bitwise equal; the sampled
    sums (per-tile partials) bitwise equal at the same tile size and within
    1e-13 across tile sizes."""
    st, w, lo, hi, dt, rc = _dense_state(seed=8, mode=mode)
    g = (0., 0., -9.81)
    kw = dict(dt=dt, rc=rc, lo=lo, hi=hi, gravity=g, mode=mode, rebuild_every=5)
    shapes = [(block, p) for block in (128, 64, 256)
              for p in (abi.FORCE_PASS_FUSED, abi.FORCE_PASS_SPLIT, abi.FORCE_PASS_WAVE)]
    systems = []
    for block, shape in shapes:
        s_ = make_system(st, w, **kw)
        s_.set_force_pass(shape)
        s_.set_shape(block, 0)
        s_.set_diag_every(7)
        systems.append(s_)
    for nsteps in (1, 23, 36):
        for s_ in systems:
            s_.step(nsteps)
    fields = ("x", "u", "omega", "force", "torque", "ss_idx", "ss_cnt", "ss_x", "ss_y", "ss_z",
              "sw_x", "sw_y", "sw_z")
    ref = systems[0].download()
    d0 = systems[0].diag_history()
    assert ref.ss_cnt.sum() > 100 and ref.ss_cnt.max() >= 4
    for (block, _), s_ in zip(shapes[1:], systems[1:]):
        d = s_.diag_history()
        if block == 128:
            assert np.array_equal(d0, d)
        else:
            assert np.array_equal(d0[:, 0], d[:, 0])
            for k in range(1, d0.shape[1]):
                scale = max(np.abs(d0[:, k]).max(), 1e-300)
                assert np.abs(d[:, k] - d0[:, k]).max() <= 1e-13 * scale, (block, k)
        got = s_.download()
        for f in fields:
            assert np.array_equal(getattr(ref, f), getattr(got, f)), (block, f)
    for s_ in systems:
        s_.build_neighbors()
        s_.compute_forces()
    ref = systems[0].download()
    for (block, _), s_ in zip(shapes[1:], systems[1:]):
        got = s_.download()
        for f in fields:
            assert np.array_equal(getattr(ref, f), getattr(got, f)), (block, f)
    for s_ in systems:
        s_.close()


def test_tile_shape_arguments(gpu_lib):
    st, w, lo, hi, dt, rc = _dense_state(seed=2)
    sysm = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi)
    assert sysm.shape() == (128, 8)
    for bad in ((32, 0), (512, 0), (0, -1), (0, 33)):
        with pytest.raises(abi.DemError):
            sysm.set_shape(*bad)
    assert sysm.shape() == (128, 8)
    sysm.set_shape(256, 1)
    assert sysm.shape() == (256, 1)
    sysm.step(3)
    sysm.set_shape(0, 0)
    assert sysm.shape() == (128, 8)
    sysm.close()


def test_kernel_timing_samples_force_passes(gpu_lib):
    """dem_set_timing(k): every k-th force pass is timed, passes that sample
    diagnostics never (and they do not advance the count); every rebuild is
    timed; timing does not change the run."""
    st, w, lo, hi, dt, rc = _dense_state(seed=6)
    kw = dict(dt=dt, rc=rc, lo=lo, hi=hi, rebuild_every=5)
    a = make_system(st, w, **kw)
    b = make_system(st, w, **kw)
    b.set_timing(1)
    b.step(12)
    ms, n = b.timing()
    assert n == [12, 3] and ms[0] > 0 and ms[1] > 0
    b.set_timing(4)
    b.step(12)  # passes 0, 4, 8 of the call; rebuilds at steps 15, 20
    assert b.timing()[1] == [3, 2]
    b.set_diag_every(4)  # global steps 24, 28, 32 sample diagnostics
    b.set_timing(4)
    b.step(12)  # eligible: 25, 26, 27, 29, 30, 31, 33, 34, 35 -> 25, 30, 35 timed
    assert b.timing()[1][0] == 3
    b.set_timing(False)
    a.step(12)
    a.step(12)
    a.set_diag_every(4)
    a.step(12)
    assert np.array_equal(a.diag_history(), b.diag_history())
    sa, sb = a.download(), b.download()
    for f in ("x", "u", "omega", "force", "torque"):
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
    a.close()
    b.close()
    # automatic cadence: the passes that check the displacements are not timed
    c = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, rebuild_every=abi.REBUILD_AUTO,
                    skin=rc - 2e-3)
    c.step(5)
    checks0 = c.counters().displacement_checks
    c.set_timing(1)
    c.step(40)
    checks = c.counters().displacement_checks - checks0
    assert checks > 0 and c.timing()[1][0] == 40 - checks
    c.close()


def test_automatic_rebuild_cadence_matches_every_step_rebuilds(gpu_lib):
    """rebuild_every = DEM_REBUILD_AUTO: the rows are rebuilt only when the
    displacement budget (skin/2) may run out, and the run equals the oracle
    rebuilding every step (the reference's cadence) within the run bar.
    Spheres start at 0.5 m/s so the budget runs out several times."""
    st, w, lo, hi = granular_box(12, 12, 12, spacing_r=2.05, jitter_r=0.02, vsigma=0.5, seed=21)
    dt, rc = 1e-6, 2.4e-3  # skin = rc - 2 r = 0.4 mm, budget 0.2 mm
    nsteps, every = 1500, 100
    g = (0., 0., -9.81)
    ref = st.copy()
    diag_ref, _ = O.run(ref, w, oracle_params(dt, gravity=g), nsteps=nsteps, rc=rc, lo=lo, hi=hi,
                        rebuild_every=1, max_n=64, diag_every=every)
    sysm = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=g, rebuild_every=abi.REBUILD_AUTO,
                       skin=rc - 2e-3)
    sysm.set_diag_every(every)
    for chunk in (1, 499, 1000):  # call boundaries inside the budget too
        sysm.step(chunk)
    c = sysm.counters()
    # 0.2 mm at up to ~1.5 m/s: a rebuild every ~100 steps or more, not every step
    assert 3 <= c.rebuilds <= nsteps // 20, c.rebuilds
    # about three checks per build once under way (a doubling series from rest first)
    assert c.displacement_checks <= 5 * c.rebuilds + 12, (c.displacement_checks, c.rebuilds)
    diag = sysm.diag_history()
    assert np.array_equal(diag[:, 0], diag_ref[:, 0])
    for k in range(1, 6):
        scale = np.abs(diag_ref[:, k]).max()
        assert np.abs(diag[:, k] - diag_ref[:, k]).max() <= TOL_RUN * scale, k
    got = sysm.download()
    assert ref.ss_cnt.sum() > 0
    # per-particle: within ten times the oracle's own sensitivity (a restart
    # one ulp away in u), or the run bar
    pert = st.copy()
    pert.u[:] = np.nextafter(pert.u, np.inf)
    O.run(pert, w, oracle_params(dt, gravity=g), nsteps=nsteps, rc=rc, lo=lo, hi=hi,
          rebuild_every=1, max_n=64)
    assert normwise(got.u, ref.u) <= max(TOL_RUN, 10. * normwise(pert.u, ref.u))
    sysm.close()


def test_automatic_cadence_is_the_same_in_every_force_pass_shape(gpu_lib):
    """The displacement check of the automatic cadence runs inside whichever
    kernels write the step (the fused kernel, the gather pass and its work
    lists, the tile-spread pass): every shape and tile size takes the same
    rebuild decisions and ends bitwise equal."""
    st, w, lo, hi = granular_box(10, 10, 10, spacing_r=2.05, jitter_r=0.02, vsigma=0.5, seed=23)
    dt, rc = 1e-6, 2.4e-3
    systems = []
    for block in (128, 64, 256):
        for shape in (abi.FORCE_PASS_FUSED, abi.FORCE_PASS_SPLIT, abi.FORCE_PASS_WAVE):
            s_ = make_system(st, w, dt=dt, rc=rc, lo=lo, hi=hi, gravity=(0., 0., -9.81),
                             rebuild_every=abi.REBUILD_AUTO, skin=rc - 2e-3)
            s_.set_force_pass(shape)
            s_.set_shape(block, 0)
            systems.append(s_)
    for chunk in (1, 299, 400):
        for s_ in systems:
            s_.step(chunk)
    c0 = systems[0].counters()
    assert c0.rebuilds >= 2 and c0.displacement_checks > c0.rebuilds
    ref = systems[0].download()
    for s_ in systems[1:]:
        c = s_.counters()
        assert (c.rebuilds, c.displacement_checks) == (c0.rebuilds, c0.displacement_checks)
        got = s_.download()
        for f in ("x", "u", "omega", "force", "torque", "ss_idx", "ss_cnt"):
            assert np.array_equal(getattr(ref, f), getattr(got, f)), f
    for s_ in systems:
        s_.close()


def test_automatic_rebuild_cadence_needs_a_skin(gpu_lib):
    st, w, lo, hi = granular_box(4, 4, 4, seed=2)
    with pytest.raises(abi.DemError):
        make_system(st, w, dt=1e-6, rc=2.2e-3, lo=lo, hi=hi, rebuild_every=abi.REBUILD_AUTO, skin=0.)
