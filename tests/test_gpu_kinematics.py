"""dem_set_kinematics_device / dem_get_kinematics_device (device-side state
hand-over, no counterpart in the reference): a context reset from device
arrays steps bit for bit like one given the same state from the host."""
import numpy as np
import pytest
import torch

from cabanadem_amd import abi
from gpu_helpers import granular_box, make_system

pytestmark = pytest.mark.gpu


def _pair(mode):
    st, w, lo, hi = granular_box(8, 7, 6, spacing_r=1.98, jitter_r=0.03, vsigma=0.2, omega_sigma=40.,
                                 seed=11, wall_offset=0.98e-3)
    kw = dict(dt=1e-6, rc=2.4e-3, lo=lo - 1e-3, hi=hi + 1e-3, gravity=(0., 0., -9.81), rebuild_every=5,
              mode=mode)
    return st, make_system(st, w, **kw), make_system(st, w, **kw)


@pytest.mark.parametrize("mode", [0, 1])
def test_device_reset_steps_like_host_reset(gpu_lib, mode):
    st, a, b = _pair(mode)
    # a state to reset to: another run's positions and velocities
    rng = np.random.default_rng(mode)
    x = st.x + rng.uniform(-2e-5, 2e-5, size=st.x.shape)
    u = rng.normal(0., 0.1, size=st.u.shape)
    # host path: everything else cleared
    h = abi.ParticleState(st.n, max_cnts=16, max_walls=6)
    h.x[:], h.u[:] = x, u
    a.upload(h, fields=("x", "u", "omega", "force", "torque", "sw_x", "sw_y", "sw_z", "ss_x", "ss_y", "ss_z",
                        "ss_idx", "ss_cnt"))
    # device path
    xd, ud = torch.from_numpy(x).cuda(), torch.from_numpy(u).cuda()
    torch.cuda.synchronize()
    b.set_kinematics_device(xd.data_ptr(), ud.data_ptr(), True)
    for s in (a, b):
        s.step(37)
    ra, rb = a.download(), b.download()
    assert ra.ss_cnt.sum() > 0 and np.abs(ra.omega).max() > 0   # contacts and wall spin occurred
    for f in ("x", "u", "omega", "force", "torque", "sw_x", "ss_cnt"):
        assert np.array_equal(getattr(ra, f), getattr(rb, f)), f
    # read-back through the device path equals the host download
    xo = torch.empty_like(xd)
    uo = torch.empty_like(ud)
    b.get_kinematics_device(xo.data_ptr(), uo.data_ptr())
    b.sync()
    assert np.array_equal(xo.cpu().numpy(), rb.x) and np.array_equal(uo.cpu().numpy(), rb.u)


def test_device_reset_velocity_only_keeps_positions(gpu_lib):
    st, a, _ = _pair(0)
    a.step(3)
    before = a.download()
    u = np.full_like(st.u, 0.05)
    ud = torch.from_numpy(u).cuda()
    torch.cuda.synchronize()
    a.set_kinematics_device(None, ud.data_ptr(), False)
    after = a.download()
    assert np.array_equal(after.x, before.x) and np.array_equal(after.u, u)
    assert np.array_equal(after.omega, before.omega)
