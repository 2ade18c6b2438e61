"""gnn_radius_graph_build_static (the radius graph in the library's own
device array, no allocation) against gnn_radius_graph_build with a workspace
of the caller's, through the C ABI: integer outputs bit-exact, on two calls
in a row (the second with a smaller input, so it runs over what the first
left in the array), plus the size limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from cabanadem_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _d3(v):
    return (C.c_double * 3)(*v)


def _build_with_workspace(x, radius, lo, hi):
    """The existing entry point, with a workspace and edge arrays of ours."""
    L, s = S.glib(), S.stream()
    n = len(x)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    wb = L.gnn_radius_graph_workspace(n, _d3(lo), _d3(hi), radius)
    assert wb > 0
    work = torch.empty(wb // 4 + 1, dtype=torch.int32, device="cuda")
    row = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    cap = 400 * n  # above the largest degree (about 180 at 3,000 points)
    send, recv, rev = (torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(3))
    ne = C.c_long(-1)
    rc = L.gnn_radius_graph_build(n, S.ptr(xd), radius, _d3(lo), _d3(hi), S.ptr(work), wb, S.ptr(row),
                                  cap, S.ptr(send), S.ptr(recv), S.ptr(rev), C.byref(ne), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    e = ne.value
    return (row.cpu().numpy(), send[:e].cpu().numpy(), recv[:e].cpu().numpy(), rev[:e].cpu().numpy())


def _build_static(x, radius, lo, hi):
    L, s = S.glib(), S.stream()
    n = len(x)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    row = torch.full((n + 1,), -7, dtype=torch.int32, device="cuda")
    cap = 400 * n  # above the largest degree (about 180 at 3,000 points)
    send, recv, rev = (torch.full((cap,), -7, dtype=torch.int32, device="cuda") for _ in range(3))
    ne = C.c_long(-1)
    print(f"static build: N={n}, radius={radius}", flush=True)
    rc = L.gnn_radius_graph_build_static(n, S.ptr(xd), radius, _d3(lo), _d3(hi), S.ptr(row), cap,
                                         S.ptr(send), S.ptr(recv), S.ptr(rev), C.byref(ne), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    e = ne.value
    assert 0 <= e <= cap
    return (row.cpu().numpy(), send[:e].cpu().numpy(), recv[:e].cpu().numpy(), rev[:e].cpu().numpy())


def test_static_workspace_graph_equals_workspace_graph_twice(gpu_lib):
    rng = np.random.default_rng(21)
    lo, hi, radius = (-1e-3, -1e-3, -1e-3), (13e-3, 13e-3, 13e-3), 2.5e-3
    first = rng.uniform(0, 12e-3, size=(3000, 3))      # rows on both sides of 32 senders
    first[100:110] = first[0]                            # coincident points
    second = rng.uniform(-2e-3, 14e-3, size=(1200, 3))   # smaller, some points outside the box
    for x in (first, second):
        want = _build_with_workspace(x, radius, lo, hi)
        got = _build_static(x, radius, lo, hi)
        assert len(want[1]) > 0
        for a, b, name in zip(got, want, ("row", "send", "recv", "rev")):
            assert a.dtype == b.dtype and np.array_equal(a, b), name
        deg = np.diff(want[0])
        if x is first:
            assert (deg > 32).any() and ((deg > 0) & (deg <= 32)).any()


def test_static_workspace_rejects_graphs_beyond_its_size(gpu_lib):
    L = S.glib()
    n = 400_000  # needs more than GNN_GRAPH_SCRATCH_BYTES of workspace
    lo, hi, radius = (0., 0., 0.), (1., 1., 1.), 0.02
    assert L.gnn_radius_graph_workspace(n, _d3(lo), _d3(hi), radius) > 32 * 1024 * 1024
    dummy = torch.zeros(8, dtype=torch.int32, device="cuda")
    ne = C.c_long(-1)
    rc = L.gnn_radius_graph_build_static(n, S.ptr(dummy), radius, _d3(lo), _d3(hi), S.ptr(dummy), 0,
                                         S.ptr(dummy), S.ptr(dummy), S.ptr(dummy), C.byref(ne), S.stream())
    assert rc == -4
    assert ne.value == -1  # rejected before any work
