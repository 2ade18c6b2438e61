"""The benchmark's own arithmetic on the CPU: the byte models behind
roofline.achieved, the surrogate's FLOP counts, and the one-GPU guard."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402
from cabanadem_amd import workloads as W  # noqa: E402


def test_reference_byte_model_sums_its_passes():
    # kick + drift pass: read x,u,omega,F,T (15 doubles) and m,I, write x,u,omega
    kick = 15 * 8 + 2 * 8 + 9 * 8
    # force pass without neighbours, contacts or walls: read x,u,omega and the
    # five material doubles, the row count and the contact count; write F,T
    # and the contact count
    force = 9 * 8 + 5 * 8 + 4 + 4 + 6 * 8 + 4
    assert W.bytes_reference_passes(0., 0., 0.) == kick + force == 380.
    # each neighbour id 4 B, each tracked contact id + 3 doubles read and
    # written, each wall contact 3 doubles read and written
    assert W.bytes_reference_passes(1., 0., 0.) - 380. == 4.
    assert W.bytes_reference_passes(0., 1., 0.) - 380. == 2 * (4 + 3 * 8)
    assert W.bytes_reference_passes(0., 0., 1.) - 380. == 2 * 3 * 8
    # the rebuild is amortised over K steps
    assert W.bytes_reference_model(5., 0., 0., 20) == pytest.approx(
        W.bytes_reference_passes(5., 0., 0.) + (32 + 4 + 4 * 5.) / 20)


def test_fused_kernel_compulsory_bytes():
    # x,r and u,omega read and written, m,I read, row count and contact word
    # read, contact word written
    assert W.bytes_fused_step(0., 0., 0.) == 2 * (32 + 48) + 16 + 4 + 4 + 4 == 188.
    assert W.bytes_fused_step(5.6, 0.03, 0.) == pytest.approx(188. + 4 * 5.6 + 56 * 0.03)


def test_surrogate_flop_counts():
    L, N, E, lat = 10, 64000, 374400, 128
    minimal = bench.forward_flops(L, N, E, receiver_in_kernel=False)
    # encoders (K = 32 then 128), per layer 3 edge products of 128^2, the
    # receiver term once per node, the node MLP (256 + 128) x 128, decoder
    enc = (N + E) * (32 * lat + lat * lat)
    per_layer = E * 3 * lat * lat + N * lat * lat + N * 3 * lat * lat
    dec = N * (lat * lat + lat * 16)
    assert minimal == 2.0 * (enc + L * per_layer + dec)
    executed = bench.forward_flops(L, N, E, receiver_in_kernel=True)
    # layers after the first form the receiver term per edge instead of per node
    assert executed - minimal == 2.0 * (L - 1) * (E - N) * lat * lat


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_json_line_has_no_nan():
    out = bench.finite_or_none({"a": float("nan"), "b": [1.0, float("inf")], "c": {"d": 2.0}})
    assert out == {"a": None, "b": [1.0, None], "c": {"d": 2.0}}


@pytest.mark.parametrize("steps", [1, 5, 20, 37, 10000])
def test_timed_window_always_holds_a_diagnostic_row(steps):
    # rows are sampled at global steps that are multiples of the interval;
    # any window of `steps` consecutive steps must contain at least one
    every = bench.diag_interval(steps)
    for start in (0, 1, 2000, 2005, 2019):
        assert any((start + s) % every == 0 for s in range(steps))
    # a short window is not sampled at every other step
    assert every >= min(steps, 20)


def test_measured_traffic_records_match_workload_mode_and_size():
    """A committed counter record is used only for the workload, arithmetic
    mode and sphere count it was measured on."""
    rec = bench.measured_traffic("plain_large", 0, 4_000_000)
    assert rec is not None and 0.5e9 < rec["bytes_per_launch"] < 2e9
    assert rec["source"].startswith("profiles/")
    assert bench.measured_traffic("plain_large", 1, 4_000_000) is None
    assert bench.measured_traffic("plain_large", 0, 1_000) is None
    assert bench.measured_traffic("plain_dense", 0, 4_000_000) is None
    packed = bench.measured_traffic("plain_packed", 1, 4_000_000)
    assert packed is not None and packed["bytes_per_launch"] > 2 * rec["bytes_per_launch"]
    assert bench.measured_traffic("plain_packed", 0, 4_000_000) is None
    assert bench.measured_traffic("bonded_block", 1, 1_000_000) is not None
    # and only for the force-pass shape and tile shape it was measured with
    assert bench.measured_traffic("plain_large", 0, 4_000_000, (0, 128, 8)) is not None
    assert bench.measured_traffic("plain_large", 0, 4_000_000, (2, 128, 8)) is None
    assert bench.measured_traffic("plain_large", 0, 4_000_000, (0, 256, 8)) is None
