"""The surrogate network's CPU restatement (oracle/gnn_oracle.py) pinned by
central finite differences of its smooth FP64 form, and its BF16 rounding
pinned by known values.  CPU only."""
import numpy as np
import pytest

import gnn_oracle as G


def small_graph(N=14, seed=0, radius=0.45):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, size=(N, 3))
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    adj = (d < radius) & ~np.eye(N, dtype=bool)
    counts = adj.sum(1)
    row = np.concatenate([[0], np.cumsum(counts)])
    recv = np.repeat(np.arange(N), counts)
    send = np.concatenate([np.nonzero(adj[i])[0] for i in range(N)])
    return row, send, recv, x


def random_params(L, seed=1, scale=0.3):
    _, n = G.layout(L)
    rng = np.random.default_rng(seed)
    flat = rng.normal(0, scale, size=n)
    for K, nout, NP, o in G.layout(L)[0]:   # the padded decoder columns stay zero
        o2 = o + K * G.LAT + G.LAT
        W2 = flat[o2:o2 + G.LAT * NP].reshape(G.LAT, NP)
        W2[:, nout:] = 0.
        flat[o2 + G.LAT * NP + nout:o2 + G.LAT * NP + NP] = 0.
    return flat


def test_bf16_rounding_known_values():
    # BF16 keeps 7 fraction bits: spacing 2^-7 at 1.0
    v = np.array([1.0, 1.0078125, 1.00390625, 1.01171875, 1.0050, -3.14159265, 1e-3])
    r = G.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0078125
    assert r[2] == 1.0            # 1 + 2^-8: tie, to the even neighbour 1.0
    assert r[3] == 1.015625       # 1 + 3 2^-8: tie, to the even neighbour 1 + 2^-6
    assert r[4] == 1.0078125      # nearest
    assert r[5] == -3.140625
    assert abs(r[6] - 1e-3) / 1e-3 < 2 ** -8


def test_backward_matches_central_differences():
    L = 2
    row, send, recv, x = small_graph()
    rng = np.random.default_rng(3)
    N, E = len(row) - 1, len(send)
    xn = rng.normal(0, 1, size=(N, G.FEAT))
    xe = rng.normal(0, 1, size=(E, G.FEAT))
    flat = random_params(L)
    d_out = rng.normal(0, 1, size=(N, 3))
    net = G.Net(flat, L, quant=False)
    net.forward(row, send, recv, xn, xe)
    g = net.backward(d_out)
    # probe a spread of parameters in every MLP (weights and biases)
    idx = []
    for K, nout, NP, o in G.layout(L)[0]:
        idx += [o + 3, o + K * G.LAT // 2 + 5, o + K * G.LAT + 7, o + K * G.LAT + G.LAT + 2 * NP + 1,
                o + K * G.LAT + G.LAT + G.LAT * NP + 1]
    # step 1e-4: truncation ~1e-8 relative here, round-off grows below 1e-5
    h = 1e-4
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        yp = G.Net(fp, L, quant=False).forward(row, send, recv, xn, xe)
        ym = G.Net(fm, L, quant=False).forward(row, send, recv, xn, xe)
        fd = ((yp - ym) * d_out).sum() / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])


def test_padding_carries_no_gradient():
    L = 1
    row, send, recv, _ = small_graph(N=8, seed=2, radius=0.6)
    N, E = len(row) - 1, len(send)
    rng = np.random.default_rng(4)
    xn = np.zeros((N, G.FEAT)); xn[:, :20] = rng.normal(size=(N, 20))
    xe = np.zeros((E, G.FEAT)); xe[:, :4] = rng.normal(size=(E, 4))
    flat = random_params(L)
    net = G.Net(flat, L, quant=False)
    net.forward(row, send, recv, xn, xe)
    g = net.backward(rng.normal(size=(N, 3)))
    K, nout, NP, o = G.layout(L)[0][0]
    gW1 = g[o:o + K * G.LAT].reshape(K, G.LAT)
    assert np.all(gW1[20:] == 0.)          # zero input features
    K, nout, NP, o = G.layout(L)[0][-1]
    o2 = o + K * G.LAT + G.LAT
    assert np.all(g[o2:o2 + G.LAT * NP].reshape(G.LAT, NP)[:, 3:] == 0.)  # padded outputs


class _Fp32Accumulating(G.Net):
    """The BF16 network with its products accumulated in FP32 (as the MFMA
    units do) instead of FP64, rounding at the same points."""

    def _mlp(self, k, x, res=None):
        W1, b1, W2, b2, _ = self.mlps[k]
        f = np.float32
        pre = x.astype(f) @ W1.astype(f) + b1.astype(f)
        h = self.q(np.maximum(pre, 0.).astype(np.float64))
        y = h.astype(f) @ W2.astype(f) + b2.astype(f)
        if res is not None:
            y = y + res.astype(f)
        return y.astype(np.float64), h


def test_bf16_network_gradient_sensitivity_to_accumulation():
    """Why the GPU's network gradients are compared to the oracle's at 1e-1:
    with BF16-stored activations, only changing the forward's accumulation
    from FP64 to FP32 moves the output by ~0.4 % and the gradients by ~5 %
    (a few ReLU masks and BF16 roundings flip and the backward amplifies
    them).  The GPU backward itself is checked tightly on its own forward."""
    L = 3
    row, send, recv, _ = small_graph(N=300, seed=5, radius=0.18)
    N, E = len(row) - 1, len(send)
    rng = np.random.default_rng(6)
    xn = np.zeros((N, G.FEAT)); xn[:, :20] = rng.normal(size=(N, 20))
    xe = np.zeros((E, G.FEAT)); xe[:, :4] = rng.normal(size=(E, 4))
    flat = random_params(L, scale=0.08)
    d_out = rng.normal(size=(N, 3))
    a = G.Net(flat, L, quant=True)
    b = _Fp32Accumulating(flat, L, quant=True)
    ya = a.forward(row, send, recv, G.bf16(xn), G.bf16(xe))
    yb = b.forward(row, send, recv, G.bf16(xn), G.bf16(xe))
    fwd = np.linalg.norm(ya - yb) / np.linalg.norm(ya)
    ga, gb = a.backward(d_out), b.backward(d_out)
    grad = np.linalg.norm(ga - gb) / np.linalg.norm(ga)
    assert 1e-4 < fwd < 2e-2
    assert grad > 3 * fwd          # the gradients are the more sensitive
    assert grad < 1e-1
