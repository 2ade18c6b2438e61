"""CPU restatement of the surrogate network (cabanadem_amd/surrogate.py).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP kernels'
forward and backward passes; the product never imports it.  Plain numpy in
FP64.  With quant=True the activations the GPU stores in BF16 (inputs,
weights, hidden layers, the latent node/edge states and the aggregation)
are rounded to BF16 at the same points, so the GPU matches it to the
rounding of its FP32 sums; with quant=False it is the smooth FP64 network
whose gradients tests check by central finite differences.

Parameters use the same flat layout as the GPU: per MLP W1 [K][128],
b1 [128], W2 [128][NP], b2 [NP] (NP = 16 for the 3-wide decoder, else 128),
MLPs in the order node encoder, edge encoder, edge MLPs 0..L-1, node MLPs
0..L-1, decoder.
"""
from __future__ import annotations

import numpy as np

LAT = 128
FEAT = 32


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def layout(L):
    """[(K, nout, NP, offset)] of the MLPs in parameter order."""
    specs = [(FEAT, LAT), (FEAT, LAT)] + [(3 * LAT, LAT)] * L + [(2 * LAT, LAT)] * L + [(LAT, 3)]
    out, off = [], 0
    for K, nout in specs:
        NP = LAT if nout > 16 else 16
        out.append((K, nout, NP, off))
        off += K * LAT + LAT + LAT * NP + NP
    return out, off


def unpack(flat, L):
    mlps = []
    for K, nout, NP, o in layout(L)[0]:
        W1 = flat[o:o + K * LAT].reshape(K, LAT)
        b1 = flat[o + K * LAT:o + K * LAT + LAT]
        o2 = o + K * LAT + LAT
        W2 = flat[o2:o2 + LAT * NP].reshape(LAT, NP)[:, :nout]
        b2 = flat[o2 + LAT * NP:o2 + LAT * NP + NP][:nout]
        mlps.append((W1, b1, W2, b2, NP))
    return mlps


class Net:
    def __init__(self, flat, L, quant=True, quant_grad=False):
        """quant_grad: round the backward pass's GEMM operands to BF16 where
        the GPU does (dY entering the hidden-layer product and the weight
        gradient of the second layer, the hidden gradient dH)."""
        self.L, self.quant, self.quant_grad = L, quant, quant_grad
        self.flat = np.asarray(flat, dtype=np.float64)
        q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
        self.q = q
        self.mlps = [(q(W1), b1.astype(np.float64), q(W2), b2.astype(np.float64), NP)
                     for W1, b1, W2, b2, NP in unpack(self.flat, L)]

    def _mlp(self, k, x, res=None):
        W1, b1, W2, b2, _ = self.mlps[k]
        pre = x @ W1 + b1
        h = self.q(np.maximum(pre, 0.))
        y = h @ W2 + b2
        if res is not None:
            y = y + res
        return y, h

    def forward(self, row, send, recv, xn, xe):
        q, L = self.q, self.L
        N = xn.shape[0]
        c = {"xn": q(xn), "xe": q(xe), "h": [], "e": [], "agg": [], "hid_e": [], "hid_n": []}
        y, c["hid_enc_n"] = self._mlp(0, c["xn"])
        c["h"].append(q(y))
        y, c["hid_enc_e"] = self._mlp(1, c["xe"])
        c["e"].append(q(y))
        for l in range(L):
            h, e = c["h"][l], c["e"][l]
            xin = np.concatenate([e, h[recv], h[send]], axis=1)
            y, hid = self._mlp(2 + l, xin, e)
            e1 = q(y)
            c["e"].append(e1)
            c["hid_e"].append(hid)
            agg = np.zeros((N, LAT))
            np.add.at(agg, recv, e1)
            agg = q(agg)
            c["agg"].append(agg)
            y, hid = self._mlp(2 + L + l, np.concatenate([h, agg], axis=1), h)
            c["h"].append(q(y))
            c["hid_n"].append(hid)
        out, c["hid_dec"] = self._mlp(2 + 2 * L, c["h"][L])
        self.c = c
        self.graph = (row, send, recv)
        return out

    def _mlp_bwd(self, k, x, hid, dy, grads):
        W1, b1, W2, b2, NP = self.mlps[k]
        K, nout = W1.shape[0], W2.shape[1]
        dyq = bf16(dy) if self.quant_grad else dy
        gW2 = hid.T @ dyq
        gb2 = dy.sum(0)
        dh = (dyq @ W2.T) * (hid > 0)
        if self.quant_grad:
            dh = bf16(dh)
        gW1 = x.T @ dh
        gb1 = dh.sum(0)
        dx = dh @ W1.T
        grads[k] = (gW1, gb1, gW2, gb2)
        return dx

    def backward(self, d_out):
        """Gradient of sum(d_out * out) with respect to the parameters (flat)."""
        L, c = self.L, self.c
        row, send, recv = self.graph
        N = c["h"][0].shape[0]
        grads = {}
        d_h = self._mlp_bwd(2 + 2 * L, c["h"][L], c["hid_dec"], d_out, grads)
        d_e = np.zeros_like(c["e"][0])
        for l in reversed(range(L)):
            h, e = c["h"][l], c["e"][l]
            dx = self._mlp_bwd(2 + L + l, np.concatenate([h, c["agg"][l]], axis=1), c["hid_n"][l], d_h, grads)
            d_agg = dx[:, LAT:]
            d_h = d_h + dx[:, :LAT]
            d_e = d_e + d_agg[recv]
            xin = np.concatenate([e, h[recv], h[send]], axis=1)
            dx = self._mlp_bwd(2 + l, xin, c["hid_e"][l], d_e, grads)
            d_e = d_e + dx[:, :LAT]
            np.add.at(d_h, recv, dx[:, LAT:2 * LAT])
            np.add.at(d_h, send, dx[:, 2 * LAT:])
        self._mlp_bwd(1, c["xe"], c["hid_enc_e"], d_e, grads)
        self._mlp_bwd(0, c["xn"], c["hid_enc_n"], d_h, grads)
        flat = np.zeros_like(self.flat)
        for k, (K, nout, NP, o) in enumerate(layout(L)[0]):
            gW1, gb1, gW2, gb2 = grads[k]
            flat[o:o + K * LAT] = gW1.ravel()
            flat[o + K * LAT:o + K * LAT + LAT] = gb1
            o2 = o + K * LAT + LAT
            W2full = np.zeros((LAT, NP))
            W2full[:, :nout] = gW2
            flat[o2:o2 + LAT * NP] = W2full.ravel()
            flat[o2 + LAT * NP:o2 + LAT * NP + nout] = gb2
        return flat
