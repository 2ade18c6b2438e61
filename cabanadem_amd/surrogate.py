"""Learned surrogate of the DEM: a message-passing network over the DEM
neighbour graph that predicts each particle's mean acceleration over a
macro-step of 100 DEM steps (no counterpart in the reference).

Every arithmetic step of training and inference runs in the HIP kernels of
libcabanadem_gnn.so (include/cabanadem_gnn.h): fused two-layer MLPs on the
BF16 MFMA units with FP32 accumulation, their backward passes and weight
gradients, the graph aggregation and its adjoints, Adam, the features and
the macro-step integration.  This module lays out the parameters and
activations (device memory from torch, which is only the allocator here),
strings the kernels into the network, and runs the DEM (libcabanadem_hip.so)
as data generator, judge and fallback.

Network (encode-process-decode):
  node encoder   x_i (32: 5 velocities, radius, 4 wall distances)   -> h_i (128)
  edge encoder   x_ij (32: (x_i - x_j)/R, |x_i - x_j|/R)           -> e_ij (128)
  10 layers      e_ij += MLP_e([e_ij, h_i, h_j]);  h_i += MLP_n([h_i, sum_j e_ij])
  decoder        h_i -> a_i (3), the acceleration minus gravity over the
                 macro-step, in units of a_scale
Each MLP is Linear(K, 128) - ReLU - Linear(128, out).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import abi

GNN_LIB_PATH = Path(os.environ.get("CABANADEM_GNN_LIB") or
                    Path(__file__).resolve().parent / "lib" / "libcabanadem_gnn.so")
LAT = 128
FEAT = 32
HIST = 5
GNN_ABI_VERSION = 5

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


class GnnMlpIo(C.Structure):
    _fields_ = [("M", C.c_int), ("nseg", C.c_int), ("src", _vp * 3), ("idx", _vp * 3),
                ("width", C.c_int * 3)]


class GnnMlpWeights(C.Structure):
    _fields_ = [("K", C.c_int), ("nout", C.c_int), ("nout_pad", C.c_int),
                ("W1t", _vp), ("W1", _vp), ("b1", _vp), ("W2t", _vp), ("W2", _vp), ("b2", _vp)]


_SIG = {
    "gnn_abi_version": (C.c_int, []),
    "gnn_mlp_forward": (C.c_int, [C.POINTER(GnnMlpIo), C.POINTER(GnnMlpWeights), _vp, _vp, _vp, _vp,
                                  C.c_int, _vp]),
    "gnn_mlp_backward": (C.c_int, [C.POINTER(GnnMlpIo), C.POINTER(GnnMlpWeights), _vp, _vp, _vp,
                                   C.POINTER(_vp), C.c_int, C.c_int, _vp]),
    "gnn_weight_grad": (C.c_int, [C.POINTER(GnnMlpIo), _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int,
                                  _vp, _vp, C.c_int, _vp]),
    "gnn_segment_sum": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "gnn_edge_grad_to_nodes": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp]),
    "gnn_broadcast_add": (C.c_int, [_vp, _vp, C.c_long, _vp, _vp]),
    "gnn_adam": (C.c_int, [_vp, _vp, _vp, _vp, C.c_long, C.c_float, C.c_float, C.c_float, C.c_float,
                           C.c_float, _vp]),
    "gnn_cast_weights": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "gnn_node_features": (C.c_int, [C.c_int, _vp, _vp, C.c_int, _vp, C.c_float, C.c_float, C.c_float,
                                    _vp, C.c_int, C.c_float, _vp, _vp]),
    "gnn_edge_features": (C.c_int, [C.c_long, _vp, _vp, _vp, C.c_double, _vp, _vp]),
    "gnn_integrate": (C.c_int, [C.c_int, _vp, C.c_float, _vp, C.c_double, _vp, _vp, C.c_int, _vp]),
    "gnn_mse": (C.c_int, [_vp, _vp, C.c_long, _vp, _vp, _vp]),
    "gnn_node_p": (C.c_int, [C.c_int, _vp, C.POINTER(GnnMlpWeights), _vp, _vp]),
    "gnn_edge_mlp_forward": (C.c_int, [C.c_long, _vp, _vp, _vp, _vp, _vp, C.POINTER(GnnMlpWeights), _vp,
                                       _vp, C.c_int, _vp]),
    "gnn_edge_mlp_forward_encoded": (C.c_int, [C.c_long, _vp, C.POINTER(GnnMlpWeights), _vp, _vp, _vp, _vp,
                                               C.POINTER(GnnMlpWeights), _vp, C.c_int, _vp]),
    "gnn_node_layer_forward": (C.c_int, [C.c_int, _vp, _vp, _vp, C.POINTER(GnnMlpWeights), _vp, _vp, _vp,
                                         C.POINTER(GnnMlpWeights), _vp, _vp]),
    "gnn_radius_graph_workspace": (C.c_long, [C.c_int, _dp, _dp, C.c_double]),
    "gnn_radius_graph_count": (C.c_int, [C.c_int, _vp, C.c_double, _dp, _dp, _vp, C.c_long, _vp,
                                         C.POINTER(C.c_long), _vp]),
    "gnn_radius_graph_fill": (C.c_int, [C.c_int, _vp, C.c_double, _dp, _dp, _vp, C.c_long, _vp, C.c_long,
                                        _vp, _vp, _vp, _vp]),
    "gnn_radius_graph_build": (C.c_int, [C.c_int, _vp, C.c_double, _dp, _dp, _vp, C.c_long, _vp, C.c_long,
                                         _vp, _vp, _vp, C.POINTER(C.c_long), _vp]),
    "gnn_radius_graph_build_static": (C.c_int, [C.c_int, _vp, C.c_double, _dp, _dp, _vp, C.c_long, _vp,
                                                _vp, _vp, C.POINTER(C.c_long), _vp]),
}

_GLIB = None


def glib():
    global _GLIB
    if _GLIB is None:
        if not GNN_LIB_PATH.exists():
            raise RuntimeError(f"{GNN_LIB_PATH} is missing: build it with "
                               "`python -m cabanadem_amd.build gnn` (there is no fallback)")
        L = C.CDLL(str(GNN_LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        if L.gnn_abi_version() != GNN_ABI_VERSION:
            raise RuntimeError("libcabanadem_gnn.so ABI mismatch: rebuild it")
        _GLIB = L
    return _GLIB


def _torch():
    import torch  # device memory and the stream only
    return torch


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------------------
# graph
# ---------------------------------------------------------------------------

@dataclass
class Graph:
    """Receiver-major edge list on the device: the edges of node i are
    [row[i], row[i+1]); edge e is send[e] -> recv[e]; rev[e] is the edge
    recv[e] -> send[e]."""
    N: int
    E: int
    row: object
    send: object
    recv: object
    rev: object

    @staticmethod
    def from_rows(counts, ids, device="cuda"):
        torch = _torch()
        counts = np.asarray(counts, dtype=np.int64)
        N = counts.shape[0]
        row = np.zeros(N + 1, dtype=np.int64)
        np.cumsum(counts, out=row[1:])
        mask = np.arange(ids.shape[1])[None, :] < counts[:, None]
        send = ids[mask].astype(np.int64)
        recv = np.repeat(np.arange(N, dtype=np.int64), counts)
        key = recv * N + send            # ascending: rows ascending, ids sorted in a row
        rev = np.searchsorted(key, send * N + recv)
        if len(key) and not np.array_equal(key[np.minimum(rev, len(key) - 1)], send * N + recv):
            raise ValueError("the neighbour graph is not symmetric")
        t = lambda a: torch.from_numpy(a.astype(np.int32)).to(device)
        return Graph(N, int(row[-1]), t(row), t(send), t(recv), t(rev))


class GraphBuilder:
    """Radius graph of a particle set, built on the device
    (gnn_radius_graph_count / _fill): the pairs with |x_i - x_j|^2 <= radius^2,
    the DEM's neighbour criterion, receiver-major with ascending senders.
    The Graph that build() returns views this builder's buffers and stays
    valid until its next build()."""

    def __init__(self, n, radius, lo, hi):
        torch = _torch()
        self.n, self.radius = n, float(radius)
        self.lo = (C.c_double * 3)(*[float(v) for v in lo])
        self.hi = (C.c_double * 3)(*[float(v) for v in hi])
        wb = glib().gnn_radius_graph_workspace(n, self.lo, self.hi, self.radius)
        if wb < 0:
            raise ValueError(f"radius graph: degenerate box or too many cells ({lo}, {hi}, {radius})")
        self.wbytes = wb
        self.work = torch.empty(wb // 4 + 1, dtype=torch.int32, device="cuda")
        self.row = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        self.cap = 0
        self.send = self.recv = self.rev = None

    def build(self, x):
        """x: positions, double [n][3], a device tensor (or a host array,
        copied to the device)."""
        torch = _torch()
        if isinstance(x, np.ndarray):
            x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
        if x.dtype != torch.float64 or tuple(x.shape) != (self.n, 3) or not x.is_contiguous() or not x.is_cuda:
            raise ValueError("radius graph: x must be a contiguous float64 [n][3] device tensor")
        L, s = glib(), stream()
        E = C.c_long(0)
        # count and fill in one call while the edge arrays are large enough
        rc = L.gnn_radius_graph_build(self.n, ptr(x), self.radius, self.lo, self.hi, ptr(self.work),
                                      self.wbytes, ptr(self.row), self.cap, ptr(self.send), ptr(self.recv),
                                      ptr(self.rev), C.byref(E), s)
        if rc < 0:
            _ok(rc, "radius graph build")
        E = int(E.value)
        if rc == 1:
            self.cap = max(E * 5 // 4, 64)
            z = lambda: torch.empty(self.cap, dtype=torch.int32, device="cuda")
            self.send, self.recv, self.rev = z(), z(), z()
            _ok(L.gnn_radius_graph_fill(self.n, ptr(x), self.radius, self.lo, self.hi, ptr(self.work),
                                        self.wbytes, ptr(self.row), E, ptr(self.send), ptr(self.recv),
                                        ptr(self.rev), s), "radius graph fill")
        if self.send is None:
            self.send = self.recv = self.rev = torch.empty(1, dtype=torch.int32, device="cuda")
        return Graph(self.n, E, self.row, self.send[:E], self.recv[:E], self.rev[:E])


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------

class MLP:
    """Slots of one two-layer MLP in the flat FP32 parameter vector:
    W1 [K][128], b1 [128], W2 [128][NP], b2 [NP] (NP = 16 or 128)."""

    def __init__(self, K, nout, off):
        self.K, self.nout = K, nout
        self.NP = 128 if nout > 16 else 16
        self.o_W1 = off
        self.o_b1 = self.o_W1 + K * LAT
        self.o_W2 = self.o_b1 + LAT
        self.o_b2 = self.o_W2 + LAT * self.NP
        self.end = self.o_b2 + self.NP

    def alloc(self, dev):
        torch = _torch()
        bf = torch.bfloat16
        self.W1 = torch.zeros((self.K, LAT), dtype=bf, device=dev)
        self.W1t = torch.zeros((LAT, self.K), dtype=bf, device=dev)
        self.W2 = torch.zeros((LAT, self.NP), dtype=bf, device=dev)
        self.W2t = torch.zeros((self.NP, LAT), dtype=bf, device=dev)

    def views(self, flat):
        return (flat[self.o_W1:self.o_b1].view(self.K, LAT), flat[self.o_b1:self.o_W2],
                flat[self.o_W2:self.o_b2].view(LAT, self.NP), flat[self.o_b2:self.end])

    def weights(self, flat):
        """The gnn_mlp_weights of this MLP (cached: it only holds addresses)."""
        key = (flat.data_ptr(), self.W1t.data_ptr(), self.W2t.data_ptr())
        if getattr(self, "_wkey", None) != key:
            w = GnnMlpWeights()
            w.K, w.nout, w.nout_pad = self.K, self.nout, self.NP
            w.W1t, w.W1, w.W2t, w.W2 = ptr(self.W1t), ptr(self.W1), ptr(self.W2t), ptr(self.W2)
            w.b1 = C.c_void_p(flat.data_ptr() + 4 * self.o_b1)
            w.b2 = C.c_void_p(flat.data_ptr() + 4 * self.o_b2)
            self._w, self._wkey = w, key
        return self._w


def io(M, *segs):
    """segs: (bf16 tensor [rows][width], index tensor or None)."""
    s = GnnMlpIo()
    s.M, s.nseg = M, len(segs)
    for k, (t, idx) in enumerate(segs):
        s.src[k] = t.data_ptr()
        s.idx[k] = None if idx is None else idx.data_ptr()
        s.width[k] = t.shape[1]
    return s


class Network:
    def __init__(self, n_layers=10, seed=0, device="cuda", residual_scale=0.1, factorised=True,
                 receiver_in_kernel=True):
        """factorised: run the edge MLPs' forward through gnn_edge_mlp_forward
        (first layer split into e W1a + h[recv] W1b + h[send] W1c) rather than
        gnn_mlp_forward over the three gathered sources; the same function up
        to FP32 summation order.  receiver_in_kernel: layers after the first
        form h[recv] W1b in the edge kernel; otherwise the node pass stores
        P = h W1b per node for the next layer's edge kernel to read."""
        torch = _torch()
        self.L = n_layers
        self.factorised = factorised
        self.receiver_in_kernel = receiver_in_kernel
        self.dev = device
        off = 0
        self.enc_n = MLP(FEAT, LAT, off); off = self.enc_n.end
        self.enc_e = MLP(FEAT, LAT, off); off = self.enc_e.end
        # parameter order: encoders, the L edge MLPs, the L node MLPs, decoder
        self.edge, self.node = [], []
        for _ in range(n_layers):
            m = MLP(3 * LAT, LAT, off); off = m.end; self.edge.append(m)
        for _ in range(n_layers):
            m = MLP(2 * LAT, LAT, off); off = m.end; self.node.append(m)
        self.dec = MLP(LAT, 3, off); off = self.dec.end
        self.mlps = [self.enc_n, self.enc_e] + self.edge + self.node + [self.dec]
        self.n_params = off
        rng = np.random.default_rng(seed)
        p = np.zeros(off, dtype=np.float32)
        real_in = {id(self.enc_n): 3 * HIST + 1 + 4, id(self.enc_e): 4}
        for m in self.mlps:
            fan1 = real_in.get(id(m), m.K)
            b1 = math.sqrt(6. / fan1)
            W1 = np.zeros((m.K, LAT), dtype=np.float32)
            W1[:fan1] = rng.uniform(-b1, b1, size=(fan1, LAT))
            scale = residual_scale if m in self.edge or m in self.node else 1.0
            b2 = scale * math.sqrt(6. / LAT) / (1. if m is not self.dec else 4.)
            W2 = np.zeros((LAT, m.NP), dtype=np.float32)
            W2[:, :m.nout] = rng.uniform(-b2, b2, size=(LAT, m.nout))
            p[m.o_W1:m.o_b1] = W1.ravel()
            p[m.o_W2:m.o_b2] = W2.ravel()
        self.flat = torch.from_numpy(p).to(device)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.t = 0
        for m in self.mlps:
            m.alloc(device)
        self.cast()
        self.cap_N = self.cap_E = 0
        self.h = self.e = None
        self.part = self.bpart = None
        self.nchunks = 64

    # -- parameters --------------------------------------------------------
    def cast(self):
        L, s = glib(), stream()
        for m in self.mlps:
            base = self.flat.data_ptr()
            _ok(L.gnn_cast_weights(C.c_void_p(base + 4 * m.o_W1), m.K, LAT, LAT, ptr(m.W1), ptr(m.W1t), s),
                "cast W1")
            _ok(L.gnn_cast_weights(C.c_void_p(base + 4 * m.o_W2), LAT, m.nout, m.NP, ptr(m.W2), ptr(m.W2t),
                                   s), "cast W2")

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.t += 1
        lr_t = lr * math.sqrt(1. - b2 ** self.t) / (1. - b1 ** self.t)
        _ok(glib().gnn_adam(ptr(self.flat), ptr(self.grad), ptr(self.m), ptr(self.v), self.n_params,
                            lr_t, b1, b2, eps, wd, stream()), "adam")
        self.cast()

    def state_dict(self):
        return {"flat": self.flat.detach().cpu().numpy(), "L": self.L}

    def load(self, flat):
        torch = _torch()
        self.flat.copy_(torch.from_numpy(np.asarray(flat, dtype=np.float32)))
        self.cast()

    # -- buffers -----------------------------------------------------------
    def _ensure(self, N, E, train):
        """Buffers for N nodes and E edges, in two groups that grow separately
        with 25 % headroom (the graph's edge count drifts from step to step,
        and the node side of a forward runs before the graph exists); the
        hidden layers and the gradient buffers exist only once the network
        has trained."""
        torch = _torch()
        first_train = train and self.part is None
        grow_n, grow_e = N > self.cap_N or self.h is None, E > self.cap_E or self.e is None
        if not (grow_n or grow_e or first_train):
            return
        train = train or self.part is not None
        bf, f32, dev = torch.bfloat16, torch.float32, self.dev
        L = self.L
        z = lambda *s, dt=bf: torch.zeros(s, dtype=dt, device=dev)
        if grow_n or first_train:
            N = max(N * 5 // 4, self.cap_N) if grow_n else self.cap_N
            self.cap_N = N
            self.h = [z(N, LAT) for _ in range(L + 1)]
            self.agg = [z(N, LAT) for _ in range(L)]
            self.out = z(N, 3, dt=f32)
            self.P = z(N, LAT, dt=f32)
            self.hid_n = [z(N, LAT) for _ in range(L)] if train else [None] * L
            self.hid_enc_n = z(N, LAT) if train else None
            self.hid_dec = z(N, LAT) if train else None
            if train:
                self.d_h = z(N, LAT, dt=f32)
                self.d_agg = z(N, LAT, dt=f32)
                self.dH_n = z(N, LAT)
        if grow_e or first_train:
            E = max(E * 5 // 4, self.cap_E) if grow_e else self.cap_E
            self.cap_E = E
            self.e = [z(E, LAT) for _ in range(L + 1)]
            self.hid_e = [z(E, LAT) for _ in range(L)] if train else [None] * L
            self.hid_enc_e = z(E, LAT) if train else None
            if train:
                self.d_e = z(E, LAT, dt=f32)
                self.d_hi = z(E, LAT, dt=f32)
                self.d_hj = z(E, LAT, dt=f32)
                self.dH_e = z(E, LAT)
        if train and self.part is None:
            self.part = z(self.nchunks * 3 * LAT * LAT, dt=f32)
            self.bpart = z(self.nchunks * LAT, dt=f32)

    # -- forward -----------------------------------------------------------
    def _fwd(self, mlp, inp, res, hid, out, outf, train):
        # the hidden layers are kept for the backward pass only
        _ok(glib().gnn_mlp_forward(C.byref(inp), C.byref(mlp.weights(self.flat)), ptr(res),
                                   ptr(hid) if train else None, ptr(out), ptr(outf), 0, stream()),
            "mlp forward")

    def forward(self, g: Graph, xn, xe, train=False):
        """xn bf16 [N][32], xe bf16 [E][32] -> self.out float [N][3]."""
        self.forward_nodes(g.N, xn, train)
        return self.forward_edges(g, xe, train)

    def forward_nodes(self, N, xn, train=False):
        """The part of the forward that needs no graph: the node encoder and
        layer 0's receiver term P = h W1b (factorised).  It can be queued
        before the graph is built, so that it runs while the host waits for
        the graph's edge count."""
        self._ensure(N, 0, train)
        self._fwd(self.enc_n, io(N, (xn, None)), None, self.hid_enc_n, self.h[0], None, train)
        if self.factorised:
            _ok(glib().gnn_node_p(N, ptr(self.h[0]), C.byref(self.edge[0].weights(self.flat)), ptr(self.P),
                                  stream()), "node P")

    def forward_edges(self, g: Graph, xe, train=False):
        """The rest of the forward, after forward_nodes on the same N."""
        self._ensure(g.N, g.E, train)
        Lb, s, N, E = glib(), stream(), g.N, g.E
        F = self.flat
        fwd = lambda mlp, inp, res, hid, out, outf: self._fwd(mlp, inp, res, hid, out, outf, train)
        # inference fuses the edge encoder into the first edge MLP (e[0] is
        # then never formed); training keeps it for the backward pass
        fuse_enc = E > 0 and self.factorised and not train
        if E > 0 and not fuse_enc:
            fwd(self.enc_e, io(E, (xe, None)), None, self.hid_enc_e, self.e[0], None)
        if E > 0 and self.factorised:
            # per layer: the edge MLP, then the node side in one pass (the
            # aggregation, the node MLP and, without receiver_in_kernel, the
            # next layer's P = h W1b).  Layer 0 reads P from gnn_node_p (its
            # kernel holds the encoder's weights instead of W1b); the later
            # layers form h[recv] W1b in the edge kernel unless told not to.
            for l in range(self.L):
                h, e = self.h[l], self.e[l]
                P = self.P if (l == 0 or not self.receiver_in_kernel) else None
                if l == 0 and fuse_enc:
                    _ok(Lb.gnn_edge_mlp_forward_encoded(E, ptr(xe), C.byref(self.enc_e.weights(F)), ptr(g.recv),
                                                        ptr(g.send), ptr(self.P), ptr(h),
                                                        C.byref(self.edge[0].weights(F)), ptr(self.e[1]), 0, s),
                        "encoded edge MLP forward")
                else:
                    _ok(Lb.gnn_edge_mlp_forward(E, ptr(e), ptr(g.recv), ptr(g.send), ptr(P), ptr(h),
                                                C.byref(self.edge[l].weights(F)),
                                                ptr(self.hid_e[l]) if train else None, ptr(self.e[l + 1]), 0, s),
                        "edge MLP forward")
                nxt = (C.byref(self.edge[l + 1].weights(F))
                       if l + 1 < self.L and not self.receiver_in_kernel else None)
                _ok(Lb.gnn_node_layer_forward(N, ptr(h), ptr(self.e[l + 1]), ptr(g.row),
                                              C.byref(self.node[l].weights(F)),
                                              ptr(self.agg[l]) if train else None,
                                              ptr(self.hid_n[l]) if train else None, ptr(self.h[l + 1]), nxt,
                                              ptr(self.P), s), "node layer forward")
        else:
            for l in range(self.L):
                h, e = self.h[l], self.e[l]
                if E > 0:
                    fwd(self.edge[l], io(E, (e[:E], None), (h, g.recv), (h, g.send)), e, self.hid_e[l],
                        self.e[l + 1], None)
                _ok(Lb.gnn_segment_sum(ptr(self.e[l + 1]), ptr(g.row), N, ptr(self.agg[l]), s), "segment sum")
                fwd(self.node[l], io(N, (h[:N], None), (self.agg[l], None)), h, self.hid_n[l], self.h[l + 1],
                    None)
        fwd(self.dec, io(N, (self.h[self.L][:N], None)), None, self.hid_dec, None, self.out)
        return self.out[:N]

    # -- backward ----------------------------------------------------------
    def _wgrad(self, mlp, inp, hid, dH, dY32, M, accumulate=0):
        Lb, s = glib(), stream()
        base = self.grad.data_ptr()
        # W2, b2: X = hidden, G = dY (float [M][nout])
        _ok(Lb.gnn_weight_grad(C.byref(io(M, (hid, None))), None, ptr(dY32), mlp.nout, mlp.NP,
                               ptr(self.part), ptr(self.bpart), self.nchunks,
                               C.c_void_p(base + 4 * mlp.o_W2), C.c_void_p(base + 4 * mlp.o_b2),
                               accumulate, s), "wgrad W2")
        # W1, b1: X = the MLP input, G = dH (bf16 [M][128])
        _ok(Lb.gnn_weight_grad(C.byref(inp), ptr(dH), None, LAT, LAT, ptr(self.part), ptr(self.bpart),
                               self.nchunks, C.c_void_p(base + 4 * mlp.o_W1),
                               C.c_void_p(base + 4 * mlp.o_b1), accumulate, s), "wgrad W1")

    def backward(self, g: Graph, xn, xe, d_out):
        """d_out float [N][3]: gradient of the loss at the decoder output.
        Fills self.grad (overwritten)."""
        torch = _torch()
        Lb, s, N, E = glib(), stream(), g.N, g.E
        F = self.flat
        bwd = lambda mlp, inp, dY, hid, dH, dX, acc, M: _ok(
            Lb.gnn_mlp_backward(C.byref(inp), C.byref(mlp.weights(F)), ptr(dY), ptr(hid), ptr(dH),
                                (C.c_void_p * 3)(*[None if d is None else d.data_ptr() for d in dX]),
                                acc, 0, s), "mlp backward")
        d_h = self.d_h[:N]
        # decoder
        inp = io(N, (self.h[self.L][:N], None))
        dH = self.dH_n[:N]
        bwd(self.dec, inp, d_out, self.hid_dec, dH, [d_h, None, None], 0, N)
        self._wgrad(self.dec, inp, self.hid_dec, dH, d_out, N)
        d_e = self.d_e[:E]
        d_e.zero_()
        for l in reversed(range(self.L)):
            h, e = self.h[l], self.e[l]
            # node MLP: h_{l+1} = h_l + MLP_n([h_l, agg_l])
            inp = io(N, (h[:N], None), (self.agg[l], None))
            self._wgrad_w2_only(self.node[l], self.hid_n[l], d_h, N)
            # accumulate bit 0 only: the residual source adds into d_h, d_agg is overwritten
            bwd(self.node[l], inp, d_h, self.hid_n[l], self.dH_n[:N], [d_h, self.d_agg[:N], None], 1, N)
            self._wgrad_w1_only(self.node[l], inp, self.dH_n[:N])
            if E > 0:
                # aggregation: d_e(l+1) += d_agg[recv]
                _ok(Lb.gnn_broadcast_add(ptr(self.d_agg), ptr(g.recv), E, ptr(d_e), s), "broadcast")
                # edge MLP: e_{l+1} = e_l + MLP_e([e_l, h_l[recv], h_l[send]])
                inp = io(E, (e[:E], None), (h, g.recv), (h, g.send))
                self._wgrad_w2_only(self.edge[l], self.hid_e[l], d_e, E)
                bwd(self.edge[l], inp, d_e, self.hid_e[l], self.dH_e[:E],
                    [d_e, self.d_hi[:E], self.d_hj[:E]], 1, E)
                self._wgrad_w1_only(self.edge[l], inp, self.dH_e[:E])
                _ok(Lb.gnn_edge_grad_to_nodes(ptr(self.d_hi), ptr(self.d_hj), ptr(g.row), ptr(g.rev), N,
                                              ptr(d_h), s), "edge grads to nodes")
        # encoders (their inputs are data: no input gradient)
        if E > 0:
            inp = io(E, (xe, None))
            self._wgrad_w2_only(self.enc_e, self.hid_enc_e, d_e, E)
            bwd(self.enc_e, inp, d_e, self.hid_enc_e, self.dH_e[:E], [None, None, None], 0, E)
            self._wgrad_w1_only(self.enc_e, inp, self.dH_e[:E])
        else:
            self._zero_grad(self.enc_e)
            for l in range(self.L):
                self._zero_grad(self.edge[l])
        inp = io(N, (xn, None))
        self._wgrad_w2_only(self.enc_n, self.hid_enc_n, d_h, N)
        bwd(self.enc_n, inp, d_h, self.hid_enc_n, self.dH_n[:N], [None, None, None], 0, N)
        self._wgrad_w1_only(self.enc_n, inp, self.dH_n[:N])

    def _zero_grad(self, mlp):
        self.grad[mlp.o_W1:mlp.end].zero_()

    def _wgrad_w2_only(self, mlp, hid, dY32, M):
        base = self.grad.data_ptr()
        _ok(glib().gnn_weight_grad(C.byref(io(M, (hid[:M], None))), None, ptr(dY32), mlp.nout, mlp.NP,
                                   ptr(self.part), ptr(self.bpart), self.nchunks,
                                   C.c_void_p(base + 4 * mlp.o_W2), C.c_void_p(base + 4 * mlp.o_b2), 0,
                                   stream()), "wgrad W2")

    def _wgrad_w1_only(self, mlp, inp, dH):
        base = self.grad.data_ptr()
        _ok(glib().gnn_weight_grad(C.byref(inp), ptr(dH), None, LAT, LAT, ptr(self.part), ptr(self.bpart),
                                   self.nchunks, C.c_void_p(base + 4 * mlp.o_W1),
                                   C.c_void_p(base + 4 * mlp.o_b1), 0, stream()), "wgrad W1")


def mse(pred, target):
    """Mean squared error and its gradient (float [N][3] each), on the device."""
    torch = _torch()
    n = pred.numel()
    d = torch.empty_like(pred)
    part = torch.empty((n + 255) // 256, dtype=torch.float32, device=pred.device)
    _ok(glib().gnn_mse(ptr(pred), ptr(target), n, ptr(d), ptr(part), stream()), "mse")
    return float(part.cpu().numpy().astype(np.float64).sum()) / n, d


# ---------------------------------------------------------------------------
# features, data, training, judged roll-out
# ---------------------------------------------------------------------------

MACRO = 100  # DEM steps per macro-step


@dataclass
class Scales:
    v: float        # velocity scale of the node features
    a: float        # scale of the predicted acceleration
    r_mean: float
    r_std: float
    radius: float   # graph radius


def node_features(N, x_dev, vhist_dev, rad_dev, planes_dev, n_planes, sc: Scales, xn):
    """Node features on the device (bf16 [N][32]); they need no graph."""
    _ok(glib().gnn_node_features(N, ptr(x_dev), ptr(vhist_dev), HIST, ptr(rad_dev), sc.v, sc.r_mean,
                                 sc.r_std, ptr(planes_dev), n_planes, sc.radius, ptr(xn), stream()),
        "node features")


def edge_features(g: Graph, x_dev, sc: Scales, xe):
    """Edge features on the device (bf16 [E][32])."""
    _ok(glib().gnn_edge_features(g.E, ptr(x_dev), ptr(g.recv), ptr(g.send), sc.radius, ptr(xe), stream()),
        "edge features")


def features(g: Graph, x_dev, vhist_dev, rad_dev, planes_dev, n_planes, sc: Scales, xn, xe):
    """Node and edge features on the device (bf16 [N][32], [E][32])."""
    node_features(g.N, x_dev, vhist_dev, rad_dev, planes_dev, n_planes, sc, xn)
    edge_features(g, x_dev, sc, xe)


@dataclass
class Trajectory:
    """Frames every MACRO DEM steps.  Velocities are the macro-step mean
    velocities (x_t - x_{t-1}) / Dt, not the DEM's instantaneous ones: the
    spheres of a packing rattle many times within a macro-step, and the mean
    velocity is what carries a position from frame to frame.  With them the
    surrogate's update v_{t+1} = v_t + (a + g) Dt, x_{t+1} = x_t + v_{t+1} Dt
    reproduces the DEM's positions exactly when a is the target
    (x_{t+1} - 2 x_t + x_{t-1}) / Dt^2 - g."""
    x: np.ndarray        # [T][N][3] float64 (host, for the graph build)
    x_dev: object        # [T][N][3] float64 (device)
    v_dev: object        # [T][N][3] float32 (device): (x_t - x_{t-1}) / Dt
    a_host: np.ndarray   # [T-1][N][3] float32: (v_{t+1} - v_t)/Dt - g
    rad: np.ndarray
    wl: object


def generate(wl, n_frames, log=None):
    """Runs the DEM from the workload's state and keeps every MACRO-th step:
    positions, and mean velocities over each macro-step."""
    torch = _torch()
    sysm = wl.system()
    st = abi.ParticleState(wl.state.n, max_cnts=wl.max_contacts, max_walls=6)
    xs = np.empty((n_frames, wl.state.n, 3))
    vs = np.empty((n_frames, wl.state.n, 3), dtype=np.float32)
    dt = wl.dt * MACRO
    x_prev = wl.state.x.copy()
    t0 = time.perf_counter()
    for f in range(n_frames):
        sysm.step(MACRO)
        sysm.download(st, fields=("x",))
        xs[f] = st.x
        vs[f] = (st.x - x_prev) / dt
        x_prev = xs[f]
    sysm.close()
    if log:
        log(f"[surrogate] {n_frames} frames of {wl.state.n} spheres in {time.perf_counter() - t0:.1f} s")
    g = np.asarray(wl.gravity, dtype=np.float32)
    a = (vs[1:] - vs[:-1]) / np.float32(dt) - g
    return Trajectory(xs, torch.from_numpy(xs).cuda(), torch.from_numpy(vs).cuda(), a.astype(np.float32),
                      wl.state.rad.copy(), wl)


def fit_scales(trajs, radius):
    v = np.concatenate([t.v_dev[::10].cpu().numpy().ravel() for t in trajs])
    a = np.concatenate([t.a_host[::10].ravel() for t in trajs])
    r = np.concatenate([t.rad for t in trajs])
    return Scales(v=float(max(v.std(), 1e-3)), a=float(max(a.std(), 1e-3)), r_mean=float(r.mean()),
                  r_std=float(max(r.std(), 1e-6)), radius=radius)


class Trainer:
    """noise: training inputs see a velocity history perturbed by a random walk
    whose last step has standard deviation noise * a_scale * Dt (in velocity),
    and the target is corrected so that the predicted step still lands on the
    DEM's next velocity: a' = a + (v_t - v'_t) / Dt.  A roll-out feeds the
    network its own, slightly wrong, velocities; training on such inputs
    teaches it to steer back instead of drifting (noise injection, as in
    learned particle simulators).  Evaluation never adds noise."""

    def __init__(self, net: Network, trajs, scales: Scales, seed=0, noise=0.):
        torch = _torch()
        self.net, self.trajs, self.sc = net, trajs, scales
        self.noise = float(noise)
        self.gen = torch.Generator(device="cuda")
        self.gen.manual_seed(int(seed) + 12345)
        self.rng = np.random.default_rng(seed)
        self.builders = []
        self.planes = []
        for t in trajs:
            wl = t.wl
            self.builders.append(GraphBuilder(wl.state.n, scales.radius, wl.lo, wl.hi))
            self.planes.append(torch.tensor(np.asarray(wl.extra["planes"], dtype=np.float32).ravel(),
                                            device="cuda"))
        self.rad = [torch.from_numpy(t.rad).cuda() for t in trajs]
        self.targets = [torch.from_numpy(t.a_host / np.float32(scales.a)).cuda() for t in trajs]

    def sample(self, k, t, noise=0.):
        torch = _torch()
        tr = self.trajs[k]
        g = self.builders[k].build(tr.x_dev[t])
        vh = tr.v_dev[t - HIST + 1:t + 1].contiguous()
        target = self.targets[k][t]
        if noise > 0.:
            # random walk over the history, newest step last
            Dt = tr.wl.dt * MACRO
            sd = noise * self.sc.a * Dt / math.sqrt(HIST)
            walk = torch.cumsum(torch.randn(vh.shape, device="cuda", generator=self.gen) * sd, dim=0)
            vh = (vh + walk).contiguous()
            target = (target - walk[-1] / np.float32(Dt * self.sc.a)).contiguous()
        xn = torch.empty((g.N, FEAT), dtype=torch.bfloat16, device="cuda")
        xe = torch.empty((max(g.E, 1), FEAT), dtype=torch.bfloat16, device="cuda")
        features(g, tr.x_dev[t], vh, self.rad[k], self.planes[k], len(tr.wl.extra["planes"]), self.sc, xn, xe)
        return g, xn, xe, target

    def step(self, lr, k=None, t=None):
        if k is None:
            k = int(self.rng.integers(len(self.trajs)))
        T = self.trajs[k].x.shape[0]
        if t is None:
            t = int(self.rng.integers(HIST - 1, T - 1))
        g, xn, xe, target = self.sample(k, t, self.noise)
        out = self.net.forward(g, xn, xe[:g.E], train=True)
        loss, d = mse(out, target)
        self.net.backward(g, xn, xe[:g.E], d)
        self.net.adam(lr)
        return loss

    def evaluate(self, frames):
        """Mean one-macro-step MSE of the network and of the no-network
        baseline (a = gravity only) over (trajectory, frame) pairs."""
        torch = _torch()
        ln = lb = 0.
        for k, t in frames:
            g, xn, xe, target = self.sample(k, t)
            out = self.net.forward(g, xn, xe[:g.E])
            ln += mse(out, target)[0]
            lb += mse(torch.zeros_like(target), target)[0]
        return ln / len(frames), lb / len(frames)


def horizon_errors(net, traj, sc: Scales, starts, horizons, no_network=False):
    """Free roll-out against the DEM's own trajectory: from each recorded
    frame f in `starts` (its positions and the HIST velocities up to it),
    the surrogate alone -- no judge, no fallback -- advances max(horizons)
    macro-steps, and at each horizon h the mean |x_surrogate - x_DEM| / r_mean
    over all spheres is taken against the DEM's frame f + h.  no_network:
    the method without the network (a = g over every macro-step).  Returns
    {h: mean over the starts}.  Unlike the judge's one-macro-step check this
    error accumulates, so a tolerance on it can be missed."""
    torch = _torch()
    H = max(horizons)
    T = traj.x.shape[0]
    for f in starts:
        if f - HIST + 1 < 0 or f + H >= T:
            raise ValueError(f"start frame {f} with horizon {H} lies outside the {T}-frame trajectory")
    ro = Rollout(net, traj.wl, sc, no_network=no_network)
    ro.dem.close()
    out = {h: [] for h in horizons}
    for f in starts:
        ro.x = traj.x_dev[f].clone()
        ro.vhist = traj.v_dev[f - HIST + 1:f + 1].contiguous().clone()
        for h in range(1, H + 1):
            # a roll-out that has left every finite position (an untrained
            # network can) is off by infinity from here on
            ok = bool(torch.isfinite(ro.x).all())
            if ok:
                try:
                    ro.surrogate_step()
                    ok = bool(torch.isfinite(ro.x).all())
                except RuntimeError:
                    ok = False
            if not ok:
                for hh in horizons:
                    if hh >= h:
                        out[hh].append(float("inf"))
                break
            if h in out:
                err = torch.linalg.vector_norm(ro.x - traj.x_dev[f + h], dim=1).mean()
                out[h].append(float(err) / ro.r_mean)
    return {h: float(np.mean(v)) for h, v in out.items()}


class Rollout:
    """The surrogate advancing a DEM workload by macro-steps, with the FP64
    DEM as judge and fallback: every `check_every` macro-steps the DEM
    re-simulates the same macro-step from the same state; the mean position
    error |x_surrogate - x_DEM| / r_mean and the kinetic and mechanical
    (kinetic + gravitational) energy errors are recorded (kinetic energy of
    the macro-step mean velocities on both sides) and, above `tol`
    in position, the run continues from the DEM state.  tune_dem: the judge's
    DEM context runs the autotuner's configuration (rebuild cadence, launch
    shape), picked from the warm-up state.  no_network: the method without
    the network (a = g over every macro-step), judged the same way."""

    def __init__(self, net: Network, wl, scales: Scales, check_every=10, tol=0.02, tune_dem=False,
                 no_network=False):
        torch = _torch()
        self.net, self.wl, self.sc = net, wl, scales
        self.check_every, self.tol = check_every, tol
        self.tune_dem, self.dem_tune = tune_dem, None
        self.no_network = no_network
        n = wl.state.n
        self.n = n
        self.dem = wl.system()
        self.builder = GraphBuilder(n, scales.radius, wl.lo, wl.hi)
        self.planes = torch.tensor(np.asarray(wl.extra["planes"], dtype=np.float32).ravel(), device="cuda")
        self.n_planes = len(wl.extra["planes"])
        self.rad = torch.from_numpy(wl.state.rad.copy()).cuda()
        self.r_mean = float(wl.state.rad.mean())
        self.g_dev = torch.tensor(np.asarray(wl.gravity, dtype=np.float64), device="cuda")
        self.Dt = wl.dt * MACRO
        self.xn = torch.empty((n, FEAT), dtype=torch.bfloat16, device="cuda")
        self.xe = None
        # the DEM's positions and velocities on the device (reference order)
        self.xd = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        self.ud = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        self.m_dev = torch.from_numpy(wl.state.m.copy()).cuda()

    def energies(self, x, v):
        """Translational kinetic energy and gravitational potential energy
        (-sum m g.x, zero at the origin) of positions x and velocities v."""
        ke = 0.5 * (self.m_dev * (v.double() ** 2).sum(1)).sum()
        pe = -(self.m_dev * (x @ self.g_dev)).sum()
        return float(ke), float(pe)

    def _dem_pull(self):
        """The DEM context's positions and velocities into self.xd / self.ud
        (device, reference order), synchronised."""
        self.dem.get_kinematics_device(self.xd.data_ptr(), self.ud.data_ptr())
        self.dem.sync()

    def _dem_macro(self, x_before):
        """One macro-step of the DEM context; its positions in self.xd and its
        mean velocity over the macro-step, (x - x_before) / Dt (device)."""
        self.dem.step(MACRO)
        self._dem_pull()
        return (self.xd - x_before) / self.Dt

    def warm_up(self):
        """HIST macro-steps of DEM to fill the velocity history."""
        torch = _torch()
        if self.tune_dem and self.dem_tune is None:
            self.dem_tune = self.dem.autotune(trial_steps=80)  # restores the state
        vh = []
        self._dem_pull()
        x_prev = self.xd.clone()
        for _ in range(HIST):
            vh.append(self._dem_macro(x_prev).float())
            x_prev = self.xd.clone()
        self.x = self.xd.clone()
        self.vhist = torch.stack(vh).contiguous()
        # one untimed check, so that first-call costs (kernel loads) stay out
        # of the timed checks; the check context is reset at every check
        xd, vd = self.dem_step_from(self.x, self.vhist[-1].double())
        float(torch.linalg.vector_norm(self.x - xd, dim=1).mean())
        self.energies(xd, vd)
        # and one untimed surrogate step (buffer allocation), undone
        x_keep, v_keep = self.x.clone(), self.vhist.clone()
        self.surrogate_step()
        self.x.copy_(x_keep)
        self.vhist.copy_(v_keep)

    def surrogate_step(self):
        torch = _torch()
        if self.no_network:
            out = torch.zeros((self.n, 3), dtype=torch.float32, device="cuda")
            _ok(glib().gnn_integrate(self.n, ptr(out), self.sc.a, ptr(self.g_dev), self.Dt, ptr(self.x),
                                     ptr(self.vhist), HIST, stream()), "integrate")
            return
        # the node side first: it needs no graph, so the GPU runs it while the
        # host waits for the graph's edge count
        node_features(self.n, self.x, self.vhist, self.rad, self.planes, self.n_planes, self.sc, self.xn)
        self.net.forward_nodes(self.n, self.xn)
        g = self.builder.build(self.x)
        if self.xe is None or self.xe.shape[0] < max(g.E, 1):
            self.xe = torch.empty((max(g.E, 1) * 5 // 4, FEAT), dtype=torch.bfloat16, device="cuda")
        edge_features(g, self.x, self.sc, self.xe)
        out = self.net.forward_edges(g, self.xe[:g.E])
        _ok(glib().gnn_integrate(self.n, ptr(out), self.sc.a, ptr(self.g_dev), self.Dt, ptr(self.x),
                                 ptr(self.vhist), HIST, stream()), "integrate")

    def dem_step_from(self, x, v):
        """One macro-step of DEM from positions x and velocities v (device,
        double [n][3]) with spins, forces and contact histories cleared;
        returns the DEM's positions (device, valid until the next DEM call)
        and its mean velocity over the macro-step."""
        _torch().cuda.current_stream().synchronize()
        self.dem.set_kinematics_device(x.data_ptr(), v.data_ptr(), True)
        v_mean = self._dem_macro(x)
        return self.xd, v_mean

    def dem_advance(self, x, vhist, n_macro):
        """n_macro macro-steps of DEM from positions x and the last velocity
        of vhist (device); returns the positions and the updated velocity
        history (device copies)."""
        torch = _torch()
        torch.cuda.current_stream().synchronize()
        v0 = vhist[-1].double().contiguous()
        torch.cuda.current_stream().synchronize()
        self.dem.set_kinematics_device(x.data_ptr(), v0.data_ptr(), True)
        vh = list(vhist)
        x_prev = x.clone()
        for _ in range(n_macro):
            v_mean = self._dem_macro(x_prev)
            vh = vh[1:] + [v_mean.float()]
            x_prev = self.xd.clone()
        return self.xd.clone(), torch.stack(vh).contiguous()

    def run(self, n_macro):
        torch = _torch()
        self.warm_up()
        errs, fallbacks, dem_failures = [], 0, 0
        ke_errs, e_errs = [], []
        t_sur = t_dem = 0.
        n_sur = n_dem = 0
        # the state the run last continued from after a check (host copies)
        good = (self.x.clone(), self.vhist.clone(), 0)
        for m in range(n_macro):
            if m % self.check_every == 0:
                x0 = self.x.clone()
                v0 = self.vhist[-1].double()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self.surrogate_step()
                torch.cuda.synchronize()
                t_sur += time.perf_counter() - t0
                n_sur += 1
                t0 = time.perf_counter()
                try:
                    xd, vd = self.dem_step_from(x0, v0)
                    err = float(torch.linalg.vector_norm(self.x - xd, dim=1).mean()) / self.r_mean
                except abi.DemError:
                    # the surrogate's state is not one the DEM can continue
                    # from (e.g. deep overlaps): a failed check
                    xd = vd = None
                    err = float("inf")
                    dem_failures += 1
                t_dem += time.perf_counter() - t0
                n_dem += 1
                if xd is not None:  # energies, outside the timed check
                    ke_s, pe_s = self.energies(self.x, self.vhist[-1])
                    ke_d, pe_d = self.energies(xd, vd)
                    if ke_d > 0.:
                        ke_errs.append(abs(ke_s - ke_d) / ke_d)
                    e_errs.append(abs((ke_s + pe_s) - (ke_d + pe_d)) / max(abs(ke_d + pe_d), 1e-300))
                errs.append(err)
                if err > self.tol:
                    fallbacks += 1
                    if xd is not None:
                        self.x.copy_(xd)
                        self.vhist[-1].copy_(vd)
                    else:
                        # the DEM redoes the window since the last good state
                        xg, vg, mg = good
                        xd, vh = self.dem_advance(xg, vg, m + 1 - mg)
                        self.x.copy_(xd)
                        self.vhist.copy_(vh)
                good = (self.x.clone(), self.vhist.clone(), m + 1)
            else:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self.surrogate_step()
                torch.cuda.synchronize()
                t_sur += time.perf_counter() - t0
                n_sur += 1
        fin = [e for e in errs if np.isfinite(e)]
        return {"macro_steps": n_macro, "checks": len(errs), "fallbacks": fallbacks,
                "dem_check_failures": dem_failures,
                "mean_position_error_over_r": float(np.mean(fin)) if fin else None,
                "max_position_error_over_r": float(np.max(fin)) if fin else None,
                "mean_kinetic_energy_error_rel": float(np.mean(ke_errs)) if ke_errs else None,
                "max_kinetic_energy_error_rel": float(np.max(ke_errs)) if ke_errs else None,
                "mean_mechanical_energy_error_rel": float(np.mean(e_errs)) if e_errs else None,
                "surrogate_ms_per_macro": 1e3 * t_sur / max(n_sur, 1),
                "dem_ms_per_macro": 1e3 * t_dem / max(n_dem, 1)}
