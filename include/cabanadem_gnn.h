/*
 * cabanadem_gnn.h -- C ABI of the learned-surrogate kernels
 * (libcabanadem_gnn.so, built from cabanadem_amd/csrc/gnn.hip).
 *
 * No counterpart in the reference: the surrogate is an addition on top of
 * the DEM path.  A message-passing network over the DEM neighbour graph
 * predicts per-particle accelerations; these entry points are its building
 * blocks (fused two-layer MLPs on the BF16 MFMA units with FP32 accumulation,
 * their backward passes, weight gradients, the graph aggregation, Adam).
 * The host side that strings them into a network, trains it and runs it
 * against the DEM is cabanadem_amd/surrogate.py.
 *
 * Conventions: all pointers are device pointers; BF16 arrays are bf16
 * (2 bytes, round-to-nearest-even from FP32), FP32 arrays are float.
 * Feature widths are multiples of 32 (pad with zeros); the latent width is
 * 128.  Calls enqueue on `stream` (a hipStream_t, NULL = default stream) and
 * return 0, -1 on a bad argument, -2 on a launch failure.
 */
#ifndef CABANADEM_GNN_H
#define CABANADEM_GNN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GNN_ABI_VERSION 5
#define GNN_LATENT 128

/* Input rows of an MLP: the concatenation of up to three BF16 sources;
 * source q contributes width[q] features of row idx[q][m] (or row m when
 * idx[q] is NULL). */
typedef struct gnn_mlp_io {
  int M;               /* rows */
  int nseg;            /* sources, 1..3 */
  const void* src[3];  /* bf16 [rows_q][width_q] */
  const int* idx[3];   /* [M] row of source q for MLP row m, or NULL */
  int width[3];        /* multiples of 32 */
} gnn_mlp_io;

/* y = W2 relu(W1 x + b1) + b2, x of width K, hidden 128, output nout
 * (padded to nout_pad = 16 or 128 in the BF16 copies).  W1t/W2t are the
 * transposed copies used by the forward pass, W1/W2 the row-major copies
 * used by the backward pass. */
typedef struct gnn_mlp_weights {
  int K, nout, nout_pad;
  const void* W1t;     /* bf16 [128][K] */
  const void* W1;      /* bf16 [K][128] */
  const float* b1;     /* [128] */
  const void* W2t;     /* bf16 [nout_pad][128] */
  const void* W2;      /* bf16 [128][nout_pad] */
  const float* b2;     /* [nout_pad] */
} gnn_mlp_weights;

int gnn_abi_version(void);

/* out = residual + MLP(x); hidden (bf16 [M][128], optional) keeps relu(W1 x + b1)
 * for the backward pass; out (bf16 [M][nout]) and/or out_f32 (float [M][nout]).
 * Ungathered inputs of width 32, 128 or 128 + 128 with a 128-wide BF16
 * output (and no out_f32) run on a persistent kernel that keeps the weights
 * in LDS; everything else, or grid_cap > 0, on the row-tiled kernel
 * (grid_cap then limits its workgroups; tiles of 128 rows). */
int gnn_mlp_forward(const gnn_mlp_io* io, const gnn_mlp_weights* w, const void* residual,
                    void* hidden, void* out, float* out_f32, int grid_cap, void* stream);
/* dH (bf16 [M][128]) = (dY W2^T) masked by hidden > 0; dX[q] (float
 * [M][width_q], NULL to skip) = (dH W1^T) restricted to source q, stored,
 * or added when bit q of accumulate is set. */
int gnn_mlp_backward(const gnn_mlp_io* io, const gnn_mlp_weights* w, const float* dY,
                     const void* hidden, void* dH, float* const dX[3], int accumulate,
                     int grid_cap, void* stream);
/* dW (float [K][ncols_pad]) = X^T G and db (float [ncols_pad]) = column sums
 * of G, with G bf16 or float [M][ncols]; part/bpart are scratch of
 * nchunks*K*ncols_pad and nchunks*ncols_pad floats (the rows are split into
 * nchunks chunks reduced in chunk order, so results are deterministic). */
int gnn_weight_grad(const gnn_mlp_io* io, const void* G_bf16, const float* G_f32, int ncols,
                    int ncols_pad, float* part, float* bpart, int nchunks, float* dW, float* db,
                    int accumulate, void* stream);
/* agg[i] (bf16 [N][128]) = sum of x[e] (bf16 [E][128]) over the edges
 * e in [row[i], row[i+1]), summed in FP32. */
int gnn_segment_sum(const void* x, const int* row, int N, void* agg, void* stream);
/* dh[i] += sum over the edges e of i of d_recv[e] + d_send[rev[e]] (float
 * [E][128] each, either may be NULL). */
int gnn_edge_grad_to_nodes(const float* d_recv, const float* d_send, const int* row,
                           const int* rev, int N, float* dh, void* stream);
/* d_e[e] += d_agg[recv[e]] (float [E][128], float [N][128]). */
int gnn_broadcast_add(const float* d_agg, const int* recv, long E, float* d_e, void* stream);
/* Adam with decoupled bias correction folded into lr_t; wd = L2 factor. */
int gnn_adam(float* p, const float* g, float* m, float* v, long n, float lr_t, float b1,
             float b2, float eps, float wd, void* stream);
/* FP32 master [K][NP] (columns >= N are zero) -> bf16 W [K][NP] and Wt [NP][K]. */
int gnn_cast_weights(const float* p, int K, int N, int NP, void* W, void* Wt, void* stream);
/* Node features (bf16 [N][32]): the H most recent velocities vhist
 * (float [H][N][3], oldest first) / v_scale, (r - r_mean) / r_scale, and
 * for each plane (float [n_planes][4] = unit normal, offset d) the distance
 * n . x - d / dist_scale clipped to [0, 1]; zero padding.  3H + 1 + n_planes <= 32. */
int gnn_node_features(int N, const double* x, const float* vhist, int H, const double* rad,
                      float v_scale, float r_mean, float r_scale, const float* planes, int n_planes,
                      float dist_scale, void* out, void* stream);
/* Edge features (bf16 [E][32]) of edge send[e] -> recv[e]: (x_i - x_j) / radius
 * and |x_i - x_j| / radius; zero padding. */
int gnn_edge_features(long E, const double* x, const int* recv, const int* send, double radius,
                      void* out, void* stream);
/* Surrogate macro-step: v = v_last + (a a_scale + g) dt, x += v dt; the
 * velocity history shifts by one.  a: float [N][3] (network output),
 * gravity: double[3] on the device. */
int gnn_integrate(int N, const float* a, float a_scale, const double* gravity, double dt, double* x,
                  float* vhist, int H, void* stream);
/* dY = 2 (y - t) / count, part[b] = partial sums of (y - t)^2 per block of 256. */
int gnn_mse(const float* y, const float* t, long count, float* dY, float* part, void* stream);

/* The edge MLP of a message-passing layer, inputs (e, h[recv], h[send]),
 * with the receiver's share of its first layer precomputed per node:
 * W1 = [W1a; W1b; W1c] (K = 384, nout 128) splits into
 * e W1a + (h W1b)[recv] + h[send] W1c.
 * gnn_node_p: P (float [N][128]) = h W1b for the node latents h (bf16 [N][128]).
 * gnn_edge_mlp_forward: out[e] = e[e] + W2 relu(e[e] W1a + P[recv[e]]
 * + h[send[e]] W1c + b1) + b2 (bf16 [E][128]); hidden (bf16 [E][128],
 * optional) as in gnn_mlp_forward.  P may be null: the kernel then forms
 * h[recv[e]] W1b itself from the gathered row, and P need not exist.  The
 * same function as gnn_mlp_forward over the three sources, up to FP32
 * summation order.  grid_cap > 0 limits the persistent workgroups (default
 * one per CU). */
int gnn_node_p(int N, const void* h, const gnn_mlp_weights* w, float* P, void* stream);
int gnn_edge_mlp_forward(long E, const void* e, const int* recv, const int* send, const float* P,
                         const void* h, const gnn_mlp_weights* w, void* hidden, void* out,
                         int grid_cap, void* stream);
/* gnn_edge_mlp_forward_encoded: gnn_edge_mlp_forward with e = the edge
 * encoder enc (K = 32, 128 outputs, no residual) applied to xe (bf16
 * [E][32]) inside the same kernel, so the encoded rows are never stored:
 * bitwise the result of gnn_mlp_forward(xe, enc) followed by
 * gnn_edge_mlp_forward.  Inference only (no hidden rows). */
int gnn_edge_mlp_forward_encoded(long E, const void* xe, const gnn_mlp_weights* enc, const int* recv,
                                 const int* send, const float* P, const void* h,
                                 const gnn_mlp_weights* w, void* out, int grid_cap, void* stream);

/* The node side of a message-passing layer in one pass:
 *   agg (bf16 [N][128], optional output) = the sum of e (bf16 [E][128]) over
 *     each node's edges [row[i], row[i+1]), as gnn_segment_sum;
 *   out = h + MLP([h, agg]) with w (K = 256, nout 128); hidden (optional) as
 *     in gnn_mlp_forward;
 *   when next_edge (the next layer's edge MLP, K = 384) is given,
 *     P (float [N][128]) = out W1b of it, as gnn_node_p. */
int gnn_node_layer_forward(int N, const void* h, const void* e, const int* row,
                           const gnn_mlp_weights* w, void* agg, void* hidden, void* out,
                           const gnn_mlp_weights* next_edge, float* P, void* stream);

/* Radius graph on the device: the edges j -> i for every pair i != j of the
 * N points x (double [N][3]) with |x_i - x_j|^2 <= radius^2 (the DEM
 * neighbour criterion), in the receiver-major layout above with each node's
 * senders ascending.  Points are binned into cells of side > radius over the
 * box [lo, hi] (host double[3] each; points outside are clamped into the
 * edge cells, which keeps the result exact and only costs time).
 *
 * gnn_radius_graph_workspace: bytes of device workspace for N points, or -1
 * for a degenerate box or one needing more than 2^26 cells.
 * gnn_radius_graph_count: bins the points and fills row (int [N+1]); writes
 * the edge count E to *n_edges (host) after synchronising `stream`; -3 when
 * E does not fit an int.
 * gnn_radius_graph_fill: with the same arguments and the workspace as left by
 * the count, fills send, recv and rev (int [E] each; rev[e] is the edge
 * recv[e] -> send[e]).  n_edges is the count's E (the capacity of the three
 * arrays).
 * gnn_radius_graph_build: the count and, when E <= capacity (the length of
 * send, recv and rev), the fill in one call, so no host round trip sits
 * between them; returns 1, with *n_edges set and the fill not run, when E
 * exceeds capacity (grow the arrays, then call gnn_radius_graph_fill). */
long gnn_radius_graph_workspace(int N, const double lo[3], const double hi[3], double radius);
int gnn_radius_graph_count(int N, const double* x, double radius, const double lo[3],
                           const double hi[3], void* work, long work_bytes, int* row,
                           long* n_edges, void* stream);
int gnn_radius_graph_fill(int N, const double* x, double radius, const double lo[3],
                          const double hi[3], const void* work, long work_bytes, const int* row,
                          long n_edges, int* send, int* recv, int* rev, void* stream);
int gnn_radius_graph_build(int N, const double* x, double radius, const double lo[3],
                           const double hi[3], void* work, long work_bytes, int* row,
                           long capacity, int* send, int* recv, int* rev, long* n_edges,
                           void* stream);

/* gnn_radius_graph_build_static: gnn_radius_graph_build with the workspace
 * in a device array of the library's own (GNN_GRAPH_SCRATCH_BYTES on each
 * GPU, held from the library's load until the process ends) instead of a
 * caller's buffer, so the call allocates nothing: no device memory, no
 * pinned host word.  No counterpart in the reference; it is for callers
 * that cannot allocate (serving loops, harnesses that forbid allocation).
 * The caller still passes x, row and the edge arrays.  Same results and
 * return codes as gnn_radius_graph_build, and -4 when the graph's workspace
 * (gnn_radius_graph_workspace) exceeds GNN_GRAPH_SCRATCH_BYTES, before any
 * work is queued.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other host threads or streams) race. */
#define GNN_GRAPH_SCRATCH_BYTES (32L * 1024L * 1024L)
int gnn_radius_graph_build_static(int N, const double* x, double radius, const double lo[3],
                                  const double hi[3], int* row, long capacity, int* send,
                                  int* recv, int* rev, long* n_edges, void* stream);

#ifdef __cplusplus
}
#endif
#endif
