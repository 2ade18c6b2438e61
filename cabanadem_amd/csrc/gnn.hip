// gnn.hip -- kernels of the learned surrogate: a message-passing network
// over the DEM neighbour graph (no counterpart in the reference).
//
// Graph: receiver-major edge list (CSR): edges of node i are
// [row[i], row[i+1]), edge e goes j = send[e] -> i = recv[e]; rev[e] is the
// edge i -> j.  Every MLP is two linear layers, y = R + W2 relu(W1 x + b1) + b2,
// with x the concatenation of up to three BF16 row sources (each row either
// the MLP row itself or gathered through an index array).  Weights and
// activations are BF16, every product accumulates in FP32 on the MFMA units
// (v_mfma_f32_16x16x32_bf16), gradients are FP32.
//
// MFMA 16x16x32 BF16 operand layout (lane l of the 64-lane wave):
//   A (16 x 32):  a[0..7] = A[l % 16][8 (l / 16) + 0..7]
//   B (32 x 16):  b[0..7] = B[8 (l / 16) + 0..7][l % 16]
//   D (16 x 16):  d[r]    = D[4 (l / 16) + r][l % 16],  r = 0..3
// Weights are kept in both layouts: Wt [nout][K] (forward: B columns are
// rows of Wt, 8 contiguous k) and W [K][nout] (backward data: 8 contiguous
// output columns).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>

#include "../../include/cabanadem_gnn.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256; // 4 waves
constexpr int kWaveRows = 32; // rows per wave in the row-tiled kernels
constexpr int kTileRows = 4 * kWaveRows;
constexpr int kLat = 128;     // latent width

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline bf16x8 ld8f(const float* p) {
  const float4 u = *reinterpret_cast<const float4*>(p);
  const float4 v = *reinterpret_cast<const float4*>(p + 4);
  bf16x8 r;
  r[0] = (bf16)u.x; r[1] = (bf16)u.y; r[2] = (bf16)u.z; r[3] = (bf16)u.w;
  r[4] = (bf16)v.x; r[5] = (bf16)v.y; r[6] = (bf16)v.z; r[7] = (bf16)v.w;
  return r;
}

constexpr int kWPitch = kLat + 8; // LDS row pitch (bf16) of the weight and hidden tiles

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Rows e0 .. e0 + R - 1 (those below M) of a wave's LDS tile [R][kWPitch]
// to dst [M][128] BF16, 16 bytes per lane: 4 whole rows per instruction.
template <int R>
__device__ inline void wave_store_tile(const bf16 (*tile)[kWPitch], bf16* dst, int e0, int M, int lane) {
#pragma unroll
  for (int it = 0; it < R / 4; ++it) {
    const int r = it * 4 + (lane >> 4), c = (lane & 15) * 8;
    if (e0 + r < M)
      *reinterpret_cast<bf16x8*>(dst + (long)(e0 + r) * kLat + c) = *reinterpret_cast<const bf16x8*>(&tile[r][c]);
  }
}

__device__ inline void wave_lds_sync() {
  // a wave's own LDS tile: its writes before its later reads, and back
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Segs {
  int nseg;
  const bf16* src[3];
  const int* idx[3];
  int width[3];
  int start[4]; // prefix of widths
};

__host__ Segs make_segs(const gnn_mlp_io* io) {
  Segs s{};
  s.nseg = io->nseg;
  s.start[0] = 0;
  for (int k = 0; k < 3; ++k) {
    s.src[k] = k < io->nseg ? reinterpret_cast<const bf16*>(io->src[k]) : nullptr;
    s.idx[k] = k < io->nseg ? io->idx[k] : nullptr;
    s.width[k] = k < io->nseg ? io->width[k] : 0;
    s.start[k + 1] = s.start[k] + s.width[k];
  }
  return s;
}

// 8 consecutive input features k..k+7 of row m (k a multiple of 8; widths
// are multiples of 32, so the 8 never straddle two sources).
__device__ inline bf16x8 seg_load(const Segs& s, int m, int k) {
  int q = 0;
  if (k >= s.start[1]) q = (k >= s.start[2]) ? 2 : 1;
  const int row = s.idx[q] ? s.idx[q][m] : m;
  return ld8(s.src[q] + (long)row * s.width[q] + (k - s.start[q]));
}

struct FwdArgs {
  int M, K, NO, nout;
  Segs in;
  const bf16* W1t;   // [128][K]
  const float* b1;   // [128]
  const bf16* W2t;   // [NO][128]
  const float* b2;   // [NO]
  const bf16* res;   // [M][nout] or null
  bf16* hid;         // [M][128] post-ReLU hidden, or null
  bf16* out;         // [M][nout] or null
  float* outf;       // [M][nout] or null
};

// y = res + W2 relu(W1 x + b1) + b2 for a tile of 128 rows (32 per wave).
template <int NB2> // output column blocks of 16: NO = 16 NB2
__global__ __launch_bounds__(kThreads) void k_mlp_fwd(FwdArgs a) {
  __shared__ bf16 sh[4][kWaveRows][kLat + 8]; // hidden tile per wave (+8: bank skew)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int ntiles = (a.M + kTileRows - 1) / kTileRows;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int r0 = t * kTileRows + w * kWaveRows;
    f32x4 acc[2][8];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    int rows[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const int r = r0 + rb * 16 + lr;
      rows[rb] = r < a.M ? r : a.M - 1; // clamped loads; stores are guarded
    }
    // source by source; each gathered row's address is formed once per tile,
    // so the k loop issues independent loads only
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q >= a.in.nseg) break;
      const int wq = a.in.width[q];
      const bf16* src = a.in.src[q];
      const int* idx = a.in.idx[q];
      const bf16* p0 = src + (long)(idx ? idx[rows[0]] : rows[0]) * wq + lk;
      const bf16* p1 = src + (long)(idx ? idx[rows[1]] : rows[1]) * wq + lk;
      const bf16* wt = a.W1t + (long)lr * a.K + a.in.start[q] + lk;
      for (int kk = 0; kk < wq; kk += 32) {
        const bf16x8 av0 = ld8(p0 + kk), av1 = ld8(p1 + kk);
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
          const bf16x8 bv = ld8(wt + (long)cb * 16 * a.K + kk);
          acc[0][cb] = mfma(av0, bv, acc[0][cb]);
          acc[1][cb] = mfma(av1, bv, acc[1][cb]);
        }
      }
    }
    // hidden = relu(acc + b1) -> LDS (and global when saved for backward)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        const int col = cb * 16 + lr;
        const float bias = a.b1[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lrow = rb * 16 + 4 * (lane >> 4) + r;
          float v = acc[rb][cb][r] + bias;
          v = v > 0.f ? v : 0.f;
          sh[w][lrow][col] = (bf16)v;
        }
      }
    __syncthreads();
    if (a.hid) wave_store_tile<kWaveRows>(sh[w], a.hid, r0, a.M, lane);
    f32x4 acc2[2][NB2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < NB2; ++cb) acc2[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < kLat; k0 += 32) {
      bf16x8 av[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) av[rb] = *reinterpret_cast<const bf16x8*>(&sh[w][rb * 16 + lr][k0 + lk]);
#pragma unroll
      for (int cb = 0; cb < NB2; ++cb) {
        const bf16x8 bv = ld8(a.W2t + (long)(cb * 16 + lr) * kLat + k0 + lk);
        acc2[0][cb] = mfma(av[0], bv, acc2[0][cb]);
        acc2[1][cb] = mfma(av[1], bv, acc2[1][cb]);
      }
    }
    if (NB2 == 8 && a.nout == kLat && !a.outf) {
      // 128-wide BF16 output through the wave's (now free) hidden tile, then
      // whole rows out
      wave_lds_sync();
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < NB2; ++cb) {
          const int col = cb * 16 + lr;
          const float bias = a.b2[col];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int lrow = rb * 16 + 4 * (lane >> 4) + r;
            const int srow = r0 + lrow < a.M ? r0 + lrow : a.M - 1;
            float v = acc2[rb][cb][r] + bias;
            if (a.res) v += (float)a.res[(long)srow * kLat + col];
            sh[w][lrow][col] = (bf16)v;
          }
        }
      wave_lds_sync();
      if (a.out) wave_store_tile<kWaveRows>(sh[w], a.out, r0, a.M, lane);
    } else {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < NB2; ++cb) {
          const int col = cb * 16 + lr;
          if (col >= a.nout) continue;
          const float bias = a.b2[col];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int grow = r0 + rb * 16 + 4 * (lane >> 4) + r;
            if (grow >= a.M) continue;
            float v = acc2[rb][cb][r] + bias;
            if (a.res) v += (float)a.res[(long)grow * a.nout + col];
            if (a.out) a.out[(long)grow * a.nout + col] = (bf16)v;
            if (a.outf) a.outf[(long)grow * a.nout + col] = v;
          }
        }
    }
    __syncthreads(); // the hidden tile is rewritten by the next tile
  }
}

struct BwdArgs {
  int M, K, NO, nout;
  const float* dY;   // [M][nout]
  const bf16* hid;   // [M][128] post-ReLU hidden of the forward pass
  const bf16* W2;    // [128][NO]
  const bf16* W1;    // [K][128]
  bf16* dH;          // [M][128] gradient at the hidden pre-activation (out)
  float* dX[3];      // [M][width_q] gradient of each input source, or null
  int width[3];
  int start[4];
  int accumulate;    // bit q: dX[q] += instead of =
};

// dH = (dY W2^T) * [hid > 0];  dX = dH W1^T, split by source.
template <int NB2>
__global__ __launch_bounds__(kThreads) void k_mlp_bwd(BwdArgs a) {
  __shared__ bf16 sh[4][kWaveRows][kLat + 8];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int ntiles = (a.M + kTileRows - 1) / kTileRows;
  const int NOp = NB2 * 16;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int r0 = t * kTileRows + w * kWaveRows;
    int rows[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const int r = r0 + rb * 16 + lr;
      rows[rb] = r < a.M ? r : a.M - 1;
    }
    // dH: inner dimension = output columns (padded to 32 with zeros)
    f32x4 acc[2][8];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int n0 = 0; n0 < NOp; n0 += 32) {
      bf16x8 av[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int n = n0 + lk;
        if (n + 8 <= a.nout) {
          av[rb] = ld8f(a.dY + (long)rows[rb] * a.nout + n);
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q)
            av[rb][q] = (bf16)((n + q < a.nout) ? a.dY[(long)rows[rb] * a.nout + n + q] : 0.f);
        }
      }
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        // B[n][k] = W2[k][n]: 8 consecutive n of row k = cb*16 + lr
        bf16x8 bv;
        if (n0 + lk + 8 <= NOp) bv = ld8(a.W2 + (long)(cb * 16 + lr) * NOp + n0 + lk);
        else {
#pragma unroll
          for (int q = 0; q < 8; ++q) bv[q] = (bf16)0.f;
        }
        acc[0][cb] = mfma(av[0], bv, acc[0][cb]);
        acc[1][cb] = mfma(av[1], bv, acc[1][cb]);
      }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        const int col = cb * 16 + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lrow = rb * 16 + 4 * (lane >> 4) + r;
          const int grow = r0 + lrow;
          const int srow = grow < a.M ? grow : a.M - 1;
          const float hv = (float)a.hid[(long)srow * kLat + col];
          const bf16 g = (bf16)(hv > 0.f ? acc[rb][cb][r] : 0.f);
          sh[w][lrow][col] = g;
          if (grow < a.M) a.dH[(long)grow * kLat + col] = g;
        }
      }
    __syncthreads();
    // dX = dH W1^T: inner = 128 hidden units, output columns = K inputs
    for (int c0 = 0; c0 < a.K; c0 += 16 * 8) {
      f32x4 acc2[2][8];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) acc2[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < kLat; k0 += 32) {
        bf16x8 av[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = *reinterpret_cast<const bf16x8*>(&sh[w][rb * 16 + lr][k0 + lk]);
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
          const int col = c0 + cb * 16 + lr;
          bf16x8 bv;
          if (col < a.K) bv = ld8(a.W1 + (long)col * kLat + k0 + lk); // B[k][col] = W1[col][k]
          else {
#pragma unroll
            for (int q = 0; q < 8; ++q) bv[q] = (bf16)0.f;
          }
          acc2[0][cb] = mfma(av[0], bv, acc2[0][cb]);
          acc2[1][cb] = mfma(av[1], bv, acc2[1][cb]);
        }
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
          const int col = c0 + cb * 16 + lr;
          if (col >= a.K) continue;
          const int q = (col >= a.start[1]) ? ((col >= a.start[2]) ? 2 : 1) : 0;
          float* dst = a.dX[q];
          if (!dst) continue;
          const int wq = a.width[q], cc = col - a.start[q];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int grow = r0 + rb * 16 + 4 * (lane >> 4) + r;
            if (grow >= a.M) continue;
            float* p = dst + (long)grow * wq + cc;
            *p = ((a.accumulate >> q) & 1) ? *p + acc2[rb][cb][r] : acc2[rb][cb][r];
          }
        }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Edge MLP with the receiver's share of its first layer precomputed per node:
//   pre = e W1a + (h W1b)[recv] + h[send] W1c + b1,
// W1 = [W1a; W1b; W1c] (rows 0..127, 128..255, 256..383 of W1 [384][128]).
// k_node_p forms P = h W1b (float [N][128]) once per layer; consecutive edges
// share their receiver (the list is receiver-major), so P rows are read from
// cache.  The sender's share goes through the MFMA units from the gathered
// BF16 row h[send] (256 B per edge, the node table fits an XCD's L2).
//
// The edge kernel computes transposed products, hidden^T = W1a^T e^T +
// W1c^T h[send]^T and out^T = W2^T hidden^T: the weights are the A operand,
// in LDS for the block, and 16 edges are the columns of each MFMA tile, so a
// lane's 4 accumulators are 4 consecutive units of one edge (one float4 of
// P and of each bias per lane and tile).
// ---------------------------------------------------------------------------

struct NodePArgs {
  int N;
  const bf16* h;     // [N][128]
  const bf16* W1t;   // [128][384]: row = hidden unit
  float* P;          // [N][128]
};

// P[n][c] = sum_k h[n][k] W1[128 + k][c].  A wave owns 32 nodes; the
// output goes through LDS so that whole rows are stored.
__global__ __launch_bounds__(kThreads) void k_node_p(NodePArgs a) {
  __shared__ float sp[4][kWaveRows][kLat + 4]; // output staging per wave
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int r0 = (blockIdx.x * 4 + w) * kWaveRows;
  if (r0 >= a.N) return;
  int rows[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int r = r0 + rb * 16 + lr;
    rows[rb] = r < a.N ? r : a.N - 1;
  }
  f32x4 acc[2][8];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    bf16x8 av[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) av[rb] = ld8(a.h + (long)rows[rb] * kLat + ks * 32 + lk);
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
      // B[k][c] = W1[128 + k][c] = W1t[c][128 + k]
      const bf16x8 bv = ld8(a.W1t + (long)(cb * 16 + lr) * (3 * kLat) + kLat + ks * 32 + lk);
      acc[0][cb] = mfma(av[0], bv, acc[0][cb]);
      acc[1][cb] = mfma(av[1], bv, acc[1][cb]);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) sp[w][rb * 16 + 4 * (lane >> 4) + r][cb * 16 + lr] = acc[rb][cb][r];
  wave_lds_sync();
  // whole rows out: 2 rows of 128 floats per instruction
#pragma unroll
  for (int it = 0; it < kWaveRows / 2; ++it) {
    const int r = it * 2 + (lane >> 5), c = (lane & 31) * 4;
    if (r0 + r < a.N)
      *reinterpret_cast<float4*>(a.P + (long)(r0 + r) * kLat + c) = *reinterpret_cast<const float4*>(&sp[w][r][c]);
  }
}

struct EdgeFwdArgs {
  int E;
  const bf16* e;      // [E][128]
  const int* recv;
  const int* send;
  const float* P;     // [N][128]
  const bf16* h;      // [N][128]
  const bf16* W1t;    // [128][384]
  const float* b1;    // [128]
  const bf16* W2t;    // [128][128]
  const float* b2;    // [128]
  bf16* hid;          // [E][128] post-ReLU hidden, or null
  bf16* out;          // [E][128]: e + MLP
  // with the encoder fused in (e not read): e = We2 relu(We1 xe + be1) + be2
  const bf16* xe;     // [E][32]
  const bf16* We1t;   // [128][32]
  const float* be1;
  const bf16* We2t;   // [128][128]
  const float* be2;
};

// A wave tile of 16 EB edges: their rows, receivers, and the e^T (or, with
// the encoder fused in, xe^T), h[send]^T and (receiver term in the kernel)
// h[recv]^T fragments.
template <int EB>
struct EdgeTile {
  int ed[EB], ri[EB];
  bf16x8 Be[EB][4], Bh[EB][4], Br[EB][4], Bx[EB];
};

template <int EB, bool ENC, bool PIN>
__device__ inline void edge_tile_load(const EdgeFwdArgs& a, int e0, int lr, int lk, EdgeTile<EB>& T) {
#pragma unroll
  for (int eb = 0; eb < EB; ++eb) {
    const int ee = e0 + eb * 16 + lr;
    T.ed[eb] = ee < a.E ? ee : a.E - 1; // clamped loads; stores are guarded
    T.ri[eb] = a.recv[T.ed[eb]];
    const int sj = a.send[T.ed[eb]];
    if (ENC) T.Bx[eb] = ld8(a.xe + (long)T.ed[eb] * 32 + lk);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      if (!ENC) T.Be[eb][ks] = ld8(a.e + (long)T.ed[eb] * kLat + ks * 32 + lk);
      T.Bh[eb][ks] = ld8(a.h + (long)sj * kLat + ks * 32 + lk);
      if (PIN) T.Br[eb][ks] = ld8(a.h + (long)T.ri[eb] * kLat + ks * 32 + lk);
    }
  }
}

// Row order of the weight tiles in LDS.  MFMA block ub of 16 A rows yields,
// in lane (lr, lg), the D rows 4 lg .. 4 lg + 3 for column (edge) lr.  Placing
// unit o(p) = 32 s + 8 g + 4 t + r at A row p = 16 (2 s + t) + 4 g + r puts
// units 32 s + 8 lg .. + 7 of edge lr in lane (lr, lg), blocks 2 s and 2 s + 1:
// exactly that lane's B fragment for k step s of the next product and its
// e^T fragment for channels 32 s .. 32 s + 31.  So the hidden layer feeds
// layer 2 from registers, the residual e is a register add, and each lane
// stores 16 contiguous bytes of a row.
__device__ inline int unit_of_row(int p) {
  const int blk = p >> 4, g = (p >> 2) & 3, r = p & 3;
  return 32 * (blk >> 1) + 8 * g + 4 * (blk & 1) + r;
}

// Two D fragments (4 FP32 each) as one B fragment of 8 BF16, optionally
// through the ReLU.
template <bool RELU>
__device__ inline bf16x8 pack8(f32x4 lo, f32x4 hi) {
  if (RELU) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      lo[r] = lo[r] > 0.f ? lo[r] : 0.f;
      hi[r] = hi[r] > 0.f ? hi[r] : 0.f;
    }
  }
  const bf16x4 l = __builtin_convertvector(lo, bf16x4), h = __builtin_convertvector(hi, bf16x4);
  return __builtin_shufflevector(l, h, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Persistent and wave-granular: every wave takes tiles of 16 EB edges with
// stride (waves in the grid) and loads the next tile's rows while it
// multiplies the current one.  W1a^T, W1c^T and W2^T sit in LDS for the
// block's waves (rows in unit_of_row order), each weight fragment read once
// per tile and applied to its EB edge blocks; nothing else goes through LDS.
// With a grid that is a multiple of 8, XCD (blockIdx % 8) takes a contiguous
// eighth of the tiles, so the rows its edges gather stay in that XCD's L2.
// ENC: the edge encoder runs first on each tile (its weights in LDS too),
// in the dense kernel's arithmetic, so e is never stored or read.
// PIN: the receiver term h[recv] W1b is formed on the MFMA units from the
// gathered row h[recv] (W1b in LDS too) instead of read from P, so P is
// neither stored by the node pass nor read here.
template <int EB, int WAVES, bool ENC, bool PIN>
__global__ __launch_bounds__(64 * WAVES, 1) void k_edge_fwd(EdgeFwdArgs a) {
  static_assert(!(ENC && PIN), "the encoder's weights and W1b do not fit in LDS together");
  constexpr int TE = 16 * EB;
  __shared__ bf16 swa[kLat][kWPitch]; // W1a^T: [row][k]
  __shared__ bf16 swc[kLat][kWPitch]; // W1c^T: [row][k]
  __shared__ bf16 sw2[kLat][kWPitch]; // W2^T:  [row][k]
  __shared__ bf16 se1[ENC ? kLat : 1][32 + 8];   // encoder W1^T: [row][k]
  __shared__ bf16 se2[ENC ? kLat : 1][kWPitch];  // encoder W2^T: [row][k]
  __shared__ bf16 swb[PIN ? kLat : 1][kWPitch];  // W1b^T: [row][k]
  for (int c = threadIdx.x; c < kLat * 16; c += 64 * WAVES) {
    const int r = c >> 4, k8 = (c & 15) * 8, u = unit_of_row(r);
    *reinterpret_cast<bf16x8*>(&swa[r][k8]) = ld8(a.W1t + (long)u * (3 * kLat) + k8);
    *reinterpret_cast<bf16x8*>(&swc[r][k8]) = ld8(a.W1t + (long)u * (3 * kLat) + 2 * kLat + k8);
    *reinterpret_cast<bf16x8*>(&sw2[r][k8]) = ld8(a.W2t + (long)u * kLat + k8);
    if (PIN) *reinterpret_cast<bf16x8*>(&swb[r][k8]) = ld8(a.W1t + (long)u * (3 * kLat) + kLat + k8);
    if (ENC) {
      *reinterpret_cast<bf16x8*>(&se2[r][k8]) = ld8(a.We2t + (long)u * kLat + k8);
      if (k8 < 32) *reinterpret_cast<bf16x8*>(&se1[r][k8]) = ld8(a.We1t + (long)u * 32 + k8);
    }
  }
  __syncthreads(); // the only block barrier: the loop below is per wave
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4, lk = 8 * lg;
  const int ntiles = (a.E + TE - 1) / TE;
  int t, stride, t_end;
  if (gridDim.x % 8 == 0) {
    const int per = (ntiles + 7) / 8;
    const int x = blockIdx.x % 8;
    t = x * per + (blockIdx.x / 8) * WAVES + w;
    stride = (gridDim.x / 8) * WAVES;
    t_end = min(ntiles, (x + 1) * per);
  } else {
    t = blockIdx.x * WAVES + w;
    stride = gridDim.x * WAVES;
    t_end = ntiles;
  }
  if (t >= t_end) return;
  EdgeTile<EB> cur, nxt;
  edge_tile_load<EB, ENC, PIN>(a, t * TE, lr, lk, cur);
  for (; t < t_end; t += stride) {
    const int e0 = t * TE;
    const int tn = t + stride;
    if (tn < t_end) edge_tile_load<EB, ENC, PIN>(a, tn * TE, lr, lk, nxt);
    // keeps the weight fragments' LDS reads inside the loop: hoisted out of
    // it as loop-invariant they would need 384 registers
    __asm__ volatile("" ::: "memory");
    if (ENC) {
      // e^T fragments = encoder(xe^T), as k_dense_fwd<32> computes them
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) {
        f32x4 q[8];
#pragma unroll
        for (int ub = 0; ub < 8; ++ub) {
          const float4 bb = *reinterpret_cast<const float4*>(a.be1 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
          q[ub] = mfma(*reinterpret_cast<const bf16x8*>(&se1[ub * 16 + lr][lk]), cur.Bx[eb],
                       f32x4{bb.x, bb.y, bb.z, bb.w});
        }
        bf16x8 qb[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) qb[s4] = pack8<true>(q[2 * s4], q[2 * s4 + 1]);
#pragma unroll
        for (int ub = 0; ub < 8; ++ub) {
          const float4 bb = *reinterpret_cast<const float4*>(a.be2 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
          q[ub] = f32x4{bb.x, bb.y, bb.z, bb.w};
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
          for (int ub = 0; ub < 8; ++ub)
            q[ub] = mfma(*reinterpret_cast<const bf16x8*>(&se2[ub * 16 + lr][ks * 32 + lk]), qb[ks], q[ub]);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) cur.Be[eb][s4] = pack8<false>(q[2 * s4], q[2 * s4 + 1]);
      }
    }
    // layer 1 on top of b1 + P[recv] (PIN: b1 alone, the receiver term comes
    // from h[recv] below): block 2 s + t holds units 32 s + 8 lg + 4 t .. + 3
    f32x4 acc[8][EB];
#pragma unroll
    for (int ub = 0; ub < 8; ++ub) {
      const int u = 32 * (ub >> 1) + lk + 4 * (ub & 1);
      const float4 bb = *reinterpret_cast<const float4*>(a.b1 + u);
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) {
        if (PIN) {
          acc[ub][eb] = f32x4{bb.x, bb.y, bb.z, bb.w};
        } else {
          const float4 p = *reinterpret_cast<const float4*>(a.P + (long)cur.ri[eb] * kLat + u);
          acc[ub][eb] = f32x4{bb.x + p.x, bb.y + p.y, bb.z + p.z, bb.w + p.w};
        }
      }
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub) {
        const bf16x8 Aa = *reinterpret_cast<const bf16x8*>(&swa[ub * 16 + lr][ks * 32 + lk]);
        const bf16x8 Ac = *reinterpret_cast<const bf16x8*>(&swc[ub * 16 + lr][ks * 32 + lk]);
#pragma unroll
        for (int eb = 0; eb < EB; ++eb) {
          acc[ub][eb] = mfma(Aa, cur.Be[eb][ks], acc[ub][eb]);
          acc[ub][eb] = mfma(Ac, cur.Bh[eb][ks], acc[ub][eb]);
        }
        if (PIN) {
          const bf16x8 Ab = *reinterpret_cast<const bf16x8*>(&swb[ub * 16 + lr][ks * 32 + lk]);
#pragma unroll
          for (int eb = 0; eb < EB; ++eb) acc[ub][eb] = mfma(Ab, cur.Br[eb][ks], acc[ub][eb]);
        }
      }
    // relu -> BF16: the hidden^T fragment of k step s is blocks 2 s, 2 s + 1
    bf16x8 hb[EB][4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) hb[eb][s4] = pack8<true>(acc[2 * s4][eb], acc[2 * s4 + 1][eb]);
    if (a.hid) {
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) {
        const int ee = e0 + eb * 16 + lr;
        if (ee < a.E)
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4)
            *reinterpret_cast<bf16x8*>(a.hid + (long)ee * kLat + 32 * s4 + lk) = hb[eb][s4];
      }
    }
    // layer 2 on top of b2 + e (the residual, from the e^T fragments)
    f32x4 acc2[8][EB];
#pragma unroll
    for (int ub = 0; ub < 8; ++ub) {
      const float4 bb = *reinterpret_cast<const float4*>(a.b2 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) {
        const bf16x8 ev = cur.Be[eb][ub >> 1];
        const int j0 = 4 * (ub & 1);
        acc2[ub][eb] = f32x4{bb.x + (float)ev[j0], bb.y + (float)ev[j0 + 1], bb.z + (float)ev[j0 + 2],
                             bb.w + (float)ev[j0 + 3]};
      }
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub) {
        const bf16x8 A2 = *reinterpret_cast<const bf16x8*>(&sw2[ub * 16 + lr][ks * 32 + lk]);
#pragma unroll
        for (int eb = 0; eb < EB; ++eb) acc2[ub][eb] = mfma(A2, hb[eb][ks], acc2[ub][eb]);
      }
    // 16 contiguous bytes per lane and k step: units 32 s + 8 lg .. + 7 of edge lr
#pragma unroll
    for (int eb = 0; eb < EB; ++eb) {
      const int ee = e0 + eb * 16 + lr;
      if (ee < a.E)
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
          *reinterpret_cast<bf16x8*>(a.out + (long)ee * kLat + 32 * s4 + lk) =
              pack8<false>(acc2[2 * s4][eb], acc2[2 * s4 + 1][eb]);
    }
    if (tn < t_end) cur = nxt;
  }
}

// Launch shape of the edge kernel: 16 EB edges per wave tile, WAVES waves per
// block (one block per CU: three 34 KB weight tiles, plus W1b or the
// encoder's weights, fill the LDS).
#ifndef GNN_EDGE_EB
#define GNN_EDGE_EB 1
#endif
#ifndef GNN_EDGE_WAVES
#define GNN_EDGE_WAVES 8
#endif
constexpr int kEdgeEB = GNN_EDGE_EB, kEdgeWaves = GNN_EDGE_WAVES;

// ---------------------------------------------------------------------------
// Dense MLP y = res + W2 relu(W1 [x0, x1] + b1) + b2 with 128 outputs, the
// inputs ungathered rows (K = 32, 128 or 256: x1 holds columns 128..255).
// The edge kernel's scheme: persistent, wave-granular, weights in LDS for
// the whole block (rows in unit_of_row order), transposed products (rows are
// the 16 MFMA columns), the next tile's rows loaded while the current one
// multiplies, and the hidden layer and output in registers.
// ---------------------------------------------------------------------------

struct DenseFwdArgs {
  int M;
  const bf16* x0;     // [M][min(K, 128)]
  const bf16* x1;     // [M][128] (K = 256) or null
  const bf16* W1t;    // [128][K]
  const float* b1;
  const bf16* W2t;    // [128][128]
  const float* b2;
  const bf16* res;    // [M][128] or null
  bf16* hid;          // [M][128] or null
  bf16* out;          // [M][128]
};

template <int K, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 1) void k_dense_fwd(DenseFwdArgs a) {
  constexpr int KS = K / 32;          // k steps of the first layer
  constexpr int W0 = K < kLat ? K : kLat; // width of x0
  __shared__ bf16 sw1[kLat][K + 8];   // W1^T: [row][k], rows in unit_of_row order
  __shared__ bf16 sw2[kLat][kWPitch]; // W2^T: [row][k], likewise
  for (int c = threadIdx.x; c < kLat * (K / 8); c += 64 * WAVES) {
    const int r = c / (K / 8), k8 = (c % (K / 8)) * 8;
    *reinterpret_cast<bf16x8*>(&sw1[r][k8]) = ld8(a.W1t + (long)unit_of_row(r) * K + k8);
  }
  for (int c = threadIdx.x; c < kLat * 16; c += 64 * WAVES) {
    const int r = c >> 4, k8 = (c & 15) * 8;
    *reinterpret_cast<bf16x8*>(&sw2[r][k8]) = ld8(a.W2t + (long)unit_of_row(r) * kLat + k8);
  }
  __syncthreads(); // the only block barrier
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4, lk = 8 * lg;
  const int ntiles = (a.M + 15) / 16;
  const int stride = gridDim.x * WAVES;
  int t = blockIdx.x * WAVES + w;
  if (t >= ntiles) return;
  auto load_rows = [&](int m0, bf16x8 (&B)[KS]) {
    const int m = m0 + lr < a.M ? m0 + lr : a.M - 1;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = ks * 32 + lk;
      B[ks] = ks * 32 < W0 ? ld8(a.x0 + (long)m * W0 + k) : ld8(a.x1 + (long)m * kLat + (k - kLat));
    }
  };
  bf16x8 cur[KS], nxt[KS];
  load_rows(t * 16, cur);
  for (; t < ntiles; t += stride) {
    // keeps the weight fragments' LDS reads inside the loop
    __asm__ volatile("" ::: "memory");
    const int m0 = t * 16;
    const int tn = t + stride;
    if (tn < ntiles) load_rows(tn * 16, nxt);
    const bool live = m0 + lr < a.M;
    const int mrow = live ? m0 + lr : a.M - 1;
    // layer 1: block 2 s + t holds units 32 s + 8 lg + 4 t .. + 3
    f32x4 acc[8];
#pragma unroll
    for (int ub = 0; ub < 8; ++ub) {
      const float4 bb = *reinterpret_cast<const float4*>(a.b1 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
      acc[ub] = f32x4{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub)
        acc[ub] = mfma(*reinterpret_cast<const bf16x8*>(&sw1[ub * 16 + lr][ks * 32 + lk]), cur[ks], acc[ub]);
    bf16x8 hb[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) hb[s4] = pack8<true>(acc[2 * s4], acc[2 * s4 + 1]);
    if (a.hid && live)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) *reinterpret_cast<bf16x8*>(a.hid + (long)mrow * kLat + 32 * s4 + lk) = hb[s4];
    // layer 2 on top of b2 (+ the residual row)
    f32x4 acc2[8];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int u = 32 * s4 + lk;
      const float4 b0 = *reinterpret_cast<const float4*>(a.b2 + u);
      const float4 b1 = *reinterpret_cast<const float4*>(a.b2 + u + 4);
      acc2[2 * s4] = f32x4{b0.x, b0.y, b0.z, b0.w};
      acc2[2 * s4 + 1] = f32x4{b1.x, b1.y, b1.z, b1.w};
      if (a.res) {
        const bf16x8 rv = ld8(a.res + (long)mrow * kLat + u);
        acc2[2 * s4] += f32x4{(float)rv[0], (float)rv[1], (float)rv[2], (float)rv[3]};
        acc2[2 * s4 + 1] += f32x4{(float)rv[4], (float)rv[5], (float)rv[6], (float)rv[7]};
      }
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub)
        acc2[ub] = mfma(*reinterpret_cast<const bf16x8*>(&sw2[ub * 16 + lr][ks * 32 + lk]), hb[ks], acc2[ub]);
    if (live)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
        *reinterpret_cast<bf16x8*>(a.out + (long)mrow * kLat + 32 * s4 + lk) =
            pack8<false>(acc2[2 * s4], acc2[2 * s4 + 1]);
    if (tn < ntiles) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) cur[ks] = nxt[ks];
    }
  }
}

// ---------------------------------------------------------------------------
// Node side of a message-passing layer in one pass:
//   agg = sum of the node's edge rows (gnn_segment_sum, same order),
//   h'  = h + W2 relu(W1 [h, agg] + b1) + b2,
//   P'  = h' W1b of the next layer's edge MLP (gnn_node_p),
// with the dense kernel's scheme: weights in LDS (rows in unit_of_row
// order, the next layer's W1b beside them), transposed products, and the
// layers chained in registers.
// ---------------------------------------------------------------------------

struct NodeFwdArgs {
  int N;
  const bf16* h;      // [N][128]
  const bf16* e;      // [E][128]: this layer's edge latents, receiver-major
  const int* row;     // [N + 1]
  const bf16* W1t;    // [128][256]
  const float* b1;
  const bf16* W2t;    // [128][128]
  const float* b2;
  bf16* agg;          // [N][128] or null
  bf16* hid;          // [N][128] or null
  bf16* out;          // [N][128]
  const bf16* Wn1t;   // next layer's edge W1t [128][384], or null
  float* P;           // [N][128] when Wn1t is set
};

// Launch shape of the node pass: WAVES waves per block (one block per CU:
// the weights fill 137 KB of LDS) and AGG edge rows in flight per lane in the
// aggregation.  Four waves per SIMD with two rows in flight measured best at
// 64k nodes (two waves with six rows in flight were 4 % slower).
#ifndef GNN_NODE_WAVES
#define GNN_NODE_WAVES 16
#endif
#ifndef GNN_NODE_AGG
#define GNN_NODE_AGG 2
#endif

template <int WAVES, int AGG>
__global__ __launch_bounds__(64 * WAVES, 1) void k_node_fwd(NodeFwdArgs a) {
  constexpr int K = 2 * kLat, KS = K / 32;
  __shared__ bf16 sw1[kLat][K + 8];   // W1^T: [row][k], rows in unit_of_row order
  __shared__ bf16 sw2[kLat][kWPitch]; // W2^T: [row][k], likewise
  __shared__ bf16 swn[kLat][kWPitch]; // next layer's W1b^T: [row][k], likewise
  for (int c = threadIdx.x; c < kLat * (K / 8); c += 64 * WAVES) {
    const int r = c / (K / 8), k8 = (c % (K / 8)) * 8;
    *reinterpret_cast<bf16x8*>(&sw1[r][k8]) = ld8(a.W1t + (long)unit_of_row(r) * K + k8);
  }
  for (int c = threadIdx.x; c < kLat * 16; c += 64 * WAVES) {
    const int r = c >> 4, k8 = (c & 15) * 8;
    *reinterpret_cast<bf16x8*>(&sw2[r][k8]) = ld8(a.W2t + (long)unit_of_row(r) * kLat + k8);
    if (a.Wn1t)
      *reinterpret_cast<bf16x8*>(&swn[r][k8]) = ld8(a.Wn1t + (long)unit_of_row(r) * (3 * kLat) + kLat + k8);
  }
  __syncthreads(); // the only block barrier
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4, lk = 8 * lg;
  const int ntiles = (a.N + 15) / 16;
  for (int t = blockIdx.x * WAVES + w; t < ntiles; t += gridDim.x * WAVES) {
    // keeps the weight fragments' LDS reads inside the loop
    __asm__ volatile("" ::: "memory");
    const int m0 = t * 16;
    const bool live = m0 + lr < a.N;
    const int m = live ? m0 + lr : a.N - 1; // clamped loads; stores are guarded
    bf16x8 B[KS];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) B[ks] = ld8(a.h + (long)m * kLat + ks * 32 + lk);
    // aggregation over the node's edges in edge order, FP32 sums (AGG edges'
    // rows in flight at a time)
    float s[4][8];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int q = 0; q < 8; ++q) s[ks][q] = 0.f;
    const int eb = a.row[m], ee = a.row[m + 1];
    int e = eb;
    for (; e + AGG - 1 < ee; e += AGG) {
      bf16x8 v[AGG][4];
#pragma unroll
      for (int u = 0; u < AGG; ++u)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) v[u][ks] = ld8(a.e + (long)(e + u) * kLat + ks * 32 + lk);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int u = 0; u < AGG; ++u) s[ks][q] += (float)v[u][ks][q];
    }
    for (; e < ee; ++e) {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 v = ld8(a.e + (long)e * kLat + ks * 32 + lk);
#pragma unroll
        for (int q = 0; q < 8; ++q) s[ks][q] += (float)v[q];
      }
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
      for (int q = 0; q < 8; ++q) B[4 + ks][q] = (bf16)s[ks][q];
      if (a.agg && live) *reinterpret_cast<bf16x8*>(a.agg + (long)m * kLat + ks * 32 + lk) = B[4 + ks];
    }
    // layer 1 on top of b1: block 2 s + t holds units 32 s + 8 lg + 4 t .. + 3
    f32x4 acc[8];
#pragma unroll
    for (int ub = 0; ub < 8; ++ub) {
      const float4 bb = *reinterpret_cast<const float4*>(a.b1 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
      acc[ub] = f32x4{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub)
        acc[ub] = mfma(*reinterpret_cast<const bf16x8*>(&sw1[ub * 16 + lr][ks * 32 + lk]), B[ks], acc[ub]);
    bf16x8 hb[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) hb[s4] = pack8<true>(acc[2 * s4], acc[2 * s4 + 1]);
    if (a.hid && live)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) *reinterpret_cast<bf16x8*>(a.hid + (long)m * kLat + 32 * s4 + lk) = hb[s4];
    // layer 2 on top of b2 + h (the residual, from the h^T fragments)
    f32x4 acc2[8];
#pragma unroll
    for (int ub = 0; ub < 8; ++ub) {
      const float4 bb = *reinterpret_cast<const float4*>(a.b2 + 32 * (ub >> 1) + lk + 4 * (ub & 1));
      const bf16x8 hv = B[ub >> 1];
      const int j0 = 4 * (ub & 1);
      acc2[ub] = f32x4{bb.x + (float)hv[j0], bb.y + (float)hv[j0 + 1], bb.z + (float)hv[j0 + 2],
                       bb.w + (float)hv[j0 + 3]};
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ub = 0; ub < 8; ++ub)
        acc2[ub] = mfma(*reinterpret_cast<const bf16x8*>(&sw2[ub * 16 + lr][ks * 32 + lk]), hb[ks], acc2[ub]);
    bf16x8 ob[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) ob[s4] = pack8<false>(acc2[2 * s4], acc2[2 * s4 + 1]);
    if (live)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) *reinterpret_cast<bf16x8*>(a.out + (long)m * kLat + 32 * s4 + lk) = ob[s4];
    if (a.Wn1t) {
      // P' = h' W1b (next layer): A = W1b^T fragments, B = h'^T from the
      // output fragments
      f32x4 accp[8];
#pragma unroll
      for (int ub = 0; ub < 8; ++ub) accp[ub] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int ub = 0; ub < 8; ++ub)
          accp[ub] = mfma(*reinterpret_cast<const bf16x8*>(&swn[ub * 16 + lr][ks * 32 + lk]), ob[ks], accp[ub]);
      if (live) {
#pragma unroll
        for (int ub = 0; ub < 8; ++ub)
          *reinterpret_cast<float4*>(a.P + (long)m * kLat + 32 * (ub >> 1) + lk + 4 * (ub & 1)) =
              make_float4(accp[ub][0], accp[ub][1], accp[ub][2], accp[ub][3]);
      }
    }
  }
}

struct WgradArgs {
  int M, K, NC;      // rows, input features, output columns (multiple of 16)
  int ncols_g;       // real columns of G (<= NC)
  Segs in;           // X sources (bf16), or
  const bf16* gb;    // G as bf16 [M][ncols_g], or
  const float* gf;   // G as fp32 [M][ncols_g]
  int chunk;         // rows per chunk
  float* part;       // [nchunks][K][NC] partial dW
  float* bpart;      // [nchunks][NC] partial column sums of G (bias gradient)
};

// Partial dW[k][n] = sum over a chunk of rows of X[m][k] G[m][n].
// Block (kb, chunk): k rows [64 kb, 64 kb + 64), wave w the 16 of them at
// 16 w.  X and G tiles of 32 rows go through LDS transposed, so each lane
// reads 8 consecutive rows m for its k (A) and its n (B).
__global__ __launch_bounds__(kThreads) void k_wgrad(WgradArgs a) {
  __shared__ bf16 sx[64][32 + 8];     // [k][m]
  __shared__ bf16 sg[kLat][32 + 8];   // [n][m]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int kb = blockIdx.x, ch = blockIdx.y;
  const int k0 = kb * 64;
  const int m_begin = ch * a.chunk;
  const int m_end = min(a.M, m_begin + a.chunk);
  const int NB = a.NC / 16;
  f32x4 acc[8];
#pragma unroll
  for (int cb = 0; cb < 8; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  // column sums of G (bias gradient, kb == 0 blocks), in FP32 from the
  // source values: thread t always loads the 8 columns starting at
  // 8 (t mod NC/8), since the block size is a multiple of NC/8
  float csum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int m0 = m_begin; m0 < m_end; m0 += 32) {
    // X tile: 32 rows x 64 k; thread t loads 8 k of row t / 8
    {
      const int rr = threadIdx.x >> 3, kk = (threadIdx.x & 7) * 8;
      const int m = m0 + rr;
      bf16x8 v;
      if (m < m_end && k0 + kk < a.K) v = seg_load(a.in, m, k0 + kk);
      else {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = (bf16)0.f;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) sx[kk + q][rr] = v[q];
    }
    // G tile: 32 rows x NC columns
    for (int e = threadIdx.x; e < 32 * (a.NC / 8); e += kThreads) {
      const int rr = e / (a.NC / 8), nn = (e % (a.NC / 8)) * 8;
      const int m = m0 + rr;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float g = 0.f;
        if (m < m_end && nn + q < a.ncols_g)
          g = a.gb ? (float)a.gb[(long)m * a.ncols_g + nn + q] : a.gf[(long)m * a.ncols_g + nn + q];
        sg[nn + q][rr] = (bf16)g;
        csum[q] += g;
      }
    }
    __syncthreads();
    const bf16x8 av = *reinterpret_cast<const bf16x8*>(&sx[w * 16 + lr][lk]);
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
      if (cb >= NB) break;
      const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sg[cb * 16 + lr][lk]);
      acc[cb] = mfma(av, bv, acc[cb]);
    }
    __syncthreads();
  }
  float* out = a.part + (long)ch * a.K * a.NC;
#pragma unroll
  for (int cb = 0; cb < 8; ++cb) {
    if (cb >= NB) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + w * 16 + 4 * (lane >> 4) + r;
      if (k < a.K) out[(long)k * a.NC + cb * 16 + lr] = acc[cb][r];
    }
  }
  if (kb == 0) {
    // reduce the per-thread column sums in a fixed order
    __shared__ float red[kThreads][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) red[threadIdx.x][q] = csum[q];
    __syncthreads();
    if (threadIdx.x < a.NC) {
      const int groups = a.NC / 8, grp = threadIdx.x / 8, q = threadIdx.x % 8;
      float s = 0.f;
      for (int t = grp; t < kThreads; t += groups) s += red[t][q];
      a.bpart[(long)ch * a.NC + threadIdx.x] = s;
    }
  }
}

// out[i] = sum over chunks of part[c][i], in chunk order.
__global__ __launch_bounds__(kThreads) void k_reduce_chunks(const float* part, int nchunks,
                                                            long count, float* out, int accumulate) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c) s += part[(long)c * count + i];
  out[i] = accumulate ? out[i] + s : s;
}

// agg[i] = sum of x[e] over the edges e of node i (FP32 sums, BF16 out).
// 16 threads per node, 8 channels each.
__global__ __launch_bounds__(kThreads) void k_seg_sum(const bf16* x, const int* row, int N,
                                                      bf16* agg) {
  const int node = blockIdx.x * (kThreads / 16) + (threadIdx.x >> 4);
  if (node >= N) return;
  const int c = (threadIdx.x & 15) * 8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int e = row[node]; e < row[node + 1]; ++e) {
    const bf16x8 v = ld8(x + (long)e * kLat + c);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] += (float)v[q];
  }
  bf16x8 o;
#pragma unroll
  for (int q = 0; q < 8; ++q) o[q] = (bf16)s[q];
  *reinterpret_cast<bf16x8*>(agg + (long)node * kLat + c) = o;
}

// dh[i] += sum over edges e of i of (d_recv[e] + d_send[rev[e]])  (either may be null)
__global__ __launch_bounds__(kThreads) void k_edge_grad_to_nodes(const float* d_recv, const float* d_send,
                                                                 const int* row, const int* rev, int N,
                                                                 float* dh) {
  const int node = blockIdx.x * (kThreads / 16) + (threadIdx.x >> 4);
  if (node >= N) return;
  const int c = (threadIdx.x & 15) * 8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int e = row[node]; e < row[node + 1]; ++e) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (d_recv) s[q] += d_recv[(long)e * kLat + c + q];
      if (d_send) s[q] += d_send[(long)rev[e] * kLat + c + q];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) dh[(long)node * kLat + c + q] += s[q];
}

// d_e[e] += d_agg[recv[e]]  (the aggregation's backward)
__global__ __launch_bounds__(kThreads) void k_bcast_add(const float* d_agg, const int* recv, long E,
                                                        float* d_e) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x; // one thread per (edge, 8 channels)
  const long e = i >> 4;
  if (e >= E) return;
  const int c = (int)(i & 15) * 8;
  const float* src = d_agg + (long)recv[e] * kLat + c;
  float* dst = d_e + e * kLat + c;
#pragma unroll
  for (int q = 0; q < 8; ++q) dst[q] += src[q];
}

// Adam on a flat FP32 parameter vector (bias-corrected step sizes precomputed).
__global__ __launch_bounds__(kThreads) void k_adam(float* p, const float* g, float* m, float* v, long n,
                                                   float lr_t, float b1, float b2, float eps,
                                                   float wd) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// FP32 master [K][NP] -> BF16 W [K][NP] and Wt [NP][K] (columns n >= N are 0).
__global__ __launch_bounds__(kThreads) void k_cast_weights(const float* p, int K, int N, int NP,
                                                           bf16* W, bf16* Wt) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= (long)K * NP) return;
  const int k = (int)(i / NP), n = (int)(i % NP);
  const bf16 v = (bf16)(n < N ? p[i] : 0.f);
  W[(long)k * NP + n] = v;
  Wt[(long)n * K + k] = v;
}

// Mean squared error over M x C values: dY = 2 (y - t) / (M C) and per-block
// partial sums of (y - t)^2.
__global__ __launch_bounds__(kThreads) void k_mse(const float* y, const float* t, long count,
                                                  float* dY, float* part) {
  __shared__ float red[kThreads];
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  float d = 0.f;
  if (i < count) {
    d = y[i] - t[i];
    dY[i] = 2.f * d / (float)count;
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Node features, one row of 32 per particle: the last H velocities (oldest
// first) scaled by 1/v_scale, the radius as (r - r_mean) / r_scale, and the
// distance to each wall plane n . x - d clipped to [0, 1] after scaling by
// 1/dist_scale; the rest zero.
__global__ __launch_bounds__(kThreads) void k_node_features(int N, const double* x, const float* vhist,
                                                            int H, const double* rad, float inv_v,
                                                            float r_mean, float inv_r, const float* planes,
                                                            int nw, float inv_d, bf16* out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  float f[32];
#pragma unroll
  for (int q = 0; q < 32; ++q) f[q] = 0.f;
  int c = 0;
  for (int h = 0; h < H; ++h)
    for (int d = 0; d < 3; ++d) f[c++] = vhist[((long)h * N + i) * 3 + d] * inv_v;
  f[c++] = ((float)rad[i] - r_mean) * inv_r;
  for (int w = 0; w < nw; ++w) {
    const float* p = planes + 4 * w;
    float dist = (float)(p[0] * x[3l * i] + p[1] * x[3l * i + 1] + p[2] * x[3l * i + 2]) - p[3];
    dist *= inv_d;
    f[c++] = dist < 0.f ? 0.f : (dist > 1.f ? 1.f : dist);
  }
  bf16x8* o = reinterpret_cast<bf16x8*>(out + (long)i * 32);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    bf16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (bf16)f[8 * g + q];
    o[g] = v;
  }
}

// Edge features, one row of 32 per edge j -> i: (x_i - x_j) / R and |x_i - x_j| / R.
__global__ __launch_bounds__(kThreads) void k_edge_features(long E, const double* x, const int* recv,
                                                            const int* send, double inv_R, bf16* out) {
  const long e = (long)blockIdx.x * kThreads + threadIdx.x;
  if (e >= E) return;
  const int i = recv[e], j = send[e];
  const double d0 = x[3l * i] - x[3l * j], d1 = x[3l * i + 1] - x[3l * j + 1], d2 = x[3l * i + 2] - x[3l * j + 2];
  const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  bf16x8 v;
  v[0] = (bf16)(float)(d0 * inv_R); v[1] = (bf16)(float)(d1 * inv_R); v[2] = (bf16)(float)(d2 * inv_R);
  v[3] = (bf16)(float)(r * inv_R);
#pragma unroll
  for (int q = 4; q < 8; ++q) v[q] = (bf16)0.f;
  bf16x8 z;
#pragma unroll
  for (int q = 0; q < 8; ++q) z[q] = (bf16)0.f;
  bf16x8* o = reinterpret_cast<bf16x8*>(out + e * 32);
  o[0] = v; o[1] = z; o[2] = z; o[3] = z;
}

// One macro-step of the surrogate: v += a dt, x += v dt (semi-implicit
// Euler), the velocity history shifted by one (oldest dropped).
__global__ __launch_bounds__(kThreads) void k_integrate(int N, const float* a, float a_scale,
                                                        const double* g, double dt, double* x,
                                                        float* vhist, int H) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const float* vl = vhist + ((long)(H - 1) * N + i) * 3;
  double v[3];
  for (int d = 0; d < 3; ++d) v[d] = (double)vl[d] + ((double)a[3l * i + d] * a_scale + g[d]) * dt;
  for (int h = 0; h + 1 < H; ++h)
    for (int d = 0; d < 3; ++d) vhist[((long)h * N + i) * 3 + d] = vhist[((long)(h + 1) * N + i) * 3 + d];
  for (int d = 0; d < 3; ++d) {
    vhist[((long)(H - 1) * N + i) * 3 + d] = (float)v[d];
    x[3l * i + d] += v[d] * dt;
  }
}

// ---------------------------------------------------------------------------
// radius graph: uniform cells of side >= radius, a counting sort of the
// nodes into cells, a count pass, an exclusive scan into row offsets and a
// fill pass that writes each node's senders in ascending order.
// ---------------------------------------------------------------------------

struct CellGrid {
  double lo[3], inv[3];
  int nc[3];
  int ncell;
};

constexpr int kScanPerBlock = 4 * kThreads; // elements per block of the scan

// The count sorts each node's senders in LDS (one row per thread, threads in
// cell-slot order) and keeps rows of up to kFillLds of them in the workspace,
// so the fill copies them; longer rows are found again by the fill and
// sorted in place in send.
constexpr int kFillLds = 32;

__device__ inline int cell_coord(double x, double lo, double inv, int nc) {
  double t = (x - lo) * inv;
  if (!(t >= 0.)) t = 0.;          // below the box, or NaN
  if (t >= (double)nc) t = nc - 1; // above it (clamped before the conversion)
  return (int)t;
}

__global__ __launch_bounds__(kThreads) void k_g_zero(int* p, long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kThreads) void k_g_bin(int N, const double* x, CellGrid g, int* cell_of,
                                                    int* ccount) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const int cx = cell_coord(x[3l * i], g.lo[0], g.inv[0], g.nc[0]);
  const int cy = cell_coord(x[3l * i + 1], g.lo[1], g.inv[1], g.nc[1]);
  const int cz = cell_coord(x[3l * i + 2], g.lo[2], g.inv[2], g.nc[2]);
  const int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
  cell_of[i] = c;
  atomicAdd(&ccount[c], 1);
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total: block-local
// scans (kScanPerBlock elements each, block totals to bsum), a scan of the
// block totals in one block, and the carry added back.
__global__ __launch_bounds__(kThreads) void k_scan_local(const int* in, long n, int* out, int* bsum) {
  __shared__ int s[kThreads];
  const long base = (long)blockIdx.x * kScanPerBlock + 4l * threadIdx.x;
  int v[4], t = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = base + q < n ? in[base + q] : 0;
    t += v[q];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int d = 1; d < kThreads; d <<= 1) {
    const int u = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += u;
    __syncthreads();
  }
  int run = s[threadIdx.x] - t; // exclusive prefix of this thread
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == kThreads - 1) bsum[blockIdx.x] = s[kThreads - 1];
}

// The block totals are scanned in 64 bits: the caller reads the exact total
// (total64) to reject a graph whose edge count does not fit an int.
__global__ __launch_bounds__(kThreads) void k_scan_totals(int* bsum, int nb, int* out_total,
                                                          long long* total64) {
  __shared__ long long s[kThreads];
  long long carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kThreads) {
    const int i = b0 + threadIdx.x;
    const long long t = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = t;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
      const long long u = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += u;
      __syncthreads();
    }
    if (i < nb) bsum[i] = (int)(carry + s[threadIdx.x] - t);
    carry += s[kThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_total = (int)carry;
    if (total64) *total64 = carry;
  }
}

__global__ __launch_bounds__(kThreads) void k_scan_carry(int* out, long n, const int* bsum) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] += bsum[i / kScanPerBlock];
}

__global__ __launch_bounds__(kThreads) void k_g_scatter(int N, const double* x, const int* cell_of,
                                                        const int* cstart, int* cursor, int* cnodes,
                                                        double4* cpos) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const int c = cell_of[i];
  const int slot = cstart[c] + atomicAdd(&cursor[c], 1);
  cnodes[slot] = i;
  cpos[slot] = make_double4(x[3l * i], x[3l * i + 1], x[3l * i + 2], 0.);
}

// Visits the nodes j != i with |x_i - x_j|^2 <= r2 in the 27 cells around
// node i's cell c (the order depends on the scatter; callers sort).  The
// candidates' positions are read from the cell-ordered copy, so a row of
// cells is one contiguous stretch of memory.
template <class F>
__device__ inline void for_each_neighbor(int i, double4 xi, int c, const CellGrid& g, double r2,
                                         const int* cstart, const int* cnodes, const double4* cpos, F f) {
  const int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
  const double xi0 = xi.x, xi1 = xi.y, xi2 = xi.z;
  for (int z = cz > 0 ? cz - 1 : 0; z <= cz + 1 && z < g.nc[2]; ++z)
    for (int y = cy > 0 ? cy - 1 : 0; y <= cy + 1 && y < g.nc[1]; ++y) {
      const int row0 = (z * g.nc[1] + y) * g.nc[0];
      const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g.nc[0] ? cx + 1 : g.nc[0] - 1;
      // cells x0..x1 of one row are contiguous in cnodes
      for (int p = cstart[row0 + x0]; p < cstart[row0 + x1 + 1]; ++p) {
        const int j = cnodes[p];
        if (j == i) continue;
        const double4 xj = cpos[p];
        const double d0 = xi0 - xj.x, d1 = xi1 - xj.y, d2 = xi2 - xj.z;
        if (d0 * d0 + d1 * d1 + d2 * d2 <= r2) f(j);
      }
    }
}

__global__ __launch_bounds__(kThreads) void k_g_count(int N, CellGrid g, double r2, const int* cell_of,
                                                      const int* cstart, const int* cnodes, const double4* cpos,
                                                      int* deg, int* rows) {
  __shared__ int sb[kThreads * (kFillLds + 1)]; // thread t's row at t (kFillLds + 1): no bank conflicts
  const int t = blockIdx.x * kThreads + threadIdx.x; // cell slot: a wave's nodes share cells
  if (t >= N) return;
  const int i = cnodes[t];
  const double4 xi = cpos[t];
  int* b = sb + threadIdx.x * (kFillLds + 1);
  int n = 0;
  for_each_neighbor(i, xi, cell_of[i], g, r2, cstart, cnodes, cpos, [&](int j) {
                      if (n < kFillLds) { // insertion into the sorted prefix b[0..n)
                        int p = n;
                        while (p > 0 && b[p - 1] > j) {
                          b[p] = b[p - 1];
                          --p;
                        }
                        b[p] = j;
                      }
                      ++n;
                    });
  deg[i] = n;
  if (n <= kFillLds)
    for (int k = 0; k < n; ++k) rows[(long)i * kFillLds + k] = b[k];
}

__global__ __launch_bounds__(kThreads) void k_g_fill(int N, const double* x, CellGrid g, double r2,
                                                     const int* cell_of, const int* cstart, const int* cnodes,
                                                     const double4* cpos, const int* rows, const int* row, int E,
                                                     int* send, int* recv) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const int e0 = row[i], e1 = row[i + 1] < E ? row[i + 1] : E;
  if (e0 < 0) return;
  if (row[i + 1] - e0 <= kFillLds) { // the count kept the sorted row (degree from the count's scan)
    for (int k = 0; k < e1 - e0; ++k) {
      send[e0 + k] = rows[(long)i * kFillLds + k];
      recv[e0 + k] = i;
    }
    return;
  }
  int n = e0;
  for_each_neighbor(i, make_double4(x[3l * i], x[3l * i + 1], x[3l * i + 2], 0.), cell_of[i], g, r2, cstart,
                    cnodes, cpos, [&](int j) {
                      if (n < e1) { // the count pass saw the same set; the bound only guards memory
                        // insertion into the sorted prefix send[e0..n)
                        int p = n;
                        while (p > e0 && send[p - 1] > j) {
                          send[p] = send[p - 1];
                          --p;
                        }
                        send[p] = j;
                        recv[n] = i;
                        ++n;
                      }
                    });
}

// rev[e] for edge j -> i: the position of i among node j's senders
// (binary search in the sorted row of j; the graph is symmetric).
// An edge whose reverse is missing (impossible for this symmetric criterion)
// gets rev[e] = e, which keeps every later gather in bounds.
__global__ __launch_bounds__(kThreads) void k_g_rev(int N, const int* row, int E, const int* send,
                                                    int* rev) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const int e1 = row[i + 1] < E ? row[i + 1] : E;
  for (int e = row[i] > 0 ? row[i] : 0; e < e1; ++e) {
    const int j = send[e];
    if (j < 0 || j >= N) {
      rev[e] = e;
      continue;
    }
    int lo = row[j], hi = row[j + 1] < E ? row[j + 1] : E;
    const int end = hi;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (send[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    rev[e] = lo < end && send[lo] == i ? lo : e;
  }
}

inline unsigned nblk(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Cell grid of the box [lo, hi] with cells of side > radius (by a relative
// 1e-6, so that rounding in the cell coordinates never puts two nodes closer
// than the radius two cells apart); -1 if the box is degenerate or needs
// more than 2^26 cells.
int make_grid(const double lo[3], const double hi[3], double radius, CellGrid* g) {
  if (!lo || !hi || !(radius > 0.)) return -1;
  long nc = 1;
  for (int d = 0; d < 3; ++d) {
    const double w = hi[d] - lo[d];
    if (!(w > 0.)) return -1;
    double k = w / (radius * (1. + 1e-6));
    if (!(k < 1024.)) k = 1024.;
    g->nc[d] = k < 1. ? 1 : (int)k;
    g->lo[d] = lo[d];
    g->inv[d] = g->nc[d] / w;
    nc *= g->nc[d];
  }
  if (nc > (1l << 26)) return -1;
  g->ncell = (int)nc;
  return 0;
}

long scan_blocks(long n) { return (n + kScanPerBlock - 1) / kScanPerBlock; }

// Workspace of the radius graph, in ints.
struct GraphWork {
  int *ccount, *cstart, *cursor, *cell_of, *cnodes, *deg, *bsum;
  long long* total64;
  double4* cpos; // positions by cell slot (cnodes order)
  int* rows;     // [N][kFillLds]: each node's sorted senders, from the count
  long ints;
};

GraphWork carve(void* base, int N, int ncell) {
  GraphWork w{};
  int* p = reinterpret_cast<int*>(base);
  long o = 0;
  auto take = [&](long n) {
    int* r = p ? p + o : nullptr;
    o += (n + 63) / 64 * 64;
    return r;
  };
  w.ccount = take(ncell);
  w.cstart = take((long)ncell + 1);
  w.cursor = take(ncell);
  w.cell_of = take(N);
  w.cnodes = take(N);
  w.deg = take(N);
  w.bsum = take(scan_blocks((long)(N > ncell ? N : ncell) + 1) + 1);
  w.total64 = reinterpret_cast<long long*>(take(2)); // offsets are 256-byte aligned
  w.cpos = reinterpret_cast<double4*>(take(8l * N));
  w.rows = take((long)kFillLds * N);
  w.ints = o;
  return w;
}

void scan(const int* in, long n, int* out, int* bsum, long long* total64, hipStream_t s) {
  const long nb = scan_blocks(n);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)(nb > 0 ? nb : 1)), dim3(kThreads), 0, s, in, n, out, bsum);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kThreads), 0, s, bsum, (int)(nb > 0 ? nb : 1), out + n,
                     total64);
  if (n > 0) hipLaunchKernelGGL(k_scan_carry, dim3(nblk(n)), dim3(kThreads), 0, s, out, n, bsum);
}

int grid_rows(int M, int grid_cap) {
  const int nt = (M + kTileRows - 1) / kTileRows;
  return grid_cap > 0 && grid_cap < nt ? grid_cap : nt;
}

// The persistent dense kernel takes ungathered inputs of width 32, 128 or
// 128 + 128 and a 128-wide BF16 output.
bool dense_shape(const gnn_mlp_io* io, const gnn_mlp_weights* wt, const void* out, const float* outf) {
  if (!out || outf || wt->nout != kLat || wt->nout_pad != kLat) return false;
  for (int k = 0; k < io->nseg; ++k)
    if (io->idx[k]) return false;
  if (io->nseg == 1) return io->width[0] == 32 || io->width[0] == kLat;
  return io->nseg == 2 && io->width[0] == kLat && io->width[1] == kLat;
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      return -1;
    cus = n;
  }
  return cus;
}

template <int K, int WAVES>
int launch_dense_k(const DenseFwdArgs& a, hipStream_t s) {
  const int cus = device_cus();
  if (cus <= 0) return -2;
  const long blocks = ((long)a.M + 16l * WAVES - 1) / (16l * WAVES);
  const int grid = blocks < cus ? (int)blocks : cus;
  hipLaunchKernelGGL((k_dense_fwd<K, WAVES>), dim3(grid), dim3(64 * WAVES), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_dense(const gnn_mlp_io* io, const gnn_mlp_weights* wt, const void* residual, void* hidden,
                 void* out, hipStream_t s) {
  DenseFwdArgs a{};
  a.M = io->M;
  a.x0 = reinterpret_cast<const bf16*>(io->src[0]);
  a.x1 = io->nseg == 2 ? reinterpret_cast<const bf16*>(io->src[1]) : nullptr;
  a.W1t = reinterpret_cast<const bf16*>(wt->W1t);
  a.b1 = wt->b1;
  a.W2t = reinterpret_cast<const bf16*>(wt->W2t);
  a.b2 = wt->b2;
  a.res = reinterpret_cast<const bf16*>(residual);
  a.hid = reinterpret_cast<bf16*>(hidden);
  a.out = reinterpret_cast<bf16*>(out);
  const int K = io->nseg == 2 ? 2 * kLat : io->width[0];
  if (K != wt->K) return -1;
  if (K == 32) return launch_dense_k<32, 16>(a, s);
  if (K == kLat) return launch_dense_k<kLat, 16>(a, s);
  return launch_dense_k<2 * kLat, 12>(a, s);
}

int check_io(const gnn_mlp_io* io) {
  if (!io || io->M <= 0 || io->nseg < 1 || io->nseg > 3) return -1;
  for (int k = 0; k < io->nseg; ++k)
    if (!io->src[k] || io->width[k] <= 0 || io->width[k] % 32) return -1;
  return 0;
}

// one block per CU (the weights in LDS, shared by its waves), persistent
template <bool ENC, bool PIN>
int launch_edge(const EdgeFwdArgs& a, int grid_cap, void* stream) {
  const int cus = device_cus();
  if (cus <= 0) return -2;
  const long per_block = 16l * kEdgeEB * kEdgeWaves;
  const long nblocks = ((long)a.E + per_block - 1) / per_block;
  int grid = grid_cap > 0 ? grid_cap : cus;
  if (grid > nblocks) grid = (int)nblocks;
  hipLaunchKernelGGL((k_edge_fwd<kEdgeEB, kEdgeWaves, ENC, PIN>), dim3(grid), dim3(64 * kEdgeWaves), 0,
                     reinterpret_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

bool edge_weights_ok(const gnn_mlp_weights* wt) {
  return wt && wt->K == 3 * kLat && wt->nout == kLat && wt->nout_pad == kLat && wt->W1t && wt->W2t && wt->b1 &&
         wt->b2;
}

} // namespace

extern "C" {

int gnn_abi_version(void) { return GNN_ABI_VERSION; }

int gnn_mlp_forward(const gnn_mlp_io* io, const gnn_mlp_weights* wt, const void* residual,
                    void* hidden, void* out, float* out_f32, int grid_cap, void* stream) {
  if (check_io(io) || !wt || (wt->nout_pad != 16 && wt->nout_pad != 128) || wt->nout > wt->nout_pad)
    return -1;
  if (grid_cap == 0 && dense_shape(io, wt, out, out_f32))
    return launch_dense(io, wt, residual, hidden, out, reinterpret_cast<hipStream_t>(stream));
  FwdArgs a{};
  a.M = io->M;
  a.in = make_segs(io);
  a.K = a.in.start[io->nseg];
  if (a.K != wt->K) return -1;
  a.NO = wt->nout_pad;
  a.nout = wt->nout;
  a.W1t = reinterpret_cast<const bf16*>(wt->W1t);
  a.b1 = wt->b1;
  a.W2t = reinterpret_cast<const bf16*>(wt->W2t);
  a.b2 = wt->b2;
  a.res = reinterpret_cast<const bf16*>(residual);
  a.hid = reinterpret_cast<bf16*>(hidden);
  a.out = reinterpret_cast<bf16*>(out);
  a.outf = out_f32;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 g(grid_rows(io->M, grid_cap)), b(kThreads);
  if (a.NO == 128) hipLaunchKernelGGL(k_mlp_fwd<8>, g, b, 0, s, a);
  else hipLaunchKernelGGL(k_mlp_fwd<1>, g, b, 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_mlp_backward(const gnn_mlp_io* io, const gnn_mlp_weights* wt, const float* dY,
                     const void* hidden, void* dH, float* const dX[3], int accumulate,
                     int grid_cap, void* stream) {
  if (check_io(io) || !wt || !dY || !hidden || !dH) return -1;
  BwdArgs a{};
  a.M = io->M;
  Segs sg = make_segs(io);
  a.K = sg.start[io->nseg];
  if (a.K != wt->K) return -1;
  a.NO = wt->nout_pad;
  a.nout = wt->nout;
  a.dY = dY;
  a.hid = reinterpret_cast<const bf16*>(hidden);
  a.W2 = reinterpret_cast<const bf16*>(wt->W2);
  a.W1 = reinterpret_cast<const bf16*>(wt->W1);
  a.dH = reinterpret_cast<bf16*>(dH);
  for (int k = 0; k < 3; ++k) {
    a.dX[k] = (dX && k < io->nseg) ? dX[k] : nullptr;
    a.width[k] = sg.width[k];
    a.start[k] = sg.start[k];
  }
  a.start[3] = sg.start[3];
  a.accumulate = accumulate;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 g(grid_rows(io->M, grid_cap)), b(kThreads);
  if (a.NO == 128) hipLaunchKernelGGL(k_mlp_bwd<8>, g, b, 0, s, a);
  else hipLaunchKernelGGL(k_mlp_bwd<1>, g, b, 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_weight_grad(const gnn_mlp_io* io, const void* G_bf16, const float* G_f32, int ncols,
                    int ncols_pad, float* part, float* bpart, int nchunks, float* dW, float* db,
                    int accumulate, void* stream) {
  if (check_io(io) || (!G_bf16 && !G_f32) || !part || !bpart || !dW || !db || nchunks <= 0)
    return -1;
  if (ncols_pad % 16 || ncols_pad > kLat || ncols > ncols_pad) return -1;
  WgradArgs a{};
  a.M = io->M;
  a.in = make_segs(io);
  a.K = a.in.start[io->nseg];
  a.NC = ncols_pad;
  a.ncols_g = ncols;
  a.gb = reinterpret_cast<const bf16*>(G_bf16);
  a.gf = G_bf16 ? nullptr : G_f32;
  a.chunk = ((io->M + nchunks - 1) / nchunks + 31) / 32 * 32;
  const int nch = (io->M + a.chunk - 1) / a.chunk;
  a.part = part;
  a.bpart = bpart;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_wgrad, dim3((a.K + 63) / 64, nch), dim3(kThreads), 0, s, a);
  // dW [K][ncols] from the padded partials: reduce the full padded block
  // and let the caller read the real columns (dW is [K][ncols_pad])
  const long cw = (long)a.K * a.NC;
  hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(cw)), dim3(kThreads), 0, s, part, nch, cw, dW, accumulate);
  hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(a.NC)), dim3(kThreads), 0, s, bpart, nch, (long)a.NC, db,
                     accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_segment_sum(const void* x, const int* row, int N, void* agg, void* stream) {
  if (!x || !row || !agg || N <= 0) return -1;
  hipLaunchKernelGGL(k_seg_sum, dim3((N + 15) / 16), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const bf16*>(x), row, N, reinterpret_cast<bf16*>(agg));
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_edge_grad_to_nodes(const float* d_recv, const float* d_send, const int* row, const int* rev,
                           int N, float* dh, void* stream) {
  if (!row || !dh || N <= 0 || (d_send && !rev)) return -1;
  hipLaunchKernelGGL(k_edge_grad_to_nodes, dim3((N + 15) / 16), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), d_recv, d_send, row, rev, N, dh);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_broadcast_add(const float* d_agg, const int* recv, long E, float* d_e, void* stream) {
  if (!d_agg || !recv || !d_e || E < 0) return -1;
  if (E == 0) return 0;
  hipLaunchKernelGGL(k_bcast_add, dim3(nblk(E * 16)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     d_agg, recv, E, d_e);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_adam(float* p, const float* g, float* m, float* v, long n, float lr_t, float b1, float b2,
             float eps, float wd, void* stream) {
  if (!p || !g || !m || !v || n <= 0) return -1;
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     p, g, m, v, n, lr_t, b1, b2, eps, wd);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_cast_weights(const float* p, int K, int N, int NP, void* W, void* Wt, void* stream) {
  if (!p || !W || !Wt || K <= 0 || N <= 0 || NP < N) return -1;
  hipLaunchKernelGGL(k_cast_weights, dim3(nblk((long)K * NP)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), p, K, N, NP, reinterpret_cast<bf16*>(W),
                     reinterpret_cast<bf16*>(Wt));
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_node_features(int N, const double* x, const float* vhist, int H, const double* rad,
                      float v_scale, float r_mean, float r_scale, const float* planes, int n_planes,
                      float dist_scale, void* out, void* stream) {
  if (N <= 0 || !x || !vhist || !rad || !out || H < 0 || n_planes < 0 || 3 * H + 1 + n_planes > 32)
    return -1;
  if (n_planes > 0 && !planes) return -1;
  hipLaunchKernelGGL(k_node_features, dim3(nblk(N)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     N, x, vhist, H, rad, 1.f / v_scale, r_mean, 1.f / r_scale, planes, n_planes,
                     1.f / dist_scale, reinterpret_cast<bf16*>(out));
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_edge_features(long E, const double* x, const int* recv, const int* send, double radius,
                      void* out, void* stream) {
  if (E < 0 || !x || !recv || !send || !out || !(radius > 0.)) return -1;
  if (E == 0) return 0;
  hipLaunchKernelGGL(k_edge_features, dim3(nblk(E)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     E, x, recv, send, 1. / radius, reinterpret_cast<bf16*>(out));
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_integrate(int N, const float* a, float a_scale, const double* gravity, double dt, double* x,
                  float* vhist, int H, void* stream) {
  if (N <= 0 || !a || !gravity || !x || !vhist || H < 1) return -1;
  hipLaunchKernelGGL(k_integrate, dim3(nblk(N)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     N, a, a_scale, gravity, dt, x, vhist, H);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_mse(const float* y, const float* t, long count, float* dY, float* part, void* stream) {
  if (!y || !t || !dY || !part || count <= 0) return -1;
  hipLaunchKernelGGL(k_mse, dim3(nblk(count)), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                     y, t, count, dY, part);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_node_p(int N, const void* h, const gnn_mlp_weights* wt, float* P, void* stream) {
  if (N <= 0 || !h || !wt || !wt->W1t || !P || wt->K != 3 * kLat) return -1;
  NodePArgs a{N, reinterpret_cast<const bf16*>(h), reinterpret_cast<const bf16*>(wt->W1t), P};
  const int blocks = (N + kTileRows - 1) / kTileRows;
  hipLaunchKernelGGL(k_node_p, dim3(blocks), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_edge_mlp_forward(long E, const void* e, const int* recv, const int* send, const float* P,
                         const void* h, const gnn_mlp_weights* wt, void* hidden, void* out, int grid_cap,
                         void* stream) {
  if (E < 0 || E > 0x7fffffffl || !edge_weights_ok(wt)) return -1;
  if (E == 0) return 0;
  if (!e || !recv || !send || !h || !out) return -1;
  EdgeFwdArgs a{};
  a.E = (int)E;
  a.e = reinterpret_cast<const bf16*>(e);
  a.recv = recv;
  a.send = send;
  a.P = P;
  a.h = reinterpret_cast<const bf16*>(h);
  a.W1t = reinterpret_cast<const bf16*>(wt->W1t);
  a.b1 = wt->b1;
  a.W2t = reinterpret_cast<const bf16*>(wt->W2t);
  a.b2 = wt->b2;
  a.hid = reinterpret_cast<bf16*>(hidden);
  a.out = reinterpret_cast<bf16*>(out);
  return P ? launch_edge<false, false>(a, grid_cap, stream) : launch_edge<false, true>(a, grid_cap, stream);
}

int gnn_edge_mlp_forward_encoded(long E, const void* xe, const gnn_mlp_weights* enc, const int* recv,
                                 const int* send, const float* P, const void* h, const gnn_mlp_weights* wt,
                                 void* out, int grid_cap, void* stream) {
  if (E < 0 || E > 0x7fffffffl || !edge_weights_ok(wt) || !enc || enc->K != 32 || enc->nout != kLat ||
      enc->nout_pad != kLat || !enc->W1t || !enc->W2t || !enc->b1 || !enc->b2)
    return -1;
  if (E == 0) return 0;
  if (!xe || !recv || !send || !P || !h || !out) return -1;
  EdgeFwdArgs a{};
  a.E = (int)E;
  a.recv = recv;
  a.send = send;
  a.P = P;
  a.h = reinterpret_cast<const bf16*>(h);
  a.W1t = reinterpret_cast<const bf16*>(wt->W1t);
  a.b1 = wt->b1;
  a.W2t = reinterpret_cast<const bf16*>(wt->W2t);
  a.b2 = wt->b2;
  a.out = reinterpret_cast<bf16*>(out);
  a.xe = reinterpret_cast<const bf16*>(xe);
  a.We1t = reinterpret_cast<const bf16*>(enc->W1t);
  a.be1 = enc->b1;
  a.We2t = reinterpret_cast<const bf16*>(enc->W2t);
  a.be2 = enc->b2;
  return launch_edge<true, false>(a, grid_cap, stream);
}

int gnn_node_layer_forward(int N, const void* h, const void* e, const int* row, const gnn_mlp_weights* wt,
                           void* agg, void* hidden, void* out, const gnn_mlp_weights* next_edge, float* P,
                           void* stream) {
  if (N <= 0 || !h || !e || !row || !wt || !out || wt->K != 2 * kLat || wt->nout != kLat ||
      wt->nout_pad != kLat || !wt->W1t || !wt->W2t || !wt->b1 || !wt->b2)
    return -1;
  if (next_edge && (next_edge->K != 3 * kLat || !next_edge->W1t || !P)) return -1;
  NodeFwdArgs a{};
  a.N = N;
  a.h = reinterpret_cast<const bf16*>(h);
  a.e = reinterpret_cast<const bf16*>(e);
  a.row = row;
  a.W1t = reinterpret_cast<const bf16*>(wt->W1t);
  a.b1 = wt->b1;
  a.W2t = reinterpret_cast<const bf16*>(wt->W2t);
  a.b2 = wt->b2;
  a.agg = reinterpret_cast<bf16*>(agg);
  a.hid = reinterpret_cast<bf16*>(hidden);
  a.out = reinterpret_cast<bf16*>(out);
  a.Wn1t = next_edge ? reinterpret_cast<const bf16*>(next_edge->W1t) : nullptr;
  a.P = next_edge ? P : nullptr;
  const int cus = device_cus();
  if (cus <= 0) return -2;
  constexpr int kWaves = GNN_NODE_WAVES;
  const long blocks = ((long)N + 16l * kWaves - 1) / (16l * kWaves);
  const int grid = blocks < cus ? (int)blocks : cus;
  hipLaunchKernelGGL((k_node_fwd<kWaves, GNN_NODE_AGG>), dim3(grid), dim3(64 * kWaves), 0, reinterpret_cast<hipStream_t>(stream),
                     a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

} // extern "C"

// Workspace of gnn_radius_graph_build_static: one array per GPU, held from
// the module's load until the process ends, shared by every call on that
// GPU (overlapping calls on one GPU race).  In host code `scratch` names a
// host-side stand-in; its device address comes from hipGetSymbolAddress,
// taken once per device with that device current.
static __device__ __align__(256) int scratch[GNN_GRAPH_SCRATCH_BYTES / sizeof(int)];

namespace {

constexpr int kMaxDevices = 64;
void* scratch_dev[kMaxDevices]; // device address of `scratch` on each GPU, once known
std::mutex scratch_mu;

int graph_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return 0;
}

// Count pass of the radius graph over the carved workspace w: bins the
// points, fills row and reads the edge total back into host_total (host
// memory the caller owns), then *n_edges.
int graph_count(int N, const double* x, double radius, const CellGrid& g, const GraphWork& w, int* row,
                long long* host_total, long* n_edges, hipStream_t s) {
  hipLaunchKernelGGL(k_g_zero, dim3(nblk(g.ncell)), dim3(kThreads), 0, s, w.ccount, (long)g.ncell);
  hipLaunchKernelGGL(k_g_zero, dim3(nblk(g.ncell)), dim3(kThreads), 0, s, w.cursor, (long)g.ncell);
  if (N > 0) {
    hipLaunchKernelGGL(k_g_bin, dim3(nblk(N)), dim3(kThreads), 0, s, N, x, g, w.cell_of, w.ccount);
    scan(w.ccount, g.ncell, w.cstart, w.bsum, nullptr, s);
    hipLaunchKernelGGL(k_g_scatter, dim3(nblk(N)), dim3(kThreads), 0, s, N, x, w.cell_of, w.cstart,
                       w.cursor, w.cnodes, w.cpos);
    hipLaunchKernelGGL(k_g_count, dim3(nblk(N)), dim3(kThreads), 0, s, N, g, radius * radius, w.cell_of,
                       w.cstart, w.cnodes, w.cpos, w.deg, w.rows);
  }
  scan(w.deg, N, row, w.bsum, w.total64, s);
  if (hipMemcpyAsync(host_total, w.total64, sizeof(long long), hipMemcpyDeviceToHost, s) != hipSuccess) return -2;
  if (hipStreamSynchronize(s) != hipSuccess) return -2;
  *n_edges = (long)*host_total;
  return *host_total > 0x7fffffffll ? -3 : 0;
}

} // namespace

extern "C" {

long gnn_radius_graph_workspace(int N, const double lo[3], const double hi[3], double radius) {
  CellGrid g;
  if (N < 0 || make_grid(lo, hi, radius, &g)) return -1;
  return carve(nullptr, N, g.ncell).ints * (long)sizeof(int);
}

int gnn_radius_graph_count(int N, const double* x, double radius, const double lo[3], const double hi[3],
                           void* work, long work_bytes, int* row, long* n_edges, void* stream) {
  CellGrid g;
  if (N < 0 || (N > 0 && !x) || !work || !row || !n_edges || make_grid(lo, hi, radius, &g)) return -1;
  const GraphWork w = carve(work, N, g.ncell);
  if (work_bytes < w.ints * (long)sizeof(int)) return -1;
  // the total comes back through a pinned word (one per host thread, kept
  // until the process ends): a direct copy, without the staging a pageable
  // destination takes
  thread_local long long* total = nullptr;
  if (!total && hipHostMalloc(reinterpret_cast<void**>(&total), sizeof(long long), hipHostMallocDefault) != hipSuccess) {
    total = nullptr;
    return -2;
  }
  return graph_count(N, x, radius, g, w, row, total, n_edges, reinterpret_cast<hipStream_t>(stream));
}

int gnn_radius_graph_fill(int N, const double* x, double radius, const double lo[3], const double hi[3],
                          const void* work, long work_bytes, const int* row, long n_edges, int* send,
                          int* recv, int* rev, void* stream) {
  CellGrid g;
  if (N < 0 || n_edges < 0 || n_edges > 0x7fffffffl || !work || !row || make_grid(lo, hi, radius, &g))
    return -1;
  if (n_edges > 0 && (!x || !send || !recv || !rev)) return -1;
  const GraphWork w = carve(const_cast<void*>(work), N, g.ncell);
  if (work_bytes < w.ints * (long)sizeof(int)) return -1;
  if (N == 0 || n_edges == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_g_fill, dim3(nblk(N)), dim3(kThreads), 0, s, N, x, g, radius * radius, w.cell_of,
                     w.cstart, w.cnodes, w.cpos, w.rows, row, (int)n_edges, send, recv);
  hipLaunchKernelGGL(k_g_rev, dim3(nblk(N)), dim3(kThreads), 0, s, N, row, (int)n_edges, send, rev);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int gnn_radius_graph_build(int N, const double* x, double radius, const double lo[3], const double hi[3],
                           void* work, long work_bytes, int* row, long capacity, int* send, int* recv, int* rev,
                           long* n_edges, void* stream) {
  if (int r = gnn_radius_graph_count(N, x, radius, lo, hi, work, work_bytes, row, n_edges, stream)) return r;
  if (*n_edges > capacity) return 1; // the caller grows the edge arrays and calls the fill
  return gnn_radius_graph_fill(N, x, radius, lo, hi, work, work_bytes, row, *n_edges, send, recv, rev, stream);
}

int gnn_radius_graph_build_static(int N, const double* x, double radius, const double lo[3],
                                  const double hi[3], int* row, long capacity, int* send, int* recv,
                                  int* rev, long* n_edges, void* stream) {
  CellGrid g;
  if (N < 0 || (N > 0 && !x) || !row || !n_edges || capacity < 0 || make_grid(lo, hi, radius, &g)) return -1;
  const long need = carve(nullptr, N, g.ncell).ints * (long)sizeof(int);
  if (need > GNN_GRAPH_SCRATCH_BYTES) return -4; // the graph is too large for the library's array
  void* work = nullptr;
  if (int r = graph_scratch(&work)) return r;
  const GraphWork w = carve(work, N, g.ncell);
  long long total = 0; // pageable: no pinned word is allocated
  if (int r = graph_count(N, x, radius, g, w, row, &total, n_edges, reinterpret_cast<hipStream_t>(stream)))
    return r;
  if (*n_edges > capacity) return 1;
  return gnn_radius_graph_fill(N, x, radius, lo, hi, work, need, row, *n_edges, send, recv, rev, stream);
}

} // extern "C"
