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
    for (int k0 = 0; k0 < a.K; k0 += 32) {
      bf16x8 av[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) av[rb] = seg_load(a.in, rows[rb], k0 + lk);
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        const bf16x8 bv = ld8(a.W1t + (long)(cb * 16 + lr) * a.K + k0 + lk);
        acc[0][cb] = mfma(av[0], bv, acc[0][cb]);
        acc[1][cb] = mfma(av[1], bv, acc[1][cb]);
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
          const bf16 hv = (bf16)v;
          sh[w][lrow][col] = hv;
          const int grow = r0 + lrow;
          if (a.hid && grow < a.M) a.hid[(long)grow * kLat + col] = hv;
        }
      }
    __syncthreads();
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

inline unsigned nblk(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int grid_rows(int M, int grid_cap) {
  const int nt = (M + kTileRows - 1) / kTileRows;
  return grid_cap > 0 && grid_cap < nt ? grid_cap : nt;
}

int check_io(const gnn_mlp_io* io) {
  if (!io || io->M <= 0 || io->nseg < 1 || io->nseg > 3) return -1;
  for (int k = 0; k < io->nseg; ++k)
    if (!io->src[k] || io->width[k] <= 0 || io->width[k] % 32) return -1;
  return 0;
}

} // namespace

extern "C" {

int gnn_abi_version(void) { return GNN_ABI_VERSION; }

int gnn_mlp_forward(const gnn_mlp_io* io, const gnn_mlp_weights* wt, const void* residual,
                    void* hidden, void* out, float* out_f32, int grid_cap, void* stream) {
  if (check_io(io) || !wt || (wt->nout_pad != 16 && wt->nout_pad != 128) || wt->nout > wt->nout_pad)
    return -1;
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

} // extern "C"
