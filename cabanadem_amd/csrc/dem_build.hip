// dem_build.hip -- neighbour-list rebuild on the GPU.
//
// Replaces Cabana::VerletList<FullNeighborTag>::build as called at
// src/CabanaDEM_Solver.hpp:60-63,87-88 (reference paths relative to the
// ExaPhysics/CabanaDEM root).  The list is the set {j != i : dx^2+dy^2+dz^2
// <= rc^2} with the squared distance summed in x, y, z order without fused
// multiply-adds, exactly as the reference's Cabana build forms it, so the
// sets are bit-exact.
//
// Sequence (one rebuild, all stream ordered):
//   1. k_cell_keys   cell key of every particle (tiled cell order, see
//                    dem_internal.h); per-cell arrival rank by atomic
//                    counting; optional skin-displacement check
//   2. scan          exclusive prefix over the cell counts -> cell starts
//   3. k_scatter     counting-sort scatter: slot = start[key] + rank
//   4. k_cell_order  within each cell, order by previous slot (the arrival
//                    order of step 1 is not deterministic; this makes the
//                    sort a stable, deterministic single-digit radix sort
//                    on the cell key) and write the inverse permutation
//   5. k_permute     gather the particle state into the new order and
//                    re-map the tracked-contact lists (j -> inv[j], re-sorted)
//   6. k_neighbors   27-cell search: for each (dz, dy) the three x-adjacent
//                    cells are one contiguous slot range; the nine ranges are
//                    visited in ascending slot order, so a row comes out in
//                    ascending j (the order the force kernel merges in).
#include "dem_internal.h"

#include <algorithm>
#include <climits>

namespace demhip {
namespace {

__device__ inline int cell_coord(double v, const Grid& g, int d) {
  double t = floor((v - g.lo[d]) / g.h[d]);
  if (!(t >= 0.0)) return 0; // below the box, or NaN
  if (t > (double)(g.nc[d] - 1)) return g.nc[d] - 1;
  return (int)t;
}

__global__ __launch_bounds__(kBlock) void k_fill_int(int* p, long count, int value) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = value;
}

// Per-cell arrival ranks with one atomic per run of equal keys in a wave:
// the slots are in the previous rebuild's cell order, so a wave's keys come
// in runs.  The run's first lane reserves the run's length and the others
// take consecutive ranks after it (any ranks that number each cell's
// particles 0 .. count - 1 will do: k_cell_order fixes the order).
__global__ __launch_bounds__(kBlock) void k_cell_keys(BuildArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool active = i < a.n; // the active lanes of a wave are a prefix
  int key = -1;
  double4 p = make_double4(0., 0., 0., 0.);
  if (active) {
    p = a.xr_in[i];
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicOr(&a.flags->nonfinite, 1);
    // ensembles: the slot's system gives the grid and the first key
    const int sys = a.sysid ? a.sysid[i] : 0;
    const Grid& g = a.sysid ? a.grids[sys] : a.grid;
    const long base = a.sysid ? a.kbase[sys] : 0;
    const int cx = cell_coord(p.x, g, 0);
    const int cy = cell_coord(p.y, g, 1);
    const int cz = cell_coord(p.z, g, 2);
    key = (int)(base + cell_key(g, cx, cy, cz));
  }
  const int prev = __shfl_up(key, 1);
  const bool head = active && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long below = lane == 63 ? ~0ull : ((2ull << lane) - 1); // lanes 0 .. lane
  const int hl = 63 - __clzll(heads & below); // this lane's run head (active lanes)
  const int nactive = __popcll(__ballot(active)); // with every lane of the wave executing
  int base_rank = 0;
  if (head) {
    const unsigned long long above = heads & ~below;
    const int next = above ? __ffsll((long long)above) - 1 : nactive;
    base_rank = atomicAdd(&a.cell_count[key], next - lane);
  }
  base_rank = __shfl(base_rank, hl >= 0 ? hl : 0);
  if (!active) return;
  a.key[i] = key;
  a.rank[i] = base_rank + (lane - hl);
  if (a.check_skin) {
    const double d0 = p.x - a.xb[3 * i + 0];
    const double d1 = p.y - a.xb[3 * i + 1];
    const double d2 = p.z - a.xb[3 * i + 2];
    if (d0 * d0 + d1 * d1 + d2 * d2 > a.skin_half_sq) atomicOr(&a.flags->skin, 1);
  }
}

// ---- exclusive scan of int counts (three launches) -------------------------
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int block_exclusive_scan(int v, int* lds, int* total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int add = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  const int incl = lds[t];
  *total = lds[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const int* in, int* out,
                                                              int* tile_sums, long count) {
  __shared__ int lds[kScanThreads];
  const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < count) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_exclusive_scan(s, lds, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < count) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_top(int* tile_sums, int ntiles,
                                                            int* out, long count) {
  __shared__ int lds[kScanThreads];
  int carry = 0;
  for (int base = 0; base < ntiles; base += kScanThreads) {
    const int t = base + (int)threadIdx.x;
    const int v = (t < ntiles) ? tile_sums[t] : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, &total);
    if (t < ntiles) tile_sums[t] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) out[count] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(int* out, const int* tile_sums, long count) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] += tile_sums[i / kScanTile];
}

hipError_t exclusive_scan(const int* in, int* out, int* tmp, long count, hipStream_t s) {
  const int ntiles = (int)((count + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kScanThreads), 0, s, in, out, tmp, count);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tmp, ntiles, out, count);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, out, tmp, count);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_scatter(BuildArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  a.perm[a.cell_start[a.key[i]] + a.rank[i]] = i;
}

// One thread per cell: insertion-sort the cell's slots by previous index
// (cells hold a handful of particles), then record old -> new.
__global__ __launch_bounds__(kBlock) void k_cell_order(BuildArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= a.ncell) return;
  const int s = a.cell_start[c], e = a.cell_start[c + 1];
  if (e - s > kMaxCellParticles) {
    atomicOr(&a.flags->cell_overflow, 1);
    for (int q = s; q < e; ++q) a.inv[a.perm[q]] = q;
    return;
  }
  for (int q = s + 1; q < e; ++q) {
    const int v = a.perm[q];
    int r = q - 1;
    while (r >= s && a.perm[r] > v) {
      a.perm[r + 1] = a.perm[r];
      --r;
    }
    a.perm[r + 1] = v;
  }
  for (int q = s; q < e; ++q) a.inv[a.perm[q]] = q;
}

__device__ inline void copy6(const double* src, double* dst) {
  const double2* s = reinterpret_cast<const double2*>(src);
  double2* d = reinterpret_cast<double2*>(dst);
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
}

// Gather the per-particle arrays into the new order.  The tracked-contact
// list of a particle keeps its entries and histories; only the neighbour
// indices change (j -> inv[j]) and the list is re-sorted by the new index.
__global__ __launch_bounds__(kBlock) void k_permute(BuildArgs a) {
  extern __shared__ int s_j[]; // [max_c][kBlock]
  const int tid = threadIdx.x;
  const int p = blockIdx.x * kBlock + tid;
  if (p >= a.n) return;
  const int o = a.perm[p];
  const double4 xr = a.xr_in[o];
  a.xr_out[p] = xr;
  copy6(a.uw_in + 6l * o, a.uw_out + 6l * p);
  if (a.permute_ft) copy6(a.ft_in + 6l * o, a.ft_out + 6l * p);
  a.mat_out[p] = a.mat_in[o];
  a.mI_out[p] = a.mI_in[o];
  a.id_out[p] = a.id_in[o];
  a.key_sorted[p] = a.key[o];
  if (a.keep_xb) {
    a.xb[3 * p + 0] = xr.x;
    a.xb[3 * p + 1] = xr.y;
    a.xb[3 * p + 2] = xr.z;
  }
  const int meta = a.cmeta_in[o];
  a.cmeta_out[p] = meta;
  const int wm = meta_walls(meta);
  if (wm) {
    for (int w = 0; w < DEM_MAX_WALLS; ++w)
      if (wm & (1 << w))
        for (int c = 0; c < 3; ++c)
          a.sw_out[(long)(w * 3 + c) * a.ld + p] = a.sw_in[(long)(w * 3 + c) * a.ld + o];
  }
  if (a.bonds_on) {
    // bond rows keep their slot order; partners follow the permutation
    const int nb = a.nbond_in[o];
    a.nbond_out[p] = nb;
    a.bEG_out[p] = a.bEG_in[o];
    if (a.vgroup_in) a.vgroup_out[p] = a.vgroup_in[o];
    for (int b = 0; b < nb; ++b) {
      const int j = a.bj_in[(long)b * a.ld + o];
      a.bj_out[(long)b * a.ld + p] = j >= 0 ? a.inv[j] : -1;
      a.bl_out[(long)b * a.ld + p] = a.bl_in[(long)b * a.ld + o];
      for (int c = 0; c < 9; ++c)
        a.bh_out[(long)(9 * b + c) * a.ld + p] = a.bh_in[(long)(9 * b + c) * a.ld + o];
    }
  } else if (a.vgroup_in) {
    a.vgroup_out[p] = a.vgroup_in[o];
  }
  if (a.fext_in)
    for (int d = 0; d < 3; ++d) a.fext_out[3l * p + d] = a.fext_in[3l * o + d];
  int oc = meta_count(meta);
  if (oc > a.max_c) oc = a.max_c; // counts are clamped on write; never index past the LDS rows
  for (int m = 0; m < oc; ++m) s_j[m * kBlock + tid] = a.inv[a.cj_in[(long)m * a.ld + o]];
  for (int m = 0; m < oc; ++m) {
    const int jm = s_j[m * kBlock + tid];
    int r = 0;
    for (int q = 0; q < oc; ++q) r += (s_j[q * kBlock + tid] < jm) ? 1 : 0;
    a.cj_out[(long)r * a.ld + p] = jm;
    for (int c = 0; c < 3; ++c)
      a.ch_out[(long)(r * 3 + c) * a.ld + p] = a.ch_in[(long)(m * 3 + c) * a.ld + o];
  }
}

__device__ inline int wave_max_int(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

__device__ inline int wave_min_int(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// A wave's own LDS tile: its writes before its later reads, and back.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef DEM_NBR_STAGE
#define DEM_NBR_STAGE 128
#endif
constexpr int kNbrStage = DEM_NBR_STAGE; // slots of one range round staged in LDS per wave

// Cabana's full Verlet list (VerletList with FullNeighborTag, built with the
// cell list, src/CabanaDEM_Solver.hpp:60-63): j != i with |x_i - x_j|^2 <= rc^2,
// rows ascending in j.  Slot p searches the 27 cells around its own as nine
// slot ranges (three consecutive x-cells of one (y, z) row are contiguous in
// cell order), walked in ascending order.
//
// LDS-staged cell tiles: in round q every lane walks its q-th range.  The
// lanes of a wave are consecutive slots, mostly of one cell row, so their
// q-th ranges lie in one short span of slots; the wave loads that span into
// LDS once (coalesced) when it holds at most kNbrStage slots, and the lanes
// test against LDS instead of each gathering its own candidates.  Wider
// spans (a wave across a y-tile or system boundary) read from global memory.
// Either way the tested values and their order are the same.
__global__ __launch_bounds__(kBlock) void k_neighbors(BuildArgs a) {
  __shared__ double4 sx[kBlock / 64][kNbrStage];
  const int p = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool active = p < a.n && !a.flags->cell_overflow;
  int cnt = 0;
  long off = p; // a.nbr offset of the row's next entry: cnt * ld + p
  int rs[9], re[9];
  double4 xi = make_double4(0., 0., 0., 0.);
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    rs[q] = INT_MAX;
    re[q] = INT_MAX;
  }
  if (active) {
    const int sys = a.sysid ? a.sysid[p] : 0;
    const Grid& g = a.sysid ? a.grids[sys] : a.grid;
    const long base = a.sysid ? a.kbase[sys] : 0;
    const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
    int key = (int)(a.key_sorted[p] - base);
    const int cx = key % nx;
    key /= nx;
    const int ly = key % g.tr;
    key /= g.tr;
    const int cz = key % nz;
    const int cy = (key / nz) * g.tr + ly;
    xi = a.xr_out[p];
    const int xlo = cx > 0 ? cx - 1 : 0;
    const int xhi = cx < nx - 1 ? cx + 1 : nx - 1;
    // the nine slot ranges, ascending by start (they never overlap)
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int z = cz + q / 3 - 1, y = cy + q % 3 - 1;
      if (z >= 0 && z < nz && y >= 0 && y < ny) {
        const long k0 = base + cell_key(g, xlo, y, z);
        rs[q] = a.cell_start[k0];
        re[q] = a.cell_start[k0 + (xhi - xlo) + 1];
      }
    }
#pragma unroll
    for (int q = 1; q < 9; ++q)
#pragma unroll
      for (int r = q; r > 0; --r)
        if (rs[r] < rs[r - 1]) {
          int t = rs[r]; rs[r] = rs[r - 1]; rs[r - 1] = t;
          t = re[r]; re[r] = re[r - 1]; re[r - 1] = t;
        }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int s = rs[q], e = re[q];
    const bool has = s != INT_MAX;
    const int lo = wave_min_int(has ? s : INT_MAX);
    const int hi = wave_max_int(has ? e : INT_MIN);
    if (lo == INT_MAX) break; // no lane has a q-th range, nor a later one
    // the same for the whole wave; readfirstlane lets the compiler branch on it
    const bool staged = __builtin_amdgcn_readfirstlane(hi - lo <= kNbrStage ? 1 : 0) != 0;
    auto test = [&](int j, const double4& xj) {
      const double d0 = xi.x - xj.x;
      const double d1 = xi.y - xj.y;
      const double d2 = xi.z - xj.z;
      const double dsqr = d0 * d0 + d1 * d1 + d2 * d2;
      if (dsqr <= a.rc2) {
        if (cnt < a.max_n) a.nbr[off] = j;
        ++cnt;
        off += a.ld;
      }
    };
    if (staged) {
      for (int t = lane; t < hi - lo; t += 64) sx[w][t] = a.xr_out[lo + t];
      wave_lds_sync();
      if (has)
        for (int j = s; j < e; ++j)
          if (j != p) test(j, sx[w][j - lo]);
      wave_lds_sync(); // the next round rewrites the tile
    } else if (has) {
      for (int j = s; j < e; ++j)
        if (j != p) test(j, a.xr_out[j]);
    }
  }
  if (active) a.nn[p] = cnt < a.max_n ? cnt : a.max_n;
  // one global atomic per wave, and only when it can raise the maximum
  const int wmax = wave_max_int(cnt);
  if (lane == 0 && wmax > 0) {
    if (wmax > *(volatile int*)&a.flags->max_nn) atomicMax(&a.flags->max_nn, wmax);
    if (wmax > a.max_n) atomicMax(&a.flags->nbr_overflow, wmax);
  }
}

// src/CabanaBondedDEM_Particles.hpp:118-137: bond i to every list neighbour
// j with r_ij < factor * r_i (the reference hard-codes factor 3), storing the
// partner and the initial length in the next slot.  Rows are ascending in j.
__global__ __launch_bounds__(kBlock) void k_setup_bonds(BondSetupArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const double4 xi = a.xr[i];
  const double cond = a.factor * xi.w;                 // :130
  int nb = 0;
  for (int k = 0; k < a.nn[i]; ++k) {
    const int j = a.nbr[(long)k * a.ld + i];
    const double4 xj = a.xr[j];
    const double p0 = xi.x - xj.x, p1 = xi.y - xj.y, p2 = xi.z - xj.z; // :120-122
    const double r2 = p0 * p0 + p1 * p1 + p2 * p2;                      // :124
    const double rij = sqrt(r2);                                        // :126
    if (rij < cond) {                                                   // :132
      if (nb < a.max_b) {
        a.bj[(long)nb * a.ld + i] = j;                                  // :133
        a.bl[(long)nb * a.ld + i] = rij;                                // :134
        for (int c = 0; c < 9; ++c) a.bh[(long)(9 * nb + c) * a.ld + i] = 0.;
      } else {
        atomicMax(&a.flags->bond_overflow, nb + 1);
      }
      ++nb;                                                             // :135
    }
  }
  a.nbond[i] = nb < a.max_b ? nb : a.max_b;
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0ull;
}

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

} // namespace

hipError_t launch_zero_u64(unsigned long long* p, int count, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, count);
  return hipGetLastError();
}

hipError_t launch_setup_bonds(const BondSetupArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_setup_bonds, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, long count, int value, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  long blocks = (count + kBlock - 1) / kBlock;
  if (blocks > 65535L * 8) blocks = 65535L * 8;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, count, value);
  return hipGetLastError();
}

hipError_t launch_rebuild(const BuildArgs& a, hipStream_t s) {
  hipError_t e = launch_fill_int(&a.flags->max_nn, 1, 0, s);
  if (e != hipSuccess) return e;
  e = launch_fill_int(a.cell_count, a.ncell, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cell_keys, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = exclusive_scan(a.cell_count, a.cell_start, a.scan_tmp, a.ncell, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_scatter, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_cell_order, dim3(nblk(a.ncell)), dim3(kBlock), 0, s, a);
  const size_t lds = (size_t)(a.max_c > 0 ? a.max_c : 1) * kBlock * sizeof(int);
  hipLaunchKernelGGL(k_permute, dim3(nblk(a.n)), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_neighbors(const BuildArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_neighbors, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

} // namespace demhip
