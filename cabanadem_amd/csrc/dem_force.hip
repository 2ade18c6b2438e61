// dem_force.hip -- the force-kernel launchers for tiles of 128 slots (and
// the bonded kernels), the dispatch over tile sizes, and the small
// per-particle kernels: force zeroing, kinematics upload and download,
// contact-list update, stand-alone Verlet stages and the diagnostics.
// The force kernels themselves live in dem_force_kernels.h.
#include "dem_force_kernels.h"

namespace demhip {
namespace {

// F = m g + f_ext, T = 0  (src/CabanaDEM_Force.hpp:39-47 plus the body and
// external forces).
__global__ __launch_bounds__(kBlock) void k_zero_forces(StepArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const double2 mI = a.mI[i];
  double ft[6] = {0., 0., 0., 0., 0., 0.};
  initial_force(a, i, mI.x, ft);
  store6(a.ft + 6l * i, ft);
}

// Positions and velocities from device arrays in the reference order (slot
// p holds particle id[p]; either array may be null); with clear, spins,
// forces, torques and every contact history start from zero (meta 0: no
// tracked contact, no wall bit).
__global__ __launch_bounds__(kBlock) void k_set_kinematics(int n, const int* id, const double* x,
                                                           const double* u, int clear, double4* xr,
                                                           double* uw, double* ft, int* cmeta) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const long I = id[p];
  if (x) {
    double4 v = xr[p];
    v.x = x[3 * I];
    v.y = x[3 * I + 1];
    v.z = x[3 * I + 2];
    xr[p] = v;
  }
  if (u)
    for (int d = 0; d < 3; ++d) uw[6l * p + d] = u[3 * I + d];
  if (clear) {
    for (int d = 0; d < 3; ++d) uw[6l * p + 3 + d] = 0.;
    for (int c = 0; c < 6; ++c) ft[6l * p + c] = 0.;
    cmeta[p] = make_meta(0, 0);
  }
}

// Positions and velocities to device arrays in the reference order.
__global__ __launch_bounds__(kBlock) void k_get_kinematics(int n, const int* id, const double4* xr,
                                                           const double* uw, double* x, double* u) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const long I = id[p];
  if (x) {
    const double4 v = xr[p];
    x[3 * I] = v.x;
    x[3 * I + 1] = v.y;
    x[3 * I + 2] = v.z;
  }
  if (u)
    for (int d = 0; d < 3; ++d) u[3 * I + d] = uw[6l * p + d];
}

// src/CabanaDEM_Force.hpp:343-445: keep the tracked contacts that still
// overlap.  The list stays ascending in j; the reference's swap-remove slot
// order is not kept (the history is compared as a map keyed by (i, j)).
__global__ __launch_bounds__(kBlock) void k_update_tangential(StepArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const long ld = a.ld;
  const double4 xi = a.xr_in[i];
  const int meta = a.cmeta_in[i];
  const int oc = meta_count(meta);
  int m = 0;
  for (int q = 0; q < oc; ++q) {
    const int j = a.cj_in[q * ld + i];
    const double4 xj = a.xr_in[j];
    const double p0 = xi.x - xj.x, p1 = xi.y - xj.y, p2 = xi.z - xj.z; // :388-390
    const double r2 = p0 * p0 + p1 * p1 + p2 * p2;
    const double overlap = xi.w + xj.w - sqrt(r2);                    // :397
    if (overlap <= 0.) continue;                                      // :399
    a.cj_out[m * ld + i] = j;
    for (int c = 0; c < 3; ++c) a.ch_out[(3 * m + c) * ld + i] = a.ch_in[(3 * q + c) * ld + i];
    ++m;
  }
  a.cmeta_out[i] = make_meta(m, meta_walls(meta));
}

// Half kick (stage1/stage3) and/or drift (stage2), in place.
__global__ __launch_bounds__(kBlock) void k_kick(StepArgs a, double4* xr, double* uw,
                                                 const double* ftp, int do_kick, int do_drift) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  double s[6];
  load6(uw + 6l * i, s);
  if (do_kick) {
    double ft[6];
    load6(ftp + 6l * i, ft);
    const double2 mI = a.mI[i];
    const double minv = 1. / mI.x, Iinv = 1. / mI.y; // src/CabanaDEM_Integrator.hpp:42,48
    for (int d = 0; d < 3; ++d) s[d] = s[d] + a.hdt * ft[d] * minv;
    for (int d = 0; d < 3; ++d) s[3 + d] = s[3 + d] + a.hdt * ft[3 + d] * Iinv;
    apply_vgroup(a, i, s);
    store6(uw + 6l * i, s);
  }
  if (do_drift) {
    double4 x = xr[i];
    x.x = x.x + a.dt * s[0];
    x.y = x.y + a.dt * s[1];
    x.z = x.z + a.dt * s[2];
    xr[i] = x;
  }
}

template <int B>
__global__ __launch_bounds__(B) void k_diag_current(StepArgs a, const double* uw) {
  const int tile = blockIdx.x;
  int first = tile * B, end = a.n;
  if (a.tiles) {
    const int4 t = a.tiles[tile];
    first = t.y; end = t.z;
  }
  const int i = first + threadIdx.x;
  double dg[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  if (i < end) {
    double s[6];
    load6(uw + 6l * i, s);
    const double m = a.mI[i].x, I = a.mI[i].y;
    dg[0] = 0.5 * m * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    dg[1] = 0.5 * I * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    dg[2] = m * s[0]; dg[3] = m * s[1]; dg[4] = m * s[2];
    dg[5] = (double)a.nn[i];
    dg[6] = (double)meta_count(a.cmeta_in[i]);
    dg[7] = (double)__popc(meta_walls(a.cmeta_in[i]));
  }
  block_reduce8<B>(dg, a.diag_part, tile);
}

// One workgroup per system: its tiles' partials summed in tile order.
__global__ __launch_bounds__(kBlock) void k_diag_systems(const double* part, const int* tile0,
                                                         double* out) {
  const int s = blockIdx.x;
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int b = tile0[s] + threadIdx.x; b < tile0[s + 1]; b += kBlock)
    for (int k = 0; k < 8; ++k) v[k] += part[(long)b * 8 + k];
  block_reduce8<kBlock>(v, out, s);
}

// Sum block partials in block order; row = (step, 8 sums).
// The per-tile partials summed in two fixed-order stages (the result does
// not depend on timing): workgroup g sums tiles [g chunk, (g + 1) chunk)
// into group partial g, then one workgroup sums the groups into the row.
__global__ __launch_bounds__(kDiagThreads) void k_diag_groups(const double* part, int nblocks, int chunk,
                                                              double* gpart) {
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  const long b0 = (long)blockIdx.x * chunk;
  const long b1 = b0 + chunk < nblocks ? b0 + chunk : nblocks;
  for (long b = b0 + threadIdx.x; b < b1; b += kDiagThreads)
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] += part[b * 8 + k];
  block_reduce8<kDiagThreads>(v, gpart, blockIdx.x);
}

__global__ __launch_bounds__(kDiagThreads) void k_diag_final(const double* gpart, int ngroups, double* row,
                                                             double step) {
  double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  if ((int)threadIdx.x < ngroups)
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = gpart[threadIdx.x * 8 + k];
  __shared__ double tmp[8];
  block_reduce8<kDiagThreads>(v, tmp, 0);
  __syncthreads();
  if (threadIdx.x == 0) {
    row[0] = step;
    for (int k = 0; k < 8; ++k) row[1 + k] = tmp[k];
  }
}

} // namespace

hipError_t launch_force_b128(bool ref, bool bond, bool ens, ForceKind kind, const StepArgs& a,
                             bool sample, int grid_blocks, hipStream_t s) {
  if (ens) return ref ? launch_force_mode<true, false, true, 128>(kind, a, sample, grid_blocks, s)
                      : launch_force_mode<false, false, true, 128>(kind, a, sample, grid_blocks, s);
  if (bond) return ref ? launch_force_mode<true, true, false, 128>(kind, a, sample, grid_blocks, s)
                       : launch_force_mode<false, true, false, 128>(kind, a, sample, grid_blocks, s);
  return ref ? launch_force_mode<true, false, false, 128>(kind, a, sample, grid_blocks, s)
             : launch_force_mode<false, false, false, 128>(kind, a, sample, grid_blocks, s);
}



hipError_t launch_force(ForceKind kind, const StepArgs& a, bool sample, int grid_blocks,
                        hipStream_t s) {
  const bool ref = a.physics_mode == DEM_MODE_REFERENCE;
  const bool ens = a.tiles != nullptr; // ensembles have no bonds
  const bool bond = a.nbond != nullptr;
  if (ens && bond) return hipErrorInvalidValue;
  if (bond && a.block != kBlock) return hipErrorInvalidValue;
  switch (a.block) {
    case 64: return launch_force_b64(ref, ens, kind, a, sample, grid_blocks, s);
    case 128: return launch_force_b128(ref, bond, ens, kind, a, sample, grid_blocks, s);
    case 256: return launch_force_b256(ref, ens, kind, a, sample, grid_blocks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zero_forces(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_forces, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_set_kinematics(int n, const int* id, const double* x, const double* u, int clear,
                                 double4* xr, double* uw, double* ft, int* cmeta, hipStream_t s) {
  hipLaunchKernelGGL(k_set_kinematics, dim3(nblk(n)), dim3(kBlock), 0, s, n, id, x, u, clear, xr, uw, ft,
                     cmeta);
  return hipGetLastError();
}

hipError_t launch_get_kinematics(int n, const int* id, const double4* xr, const double* uw, double* x,
                                 double* u, hipStream_t s) {
  hipLaunchKernelGGL(k_get_kinematics, dim3(nblk(n)), dim3(kBlock), 0, s, n, id, xr, uw, x, u);
  return hipGetLastError();
}

hipError_t launch_update_tangential(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_update_tangential, dim3(nblk(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_kick(const StepArgs& a, double4* xr, double* uw, const double* ft,
                       int do_kick, int do_drift, hipStream_t s) {
  hipLaunchKernelGGL(k_kick, dim3(nblk(a.n)), dim3(kBlock), 0, s, a, xr, uw, ft, do_kick, do_drift);
  return hipGetLastError();
}

hipError_t launch_diag_current(const StepArgs& a, const double* uw, hipStream_t s) {
  switch (a.block) {
    case 64: hipLaunchKernelGGL(k_diag_current<64>, dim3(a.ntiles), dim3(64), 0, s, a, uw); break;
    case 128: hipLaunchKernelGGL(k_diag_current<128>, dim3(a.ntiles), dim3(128), 0, s, a, uw); break;
    case 256: hipLaunchKernelGGL(k_diag_current<256>, dim3(a.ntiles), dim3(256), 0, s, a, uw); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_diag_systems(const double* part, const int* tile0, int nsys, double* out,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_diag_systems, dim3(nsys), dim3(kBlock), 0, s, part, tile0, out);
  return hipGetLastError();
}

hipError_t launch_diag_final(double* part, int nblocks, double* row, double step, hipStream_t s) {
  // the group partials follow the tile partials ([kDiagThreads][8] more)
  const int chunk = nblocks > kDiagThreads ? (nblocks + kDiagThreads - 1) / kDiagThreads : 1;
  const int ngroups = nblocks > 0 ? (nblocks + chunk - 1) / chunk : 1;
  double* gpart = part + (long)(nblocks > 0 ? nblocks : 0) * 8;
  hipLaunchKernelGGL(k_diag_groups, dim3(ngroups), dim3(kDiagThreads), 0, s, (const double*)part, nblocks,
                     chunk, gpart);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kDiagThreads), 0, s, (const double*)gpart, ngroups, row, step);
  return hipGetLastError();
}

} // namespace demhip
