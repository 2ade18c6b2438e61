// dem_internal.h -- device data layout and kernel launchers of the MI355X
// DEM time-step library.  Not part of the public ABI (include/cabanadem.h).
//
// Layout in HBM (all FP64 state stays resident; particles are stored in
// cell order and re-sorted at every neighbour rebuild):
//   xr    double4[n]        x, y, z, radius               (gathered for neighbours)
//   uw    double[n][6]      u, omega                      (gathered for contacts)
//   mI    double2[n]        m, I  (1/m and 1/I are formed per use, as the
//                           reference does, src/CabanaDEM_Integrator.hpp:42,48)
//   mat   double4[n]        (1-nu^2)/E, (2-nu)/G, 1/r, 1/m (read only on contact)
//   ft    double[n][6]      F, T
//   id    int[n]            original particle id of each sorted slot
//   nn    int[n], nbr int[max_n][ld]
//                           neighbour rows, slot-major (lane-coalesced), ascending j
//   cmeta int[n]            tracked-contact count (bits 0-7) | wall-contact bits (8-15)
//   cj    int[max_c][ld], ch double[max_c*3][ld]
//                           tracked pp contacts: compact, ascending j, with the
//                           tangential history of each
//   sw    double[DEM_MAX_WALLS*3][ld]  wall tangential history
// Cell order: cells are numbered in y-tiles of Grid::tr full x-rows,
// z-major inside a tile, so the z-neighbours of a particle sit a tile's
// area of cells away (tens of kB) instead of a whole x-y layer; the
// neighbour gathers then hit L2.
// Arrays written by the fused step kernel or by the re-sort are double
// buffered ([2]) and swapped by the host.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/cabanadem.h"

namespace demhip {

#ifndef DEM_BLOCK
#define DEM_BLOCK 128
#endif
constexpr int kBlock = DEM_BLOCK;  // threads per block of the particle kernels
// Tile (workgroup) sizes of the force kernels of systems without bonds,
// chosen at run time (dem_set_shape); bonded systems run kBlock.
constexpr int kForceBlocks[3] = {64, 128, 256};
static_assert(kForceBlocks[1] == kBlock, "index 1 is the default tile size, the bonded one");
constexpr int kMaxContacts = 32;  // upper bound of max_contacts
constexpr int kMaxNeighbors = 255;
constexpr int kXcds = 8;          // MI355X: 8 XCDs, workgroups dealt round-robin
#ifndef DEM_TILE_ROWS
#define DEM_TILE_ROWS 8
#endif
constexpr int kTileRows = DEM_TILE_ROWS;      // default y-rows per cell tile
constexpr int kMaxTileRows = 32;              // dem_set_shape accepts 1..32
// Workgroup size and number of groups of the diagnostics' second stage; the
// tile partials are followed by kDiagThreads * 8 doubles of group partials.
constexpr int kDiagThreads = 256;

// Timing events skip the system-scope fence (an L2 write-back and
// invalidation per event that the kernels after it would pay for).
#ifndef DEM_TIMING_EVENT_FLAGS
#define DEM_TIMING_EVENT_FLAGS hipEventDisableSystemFence
#endif

// Logical block index that gives each XCD one contiguous run of blocks, so
// that the neighbour gathers of one XCD stay inside its own L2.  Workgroup b
// of a launch goes to XCD b % 8; this maps XCD k to the k-th contiguous
// share of [0, nb).  Bijective for any nb.
__device__ inline int xcd_block(int b, int nb) {
  const int q = nb / kXcds, r = nb % kXcds;
  const int xcd = b % kXcds, local = b / kXcds;
  return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

struct WallConst {
  double x[3];
  double n[3];
  double a1; // (1 - nu_w^2) / E_w
  double a2; // (2 - nu_w) / G_w
  // corrected mode: the plane translates with u (x at wall time 0) and its
  // surface moves with u + w x (c - x_w) at a contact point c
  double u[3];
  double w[3];
};

struct Grid {
  double lo[3];
  double h[3];
  int nc[3];  // cells per dimension
  int nty;    // y-tiles
  int tr;     // y-rows per tile
  int pad;
};

// Per-system constants of the contact model.  A context holds one system
// or an ensemble of independent systems; each system has its own materials,
// walls and binning box, and its particles occupy one contiguous slot range
// that the re-sort never leaves (cell keys are system-major).
struct SysParams {
  double c_eta_pp, c_eta_pw;   // (-2 sqrt(5/6)) * beta of the restitution coefficient
  double fr_pp, fr_pw;
  int n_walls, pad;
  WallConst walls[DEM_MAX_WALLS];
};

// Cell key of cell (cx, cy, cz): tile-major, then z, then y inside the
// tile, then x.  Keys of x-adjacent cells of one row are consecutive.
__host__ __device__ inline long cell_key(const Grid& g, int cx, int cy, int cz) {
  const int ty = cy / g.tr, ly = cy % g.tr;
  return (((long)ty * g.nc[2] + cz) * g.tr + ly) * g.nc[0] + cx;
}
__host__ __device__ inline long cell_key_count(const Grid& g) {
  return (long)g.nty * g.nc[2] * g.tr * g.nc[0];
}

struct DevFlags {
  int nonfinite;
  int nbr_overflow;   // largest row that did not fit
  int max_nn;         // largest row of the last build
  int contact_overflow;
  int skin;
  int bond_overflow;
  int cell_overflow;  // a cell held more than kMaxCellParticles particles
  int bonds_broken;   // bonds broken since the flags were last cleared
};

// Upper bound on particles in one cell; beyond it the binning box is far too
// coarse for the cutoff and the rebuild stops with an error instead of
// running quadratic work.
constexpr int kMaxCellParticles = 4096;

__host__ __device__ inline int meta_count(int m) { return m & 0xFF; }
__host__ __device__ inline int meta_walls(int m) { return (m >> 8) & 0xFF; }
__host__ __device__ inline int make_meta(int count, int walls) { return count | (walls << 8); }

// Pointers and scalars of one launch, resolved for the current buffers.
struct StepArgs {
  int n, ld, max_n, max_c;
  int physics_mode;
  double dt, hdt;
  double g[3];
  int gravity_on;
  const SysParams* sys;  // [n_systems] (ensembles)
  SysParams sys0;        // the single system's parameters, by value: kernel-argument
                         // loads are scalar and can be repeated at will
  // Tiles of `block` slots (the force kernels' workgroup size, one of
  // kForceBlocks), each inside one system: (system, first, end, -).  Null
  // for a single system (tile t = slots [t block, min((t+1) block, n))).
  const int4* tiles;
  int ntiles;
  int block;

  const double4* xr_in;
  double4* xr_out;
  const double* uw_in;
  double* uw_out;
  const double4* mat;
  const double2* mI;
  double* ft;
  double* sw;
  const int* nn;
  const int* nbr;
  int* cmeta_in;   // written in place only by the walls-only kernel
  const int* cj_in;
  const double* ch_in;
  int* cmeta_out;
  int* cj_out;
  double* ch_out;
  DevFlags* flags;
  double* diag_part; // [blocks][8] when sampling

  // bonded DEM (src/CabanaBondedDEM_Force.hpp:52-348), slot-indexed
  int max_b;
  const int* nbond;      // bond slots in use per particle
  int* bj;               // [max_b][ld] partner slot, -1 once broken
  const double* bl;      // [max_b][ld] rest length
  double* bh;            // [9 max_b][ld] ft, mn, mt history per slot
  const double2* bEG;    // E, G per particle (bond stiffness)
  double beta_bond, bond_radius_factor, tensile_strength;
  int unbonded_contacts; // 0: no pp contacts at all in a bonded run
  // particle groups: vgroup[i] = g > 0 selects group g-1 of kind
  //   DEM_GROUP_VELOCITY: u = vgv[g-1], omega = 0 after every half kick
  //   DEM_GROUP_FIXED_FORCE: F = 0 after the force pass (torque kept)
  const int* vgroup;
  int n_vgroups;
  int vgkind[DEM_MAX_GROUPS];
  double vgv[DEM_MAX_GROUPS][3];
  const double* fext;    // [n][3] constant external force, or null
  // 1: whole force passes run split (systems without bonds): k_defer
  // finishes the particles without contacts and lists the others per tile
  // in work, then k_force_list runs those (sampling passes: k_detect, then
  // the force kernel over its contact lists); 0: one fused kernel
  int split;
  int* work;             // [ntiles][block] each tile's particles with contacts (split passes)
  int* work_tile;        // [ntiles] their number per tile
  // 1: the contact arithmetic of all lanes of a tile runs spread over the
  // tile's lanes (systems without bonds); 0: each lane walks its own list
  int wave;
  // time of this force pass since the walls' positions were set (moving
  // walls, corrected mode)
  double tw;
  // the displacement check of the automatic cadence fused into this pass,
  // or null: max |x_out - xb|^2 and max |u_out|^2 over the particles it
  // writes, as bits into disp[0], disp[1] (cleared before the pass)
  const double* xb;      // [n][3] positions at the last build
  unsigned long long* disp;
};

// Maximum that sends NaN to the top (the displacement check's maxima are
// >= 0, so their bits order as the values do).
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

struct BuildArgs {
  int n, ld, max_n, max_c;
  Grid grid;         // the single system's grid (sysid == null)
  // ensembles: system of each slot (static), each system's grid and first key
  const int* sysid;
  const Grid* grids;
  const long* kbase;
  long ncell;        // number of cell keys (padded to whole tiles, all systems)
  double rc2;
  double skin_half_sq;
  int check_skin;    // compare with the positions of the last build
  int keep_xb;       // record the positions of this build
  int permute_ft;    // F and T are live across this rebuild
  // old (cur) and new (nxt) buffers
  const double4* xr_in; double4* xr_out;
  const double* uw_in; double* uw_out;
  const double4* mat_in; double4* mat_out;
  const double2* mI_in; double2* mI_out;
  const int* id_in; int* id_out;
  const double* ft_in; double* ft_out;
  const double* sw_in; double* sw_out;
  const int* cmeta_in; int* cmeta_out;
  const int* cj_in; int* cj_out;
  const double* ch_in; double* ch_out;
  double* xb;   // positions at the last build, sorted order
  int* key;     // cell key, old order
  int* rank;    // arrival rank in the cell, old order
  int* key_sorted;
  int* cell_count;
  int* cell_start;
  int* perm;    // new -> old
  int* inv;     // old -> new
  int* scan_tmp;
  int* nn;
  int* nbr;
  DevFlags* flags;
  // bonded DEM state (permuted with the particles when bonds_on)
  int bonds_on, max_b;
  const int* nbond_in; int* nbond_out;
  const int* bj_in; int* bj_out;
  const double* bl_in; double* bl_out;
  const double* bh_in; double* bh_out;
  const double2* bEG_in; double2* bEG_out;
  const int* vgroup_in; int* vgroup_out;
  const double* fext_in; double* fext_out;
};

// Bond set-up over the current neighbour rows: every j with r_ij < factor r_i.
struct BondSetupArgs {
  int n, ld, max_b;
  double factor;
  const double4* xr;
  const int* nn;
  const int* nbr;
  int* nbond;
  int* bj;
  double* bl;
  double* bh;
  DevFlags* flags;
};
hipError_t launch_setup_bonds(const BondSetupArgs& a, hipStream_t s);

// Launchers (defined in dem_build.hip / dem_force.hip); all enqueue on `s`.
hipError_t launch_rebuild(const BuildArgs& a, hipStream_t s);
hipError_t launch_neighbors(const BuildArgs& a, hipStream_t s);
hipError_t launch_fill_int(int* p, long count, int value, hipStream_t s);
// p[0 .. count) = 0 (count <= 64)
hipError_t launch_zero_u64(unsigned long long* p, int count, hipStream_t s);

enum ForceKind {
  FK_PP_PW_FORCES = 0, // zero + update + pp + pw, write F,T
  FK_PP_ACCUM = 1,     // F,T += pp (contact list updated)
  FK_PW_ACCUM = 2,     // F,T += pw
  FK_STEP_FUSED = 3,   // forces + stage3 + stage1 + stage2 of the next step
  FK_STEP_LAST = 4,    // forces + stage3, write F,T
  FK_BOND_ACCUM = 5    // F,T += bond forces
};
// grid_blocks: workgroups of the launch; 0 = one per tile, otherwise a
// persistent grid that strides over the tiles.
hipError_t launch_force(ForceKind kind, const StepArgs& a, bool sample, int grid_blocks,
                        hipStream_t s);
hipError_t launch_zero_forces(const StepArgs& a, hipStream_t s);
hipError_t launch_update_tangential(const StepArgs& a, hipStream_t s);
hipError_t launch_set_kinematics(int n, const int* id, const double* x, const double* u, int clear,
                                 double4* xr, double* uw, double* ft, int* cmeta, hipStream_t s);
hipError_t launch_get_kinematics(int n, const int* id, const double4* xr, const double* uw, double* x,
                                 double* u, hipStream_t s);
hipError_t launch_kick(const StepArgs& a, double4* xr, double* uw, const double* ft,
                       int do_kick, int do_drift, hipStream_t s);
// Per-tile partial sums of the current state (one partial row per tile).
hipError_t launch_diag_current(const StepArgs& a, const double* uw, hipStream_t s);
// The force kernels for one tile size each (a.block), in translation units
// of their own: systems without bonds at 64 and 256 slots, and everything
// (bonded systems included) at 128.
hipError_t launch_force_b64(bool ref, bool ens, ForceKind kind, const StepArgs& a, bool sample,
                            int grid_blocks, hipStream_t s);
hipError_t launch_force_b128(bool ref, bool bond, bool ens, ForceKind kind, const StepArgs& a,
                             bool sample, int grid_blocks, hipStream_t s);
hipError_t launch_force_b256(bool ref, bool ens, ForceKind kind, const StepArgs& a, bool sample,
                             int grid_blocks, hipStream_t s);
// Sum of the tile partials part[nblocks][8] into out_row (step, then the 8
// sums); part must hold kDiagThreads * 8 more doubles after the tiles'.
hipError_t launch_diag_final(double* part, int nblocks, double* out_row, double step, hipStream_t s);
// Per-system sums of the tile partials: out[s][8] over tiles [tile0[s], tile0[s+1]).
hipError_t launch_diag_systems(const double* part, const int* tile0, int nsys, double* out,
                               hipStream_t s);

} // namespace demhip
