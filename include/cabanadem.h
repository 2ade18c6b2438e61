/*
 * cabanadem.h -- C ABI of the MI355X-native CabanaDEM time-step library
 * (libcabanadem_hip.so, built from cabanadem_amd/csrc).
 *
 * The reference exposes a header-only C++ template API and no C ABI; each
 * entry point below names the reference interface it stands in for
 * (paths relative to the ExaPhysics/CabanaDEM root).  The C++ drop-in headers
 * in this directory (CabanaDEM.hpp and friends) are the reference-shaped API
 * built on top of these calls.
 *
 * Conventions
 *  - Opaque context per simulated system; one host thread per context.
 *  - All calls return 0 on success or a negative DEM_ERR_* code; the text of
 *    the last error is available from dem_last_error().  No exception crosses
 *    the ABI.
 *  - Host arrays are in the reference layout: vectors xyz-interleaved
 *    double[n][3], histories row-major T[n][max] (src/CabanaDEM_Particles.hpp:40-46).
 *    Host pointers are borrowed for the duration of the call only.
 *  - Device work is stream ordered on the context's own stream; calls that
 *    return host data synchronise.  The context owns all device memory.
 *  - physics_mode 0 ("reference") reproduces the reference's arithmetic,
 *    including its quirks; 1 ("corrected") fixes them.  Both move the same
 *    bytes.
 */
#ifndef CABANADEM_H
#define CABANADEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_MAX_WALLS 8
#define DEM_MAX_GROUPS 8
#define DEM_MAX_BONDS 32
#define DEM_ABI_VERSION 5

enum dem_status {
  DEM_OK = 0,
  DEM_ERR_ARG = -1,            /* invalid argument / configuration */
  DEM_ERR_HIP = -2,            /* HIP runtime failure (incl. no device) */
  DEM_ERR_NEIGH_OVERFLOW = -3, /* a neighbour row exceeded max_neighbors */
  DEM_ERR_CONTACT_OVERFLOW = -4, /* more tracked contacts than max_contacts (MaxCnts) */
  DEM_ERR_NONFINITE = -5,      /* non-finite position met by the cell binning */
  DEM_ERR_SKIN = -6,           /* a particle moved more than skin/2 between rebuilds */
  DEM_ERR_BOND_OVERFLOW = -7,  /* more bonds than max_bonds (MaxBonds) */
  DEM_ERR_STATE = -8           /* call not valid in the current state */
};

enum dem_physics_mode { DEM_MODE_REFERENCE = 0, DEM_MODE_CORRECTED = 1 };

/* dem_config.rebuild_every: automatic cadence.  Every few steps the library
 * measures the largest displacement since the last build, d, and the
 * largest speed, v (one small kernel and one synchronisation), and plans the
 * next measurement M steps ahead so that d + 2 v dt (M + 1) <= 0.9 skin/2,
 * with M at most twice the previous interval (one step after a build);
 * when no step fits, the rows are rebuilt before the next force pass.  The
 * rows (cutoff >= 2 r_max + skin) then hold every pair that can touch as
 * long as speeds do not more than double between two measurements.  The
 * rebuild's skin check (DEM_ERR_SKIN) stays on as the guard against
 * anything faster.  Captured graphs are not used with this cadence.  No
 * counterpart in the reference, which rebuilds every step
 * (src/CabanaDEM_Solver.hpp:87-88). */
#define DEM_REBUILD_AUTO (-1)

/* Kinds of particle group (dem_set_groups). */
enum dem_group_kind {
  DEM_GROUP_VELOCITY = 0,   /* u = v_g and omega = 0 after every half kick (pulled faces) */
  DEM_GROUP_FIXED_FORCE = 1 /* F = 0 after the force pass, torque kept
                             * (applyBoundaryConditions,
                             * examples/tst_01_BDEM_two_particles_tensile_3D.cpp:18-66) */
};

typedef struct dem_ctx dem_ctx;

/* Static configuration of one system.  Replaces the constructor arguments
 * spread over Force (src/CabanaDEM_Force.hpp:23-28: cor_pp, cor_pw,
 * friction_pp, friction_pw), Integrator (src/CabanaDEM_Integrator.hpp:22-27:
 * dt), SolverDEM (src/CabanaDEM_Solver.hpp:32-64: delta, mesh_lo/hi,
 * cell_ratio 1.0) and the Particles template capacities MaxCnts/MaxNoWalls
 * (src/CabanaDEM_Particles.hpp:22). */
typedef struct dem_config {
  int n;               /* particles */
  int n_walls;         /* infinite walls, <= DEM_MAX_WALLS */
  int max_neighbors;   /* neighbour-row capacity; 0 = sized from the first build */
  int max_contacts;    /* tracked pp contacts per particle (MaxCnts), <= 32 */
  double dt;
  double cor_pp, cor_pw, friction_pp, friction_pw;
  double gravity[3];   /* body force per unit mass (0 reproduces the reference) */
  double cutoff;       /* neighbour cutoff (the reference's `delta`) */
  double skin;         /* > 0: flag DEM_ERR_SKIN if a particle moves > skin/2 between rebuilds */
  double lo[3], hi[3]; /* binning box (mesh_lo / mesh_hi); outside particles are clamped in */
  int rebuild_every;   /* neighbour rebuild cadence K in steps (1 = every step, as the reference),
                        * or DEM_REBUILD_AUTO: rebuild when the displacement since the last
                        * build may reach skin/2 (needs skin > 0) */
  int physics_mode;    /* DEM_MODE_* */
  int block_size;      /* force-kernel tile size: 0 (128), 64, 128 or 256 (dem_set_shape) */
  int reserved;
} dem_config;

/* One system of an ensemble (dem_create_ensemble): its particle count,
 * wall count, contact materials and binning box.  The other fields of
 * dem_config (dt, cutoff, skin, capacities, gravity, rebuild cadence,
 * physics mode) are common to the ensemble. */
typedef struct dem_system_desc {
  int n;
  int n_walls;
  double cor_pp, cor_pw, friction_pp, friction_pw;
  double lo[3], hi[3];
} dem_system_desc;

/* Host view of the particle state in the reference layout
 * (fields of src/CabanaDEM_Particles.hpp:348-370).  NULL pointers are
 * skipped on upload/download.  max_cnts / max_walls give the row length of
 * the tangential-history arrays (the MaxCnts / MaxNoWalls template
 * parameters). */
typedef struct dem_particles_view {
  int n;
  int max_cnts;
  int max_walls;
  int reserved;
  double *x, *u, *omega, *force, *torque; /* [n][3] */
  double *m, *rad, *E, *nu, *G, *I;       /* [n] */
  double *sw_x, *sw_y, *sw_z;             /* [n][max_walls] wall tangential history */
  double *ss_x, *ss_y, *ss_z;             /* [n][max_cnts]  pp tangential history */
  int *ss_idx;                            /* [n][max_cnts]  tracked neighbour ids, -1 = empty */
  int *ss_cnt;                            /* [n]            tracked contacts */
} dem_particles_view;

/* Host view of the walls (fields of src/CabanaDEM_Wall.hpp:192-201).  The
 * reference reads the wall velocity and spin but leaves them unused
 * (src/CabanaDEM_Force.hpp:468-469,529): in reference mode the walls stay
 * put.  In corrected mode a wall translates with u from the position x it
 * has at upload (one dt per dem_step step), and its surface moves with
 * u + omega x (c - x_w) at a contact point c (a belt, a turntable); the
 * normal does not turn.  u / omega NULL: static. */
typedef struct dem_walls_view {
  int n;
  int reserved;
  const double *x, *normal; /* [n][3] */
  const double *E, *nu, *G; /* [n] */
  const double *u, *omega;  /* [n][3] or NULL */
} dem_walls_view;

/* Bond parameters of a bonded system.  The reference hard-codes them in
 * computeBondedForce (src/CabanaBondedDEM_Force.hpp:186-188: beta 0.1,
 * bond radius factor 1) and in ParticlesBondedDEM's MaxBonds template
 * parameter (src/CabanaBondedDEM_Particles.hpp:29); breakage and contacts
 * between unbonded pairs are extensions. */
typedef struct dem_bond_config {
  int max_bonds;           /* bond slots per particle (MaxBonds), <= DEM_MAX_BONDS */
  int unbonded_contacts;   /* 1: Hertz contacts between overlapping unbonded pairs;
                            * 0: no pp contacts in the bonded system (the reference) */
  double beta_bond;        /* non-viscous damping coefficient (0.1) */
  double radius_factor;    /* bond radius = min(r_i, r_j) * radius_factor (1.0) */
  double tensile_strength; /* > 0: a bond breaks once E_i (r - L0) / (r_i + r_j)
                            * exceeds it; <= 0: unbreakable (the reference) */
} dem_bond_config;

/* Host view of the bond state in the reference layout (bond fields of
 * src/CabanaBondedDEM_Particles.hpp:40-47,187-201), all rows [n][max_bonds].  Slot k
 * of particle i holds partner idx (original id; -1 = empty or broken), its
 * rest length and the history of that bond.  The reference indexes the
 * history and rest length by the partner id j instead of the slot k
 * (src/CabanaBondedDEM_Force.hpp:207,231-233); this library uses the slot.
 * NULL pointers are skipped. */
typedef struct dem_bonds_view {
  int n;
  int max_bonds;
  int *idx;                      /* bond_idx */
  int *count;                    /* [n] total_no_bonds */
  double *init_len;              /* bond_init_len */
  double *ft_x, *ft_y, *ft_z;    /* tangential force history */
  double *mn_x, *mn_y, *mn_z;    /* normal (torsion) moment history */
  double *mt_x, *mt_y, *mt_z;    /* tangential (bending) moment history */
} dem_bonds_view;

/* Run-average counters (extension; the reference reports none). */
typedef struct dem_counters {
  int64_t steps;          /* steps taken by this context */
  int64_t rebuilds;       /* neighbour rebuilds */
  int max_neighbors_seen; /* largest neighbour row at the last rebuild */
  int neighbor_capacity;  /* current row capacity */
  double mean_neighbors;  /* n-bar at the last sampled step */
  double mean_contacts;   /* c-bar (tracked pp contacts) at the last sampled step */
  double mean_wall_contacts; /* w-bar at the last sampled step */
  int64_t bonds_broken;   /* bonds broken so far (tensile_strength > 0) */
  int64_t displacement_checks; /* measurements of the automatic rebuild cadence */
} dem_counters;

/* ---- lifecycle ---------------------------------------------------------- */
/* Replaces: createSolverDEM / SolverDEM::SolverDEM (src/CabanaDEM_Solver.hpp:32-64,127-138)
 * together with the device allocation of Particles::resize (src/CabanaDEM_Particles.hpp:273-297). */
int dem_create(const dem_config* cfg, dem_ctx** out);
void dem_destroy(dem_ctx* ctx);
const char* dem_last_error(const dem_ctx* ctx); /* ctx may be NULL: last global error */
/* Updates the non-structural parameters of a live context: dt, restitution
 * and friction coefficients, gravity, skin, rebuild_every, physics_mode.
 * n, max_contacts, cutoff, lo and hi must equal the creation values.
 * Replaces: constructing a new Force (src/CabanaDEM_Force.hpp:23-28) or
 * Integrator (src/CabanaDEM_Integrator.hpp:22-27) with other scalars. */
int dem_set_params(dem_ctx* ctx, const dem_config* cfg);
int dem_abi_version(void);
int dem_device_count(void);

/* ---- transfer ------------------------------------------------------------ */
/* Replaces: Particles::updateParticles (src/CabanaDEM_Particles.hpp:90-97) -- the
 * user functor runs on the host mirror and the result is uploaded here. */
int dem_upload_particles(dem_ctx* ctx, const dem_particles_view* v);
/* Replaces: host access through the Particles slices after a run
 * (slice getters src/CabanaDEM_Particles.hpp:99-271). */
int dem_download_particles(dem_ctx* ctx, dem_particles_view* v);
/* Replaces: Wall::updateParticles (src/CabanaDEM_Wall.hpp:75-82).  Sets the
 * number of walls in use (0 .. DEM_MAX_WALLS). */
int dem_upload_walls(dem_ctx* ctx, const dem_walls_view* w);

/* ---- neighbour list ------------------------------------------------------ */
/* Replaces: Cabana::VerletList<...,FullNeighborTag,...>::build as called at
 * src/CabanaDEM_Solver.hpp:87-88 (and the constructor call :60-63).  Sorts
 * the particles by cell, rebuilds the full list (every j != i with
 * |x_i - x_j|^2 <= cutoff^2) and carries the tangential history across. */
int dem_build_neighbors(dem_ctx* ctx);
/* Returns the current list in original particle ids, each row in ascending
 * id order: counts[n], ids[n][max_n].  For parity checks. */
int dem_get_neighbors(dem_ctx* ctx, int* counts, int* ids, int max_n);

/* ---- force stages ---------------------------------------------------------- */
/* No counterpart in the reference: state hand-over between this library and
 * other device code (the surrogate's judge) without host copies.  Arrays
 * are device memory in the reference order, double [n][3]; either may be
 * NULL.  dem_set_kinematics_device sets positions and velocities and, with
 * clear_history, zeroes spins, forces, torques and every pp and wall history
 * (not for bonded contexts); new positions invalidate the neighbour list.
 * dem_get_kinematics_device copies positions and velocities out.  Both are
 * stream-ordered on the context's stream: synchronise other streams first,
 * and dem_synchronize before reading the results from another stream. */
int dem_set_kinematics_device(dem_ctx* ctx, const double* x, const double* u, int clear_history);
int dem_get_kinematics_device(dem_ctx* ctx, double* x, double* u);

/* Replaces: Force::makeForceTorqueZeroOnParticle (src/CabanaDEM_Force.hpp:39-47)
 * and resetForcesAndTorques (src/CabanaBondedDEM_Force.hpp:15-50);
 * F starts at m*gravity plus the external force, if one is set. */
int dem_zero_forces(dem_ctx* ctx);
/* Replaces: Force::update_tangential_contacts (src/CabanaDEM_Force.hpp:343-445). */
int dem_update_tangential(dem_ctx* ctx);
/* Replaces: Force::computeForceFullParticleParticle (src/CabanaDEM_Force.hpp:49-341). */
int dem_force_pp(dem_ctx* ctx);
/* Replaces: Force::computeForceFullParticleInfiniteWall (src/CabanaDEM_Force.hpp:447-678). */
int dem_force_pw(dem_ctx* ctx);
/* Replaces: computeForceParticleParticleInfiniteWall (src/CabanaDEM_Force.hpp:698-710):
 * zero + update tangential + pp + pw in one kernel. */
int dem_compute_forces(dem_ctx* ctx);

/* ---- integrator -------------------------------------------------------------- */
/* Replace: Integrator::stage1 / stage2 / stage3 (src/CabanaDEM_Integrator.hpp:31-56, 59-74, 77-103). */
int dem_stage1(dem_ctx* ctx);
int dem_stage2(dem_ctx* ctx);
int dem_stage3(dem_ctx* ctx);

/* ---- whole steps ------------------------------------------------------------- */
/* Replaces: the body of SolverDEM::run (src/CabanaDEM_Solver.hpp:79-97) for
 * nsteps steps: stage1, stage2, rebuild every rebuild_every steps, forces,
 * stage3.  Consecutive steps are fused: one kernel computes the forces of
 * step s, the closing half kick of s and the opening kick and drift of s+1. */
int dem_step(dem_ctx* ctx, int nsteps);
/* Sample KE/momentum (and n-bar, c-bar, w-bar) after stage3 of every step
 * whose global index is a multiple of every (0 = off).  Extension. */
int dem_set_diag_every(dem_ctx* ctx, int every);
/* Rows of 9 doubles: step, KE_trans, KE_rot, px, py, pz, n-bar, c-bar, w-bar. */
int dem_get_diag_history(dem_ctx* ctx, double* rows, int cap, int* nrows);
/* Current KE_trans, KE_rot, px, py, pz (synchronises). */
int dem_diagnostics(dem_ctx* ctx, double out[5]);
int dem_get_counters(dem_ctx* ctx, dem_counters* out);
int dem_synchronize(dem_ctx* ctx);

/* ---- bonded DEM ---------------------------------------------------------- */
/* Turns the context into a bonded system (allocates max_bonds slots per
 * particle; no bonds yet).  Replaces: instantiating ParticlesBondedDEM
 * (src/CabanaBondedDEM_Particles.hpp:29-83).  May be called again to change
 * the scalar parameters; max_bonds must then stay the same. */
int dem_enable_bonds(dem_ctx* ctx, const dem_bond_config* cfg);
/* Bonds every particle i to each neighbour j with |x_i - x_j| < factor r_i,
 * rest length |x_i - x_j|, zero history, over the context's neighbour list
 * (rebuilt first if stale).  factor * max r must not exceed the cutoff.
 * Replaces: ParticlesBondedDEM::setupBonds (src/CabanaBondedDEM_Particles.hpp:85-171;
 * the reference hard-codes factor 3 and a 4 r_0 search radius). */
int dem_setup_bonds(dem_ctx* ctx, double factor);
/* Bond state in and out, in the reference layout and original ids. */
int dem_upload_bonds(dem_ctx* ctx, const dem_bonds_view* v);
int dem_download_bonds(dem_ctx* ctx, dem_bonds_view* v);
/* F, T += bond forces and moments; updates the bond histories and, with a
 * tensile strength, breaks bonds.  Replaces: computeBondedForce
 * (src/CabanaBondedDEM_Force.hpp:52-348).  In a bonded system dem_step,
 * dem_compute_forces and dem_force_pp include the bonds (and, with
 * unbonded_contacts, restrict pp contacts to unbonded pairs). */
int dem_bond_forces(dem_ctx* ctx);
/* Constant per-particle external force, added where the force pass resets F
 * (F = m g + f_ext).  f = [n][3] in original order; NULL removes it.
 * Replaces: applyTensileForce (examples/tst_01_BDEM_two_particles_tensile_3D.cpp:69-117). */
int dem_set_external_force(dem_ctx* ctx, const double* f);
/* Particle groups: group[i] in 0..ngroups (0 = none), original order;
 * kind[g] = DEM_GROUP_*, vel[g][3] the prescribed velocity of a
 * DEM_GROUP_VELOCITY group.  group = NULL removes all groups.
 * Replaces: applyBoundaryConditions (examples/tst_01_...cpp:18-66) and the
 * pulled end faces of a tensile test. */
int dem_set_groups(dem_ctx* ctx, const int* group, int ngroups, const int* kind,
                   const double* vel);

/* ---- ensembles (extension; no counterpart in the reference) --------------- */
/* One context holding n_systems independent systems, advanced together: one
 * fused kernel launch per step covers every system, and one neighbour
 * rebuild re-sorts all of them (cell keys are system-major, so a system's
 * particles never leave its slot range and never see another system).
 * Particle arrays in the particle-view calls are the systems' particles
 * concatenated in system order (dem_get_systems gives the offsets).  Each
 * system evolves exactly as it would alone in its own context. */
int dem_create_ensemble(const dem_config* common, int n_systems, const dem_system_desc* systems,
                        dem_ctx** out);
/* offsets: [n_systems + 1] first particle of each system, then the total. */
int dem_get_systems(dem_ctx* ctx, int* n_systems, int* offsets);
int dem_upload_system_walls(dem_ctx* ctx, int system, const dem_walls_view* w);
int dem_set_system_params(dem_ctx* ctx, int system, double cor_pp, double cor_pw,
                          double friction_pp, double friction_pw);
/* out[s][5] = KE_trans, KE_rot, px, py, pz of each system, current state. */
int dem_system_diagnostics(dem_ctx* ctx, double* out);
/* Force-kernel launch shape: grid_blocks = 0 launches one workgroup per
 * tile of 128 particles; > 0 launches a persistent grid of that many
 * workgroups that stride over the tiles.  Results do not depend on it. */
int dem_set_launch(dem_ctx* ctx, int grid_blocks);
/* Force-pass shape (systems without bonds).  Results are bitwise the same
 * in every shape; which is fastest depends on the contact density.
 *   DEM_FORCE_PASS_FUSED (default): one kernel per force pass, each lane
 *     finding its particle's contacts over the neighbour row and walking
 *     them (the contact arithmetic, the kicks);
 *   DEM_FORCE_PASS_SPLIT: the detection as a kernel of its own that writes
 *     each particle's overlapping neighbours, then the force kernel running
 *     the contact arithmetic as a wave pass (below);
 *   DEM_FORCE_PASS_WAVE: one kernel; each lane finds its contacts, then the
 *     tile (workgroup) runs the contact arithmetic of all its lanes'
 *     contacts spread over its lanes, one contact per lane per round, in
 *     the same summation order.
 * Bonded systems always run DEM_FORCE_PASS_FUSED. */
enum dem_force_pass { DEM_FORCE_PASS_FUSED = 0, DEM_FORCE_PASS_SPLIT = 1, DEM_FORCE_PASS_WAVE = 2 };
int dem_set_force_pass(dem_ctx* ctx, int mode);

/* Tile shapes.  block: the force kernels' workgroup size and tile of
 * consecutive slots, 64, 128 or 256 (0: 128; systems with bonds run 128).
 * tile_rows: the binning grid numbers its cells in y-tiles of tile_rows
 * full x-rows, z-major inside a tile (1 to 32; 0: 8), which sets how far
 * apart in memory the z-neighbours of a particle sit after the re-sort.
 * Particle states do not depend on block; sampled diagnostics sum per tile,
 * so they agree to rounding across block sizes.  A new tile_rows re-sorts
 * the particles at the next step, which changes the order of the contact
 * sums (results agree to rounding, as across rebuild cadences). */
int dem_set_shape(dem_ctx* ctx, int block, int tile_rows);
int dem_get_shape(dem_ctx* ctx, int* block, int* tile_rows);

/* Captured step sequences: with enable set (and no diagnostics sampling or
 * kernel timing on), dem_step runs whole cycles of 2 rebuild_every steps
 * that start at a multiple of rebuild_every from one hipGraph, captured on
 * first use and replayed after; the steps outside whole cycles run as
 * usual.  Results are identical either way. */
int dem_set_graph(dem_ctx* ctx, int enable);

/* Autotuner: times trial_steps steps of candidate configurations from the
 * current state and applies the fastest, restoring the state exactly before
 * each trial and after the last.  Stage 1 tries every combination of
 * neighbour-rebuild cadence (5, 10, 20, 40, 80, DEM_REBUILD_AUTO and the
 * configured one, when a skin is set; a candidate that breaks the skin
 * check is rejected), force-kernel launch grid (one workgroup per tile, or
 * persistent grids of 2 and 8 workgroups per compute unit), captured graphs
 * on or off and force-pass shape (DEM_FORCE_PASS_*, systems without bonds)
 * at the context's tile shapes; stage 2 tries, at the best of stage 1, the
 * tile shapes of dem_set_shape: blocks of 64, 128 and 256 (systems without
 * bonds) times cell tiles of 4, 8 and 16 rows.  Rebuild cadence and cell
 * tiles change the summation order of the contact sums after a re-sort, so
 * runs with different ones agree to rounding, not bitwise. */
#define DEM_TUNE_MAX_TRIALS 128
typedef struct dem_tune_result {
  int rebuild_every;   /* chosen configuration */
  int grid_blocks;
  int use_graph;
  int force_pass;      /* DEM_FORCE_PASS_* */
  int block;           /* dem_set_shape */
  int tile_rows;
  int n_trials;
  int reserved;
  double ms_per_step;  /* of the chosen configuration */
  int trial_rebuild_every[DEM_TUNE_MAX_TRIALS];
  int trial_grid_blocks[DEM_TUNE_MAX_TRIALS];
  int trial_graph[DEM_TUNE_MAX_TRIALS];
  int trial_force_pass[DEM_TUNE_MAX_TRIALS];
  int trial_block[DEM_TUNE_MAX_TRIALS];
  int trial_tile_rows[DEM_TUNE_MAX_TRIALS];
  double trial_ms_per_step[DEM_TUNE_MAX_TRIALS]; /* < 0: rejected */
} dem_tune_result;
int dem_autotune(dem_ctx* ctx, int trial_steps, dem_tune_result* out);

/* ---- kernel timing (HIP events on the context's stream) ------------------ */
/* Enables per-launch HIP-event timing of the kernels dem_step launches
 * (0: off).  enable = k > 1 times every k-th force pass only (each timed
 * pass costs the GPU a few microseconds of idle time around its events);
 * passes that sample diagnostics (another kernel variant) or check the
 * displacements of the automatic cadence (extra reads) are not timed.
 * Every rebuild is timed. */
int dem_set_timing(dem_ctx* ctx, int enable);
/* Accumulated milliseconds and launch counts for the fused step kernel
 * (slot 0) and the rebuild sequence (slot 1). */
int dem_get_timing(dem_ctx* ctx, double ms[2], int64_t launches[2]);

/* ---- synthetic inputs (host-only, no device needed) ------------------------ */
/* Jittered-lattice packing: nx*ny*nz particles at spacing `spacing` from
 * origin lo, positions jittered U(-jitter, jitter), radii U(rmin, rmax),
 * velocities N(0, vsigma) per component; m = 4/3 pi r^3 rho and
 * I = 2/5 m r^2 as the reference drivers set them (examples/01_...cpp:106-107).
 * Counter-based splitmix64 stream keyed by (seed, particle, draw). */
int dem_generate_lattice(int nx, int ny, int nz, double spacing, const double lo[3],
                         double jitter, double rmin, double rmax, double vsigma,
                         double rho, double E, double nu, uint64_t seed,
                         double* x, double* u, double* rad, double* m, double* I,
                         double* Ey, double* nuv, double* G);

#ifdef __cplusplus
}
#endif
#endif
