/*
 * dem_oracle.h -- CPU oracle for the CabanaDEM per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++/OpenMP restatement of the
 * reference solver (ExaPhysics/CabanaDEM, src/CabanaDEM_Force.hpp,
 * src/CabanaDEM_Integrator.hpp, src/CabanaDEM_Solver.hpp,
 * src/CabanaBondedDEM_Force.hpp, src/CabanaBondedDEM_Particles.hpp) in the
 * reference's own data layout: xyz-interleaved vectors double[N][3] and
 * row-major histories T[N][Max].  It is used only by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, always as the
 * checker and never as the product path.  The product (cabanadem_amd) does
 * not link or load it.
 *
 * The reference itself cannot be compiled here (it needs Kokkos >= 3.7 and
 * Cabana >= 0.6.1, neither vendored nor installed), so there is no
 * oracle/_ref build; every function below cites the reference lines it
 * restates and is pinned by the tests listed in its comment.
 *
 * Floating point: build with -O2 -ffp-contract=off so every expression is
 * evaluated in the reference's source order with no fused multiply-adds.
 */
#ifndef CABANADEM_ORACLE_H
#define CABANADEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* physics_mode: 0 = reference arithmetic (all quirks Q1-Q6 reproduced),
 *               1 = corrected (u_j, sqrt(5./6.), displacement history,
 *                   pp torque, contact-point sign, |F_t*| typo fixed). */
enum { ORACLE_MODE_REFERENCE = 0, ORACLE_MODE_CORRECTED = 1 };

/* Particle state in the reference layout (src/CabanaDEM_Particles.hpp:348-370). */
typedef struct OracleParticles {
  int n;          /* particles */
  int max_cnts;   /* MaxCnts template parameter */
  int max_walls;  /* MaxNoWalls template parameter */
  double *x, *u, *omega, *force, *torque; /* [n][3] */
  double *m, *rad, *E, *nu, *G, *I;       /* [n] */
  double *sw_x, *sw_y, *sw_z;             /* [n][max_walls] */
  double *ss_x, *ss_y, *ss_z;             /* [n][max_cnts] */
  int *ss_idx;                            /* [n][max_cnts] */
  int *ss_cnt;                            /* [n] */
} OracleParticles;

/* Infinite walls (src/CabanaDEM_Wall.hpp:192-201). */
typedef struct OracleWalls {
  int n;
  double *x, *normal; /* [n][3] */
  double *E, *nu, *G; /* [n] */
  /* extension, corrected mode only (the reference reads them at
   * src/CabanaDEM_Force.hpp:468-469 and leaves them unused, :529): the
   * plane translates with u ([n][3]) from x at wall_time 0, and its surface
   * moves with u + omega x (c - x_w(t)) at a contact point c.  NULL: static. */
  double *u, *omega;
} OracleWalls;

typedef struct OracleParams {
  double cor_pp, cor_pw, fr_pp, fr_pw; /* Force ctor, src/CabanaDEM_Force.hpp:23-28 */
  double dt;
  double gravity[3]; /* extension: F_i starts at m_i g in the force pass */
  int mode;          /* ORACLE_MODE_* */
  double wall_time;  /* extension: time since the walls' x, for moving walls
                      * (oracle_run sets it per step from its own start) */
} OracleParams;

/* Driver options for oracle_run (SolverDEM::run, src/CabanaDEM_Solver.hpp:66-98). */
typedef struct OracleRun {
  int nsteps;         /* full steps to take (the reference's loop runs num_steps+1) */
  double rc;          /* neighbour cutoff (the reference's `delta`) */
  double lo[3], hi[3];/* binning box (mesh_lo / mesh_hi) */
  int rebuild_every;  /* 1 = rebuild every step, as the reference does */
  int max_n;          /* neighbour row capacity */
  int diag_every;     /* 0 = no diagnostics */
  int threads;        /* OpenMP threads, 0 = runtime default */
} OracleRun;

/* ---- neighbour list (Cabana::VerletList FullNeighborTag set semantics) ---- */
int oracle_neighbors_cell(int n, const double *x, double rc, const double lo[3],
                          const double hi[3], int max_n, int *counts, int *ids);
int oracle_neighbors_brute(int n, const double *x, double rc, int max_n,
                           int *counts, int *ids);

/* ---- force pass ---- */
void oracle_zero_forces(OracleParticles *p, const OracleParams *prm);
void oracle_update_tangential(OracleParticles *p);
int oracle_force_pp(OracleParticles *p, const int *counts, const int *ids,
                    int max_n, const OracleParams *prm);
void oracle_force_pw(OracleParticles *p, const OracleWalls *w,
                     const OracleParams *prm);

/* ---- integrator ---- */
void oracle_stage1(OracleParticles *p, double dt);
void oracle_stage2(OracleParticles *p, double dt);
void oracle_stage3(OracleParticles *p, double dt);

/* ---- whole run; diag rows are (step, KE_trans, KE_rot, px, py, pz) ---- */
int oracle_run(OracleParticles *p, const OracleWalls *w, const OracleParams *prm,
               const OracleRun *run, double *diag, int diag_cap, int *n_diag,
               double *seconds);
void oracle_diagnostics(const OracleParticles *p, double out[5]);

/* ---- bonded DEM ---- */
typedef struct OracleBonds {
  int max_bonds;
  int *idx;        /* [n][max_bonds] */
  double *init_len;/* [n][max_bonds] */
  double *ft;      /* [n][max_bonds][3] */
  double *mn;      /* [n][max_bonds][3] */
  double *mt;      /* [n][max_bonds][3] */
  int *count;      /* [n] */
} OracleBonds;

typedef struct OracleBondParams {
  double beta_bond;       /* 0.1 in the reference */
  double radius_factor;   /* 1.0 in the reference */
  double tensile_strength;/* <= 0: unbreakable (the reference has no breakage) */
  int contacts;           /* 1: Hertz contacts between unbonded overlapping pairs */
  int mode;               /* ORACLE_MODE_* (Q1 in the bond relative velocity) */
} OracleBondParams;

int oracle_setup_bonds(int n, const double *x, const double *rad, double factor,
                       double rc, const double lo[3], const double hi[3],
                       OracleBonds *b);
void oracle_bond_force(OracleParticles *p, OracleBonds *b,
                       const OracleBondParams *bp, double dt, int *n_broken);
/* oracle_force_pp restricted to pairs that are not joined by an intact bond. */
int oracle_force_pp_unbonded(OracleParticles *p, const int *counts, const int *ids,
                             int max_n, const OracleParams *prm, const OracleBonds *b);
/* Velocity-prescribed particles: group[i] > 0 sets u_i = vel[group-1], omega_i = 0. */
void oracle_velocity_groups(OracleParticles *p, const int *group, const double *vel, int ngroups);

#ifdef __cplusplus
}
#endif
#endif
