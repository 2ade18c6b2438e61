// dem_oracle.cpp -- CPU oracle for the CabanaDEM hot path (TEST INFRASTRUCTURE).
//
// A line-by-line restatement of the reference's arithmetic in the reference's
// layout, used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline
// as the checker.  The product path never links or loads this file.
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (see oracle/Makefile).
//
// Reference files restated (paths relative to the ExaPhysics/CabanaDEM root):
//   src/CabanaDEM_Force.hpp        pp contact, tangential tracking, pw contact
//   src/CabanaDEM_Integrator.hpp   velocity-Verlet stages
//   src/CabanaDEM_Solver.hpp       stage order of one step
//   src/CabanaBondedDEM_Force.hpp  bond forces (history indexed by bond slot)
//   src/CabanaBondedDEM_Particles.hpp  bond set-up
// The neighbour list follows Cabana::VerletList<FullNeighborTag> set
// semantics as called at src/CabanaDEM_Solver.hpp:60-63,87-88: every j != i
// with dx*dx+dy*dy+dz*dz <= delta*delta.  Rows are ordered by ascending id.
#include "dem_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Neighbour list.  Linked cells of edge >= rc; cell indices are clamped into
// the box, which keeps the 27-cell search exact for particles outside it
// (floor-then-clamp moves two points at most one cell apart).  A margin of
// 1e-9 on the cell edge keeps rounding in (x-lo)/h from separating two
// points that are within rc by two cells.
// ---------------------------------------------------------------------------
struct Grid {
  double lo[3], h[3];
  int nc[3];
};

Grid make_grid(double rc, const double lo[3], const double hi[3]) {
  Grid g;
  for (int d = 0; d < 3; ++d) {
    double ext = hi[d] - lo[d];
    double cs = rc * (1.0 + 1e-9);
    int n = (int)std::floor(ext / cs);
    if (n < 1) n = 1;
    if (n > 2048) n = 2048;
    g.lo[d] = lo[d];
    g.h[d] = ext / n;
    if (g.h[d] < cs) g.h[d] = cs; // fewer, larger cells near the edges
    g.nc[d] = n;
  }
  return g;
}

inline int cell_of(const Grid& g, double v, int d) {
  double t = std::floor((v - g.lo[d]) / g.h[d]);
  if (!(t >= 0.0)) return 0; // also catches NaN
  if (t > (double)(g.nc[d] - 1)) return g.nc[d] - 1;
  return (int)t;
}

// Squared distance exactly as Cabana's VerletList forms it: a running sum
// over dimensions starting from zero, with no contraction.
inline double dist2(const double* x, int i, int j) {
  double dsqr = 0.0;
  for (int d = 0; d < 3; ++d) {
    double dx = x[3 * i + d] - x[3 * j + d];
    dsqr += dx * dx;
  }
  return dsqr;
}

// Reference mode constants: -2 * sqrt(5/6) with real division (normal
// damping, src/CabanaDEM_Force.hpp:215) and with integer division 5/6 == 0
// (tangential damping, :224, quirk Q2).
inline double eta_coef_normal() { return -2. * std::sqrt(5. / 6.); }
inline double eta_coef_tangent_ref() { return -2 * std::sqrt(5 / 6); }

// src/CabanaDEM_Force.hpp:206-210 -- beta from the restitution coefficient.
inline double beta_of(double cor) {
  double t1 = std::log(cor);
  double t2 = t1 * t1 + M_PI * M_PI;
  return t1 / std::sqrt(t2);
}

} // namespace

extern "C" {

int oracle_neighbors_cell(int n, const double* x, double rc, const double lo[3],
                          const double hi[3], int max_n, int* counts, int* ids) {
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d)
      if (!std::isfinite(x[3 * i + d])) return -1;
  Grid g = make_grid(rc, lo, hi);
  const long ncell = (long)g.nc[0] * g.nc[1] * g.nc[2];
  std::vector<int> key(n), start(ncell + 1, 0), order(n);
  for (int i = 0; i < n; ++i) {
    int c0 = cell_of(g, x[3 * i + 0], 0), c1 = cell_of(g, x[3 * i + 1], 1),
        c2 = cell_of(g, x[3 * i + 2], 2);
    key[i] = (int)(((long)c2 * g.nc[1] + c1) * g.nc[0] + c0);
    start[key[i] + 1]++;
  }
  for (long c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; ++i) order[fill[key[i]]++] = i; // ascending id per cell
  }
  const double rsqr = rc * rc;
  int maxc = 0;
#pragma omp parallel for schedule(static) reduction(max : maxc)
  for (int i = 0; i < n; ++i) {
    int c0 = key[i] % g.nc[0];
    int c1 = (key[i] / g.nc[0]) % g.nc[1];
    int c2 = key[i] / (g.nc[0] * g.nc[1]);
    std::vector<int> row;
    for (int dz = -1; dz <= 1; ++dz) {
      int z = c2 + dz;
      if (z < 0 || z >= g.nc[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        int y = c1 + dy;
        if (y < 0 || y >= g.nc[1]) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          int xx = c0 + dx;
          if (xx < 0 || xx >= g.nc[0]) continue;
          long c = ((long)z * g.nc[1] + y) * g.nc[0] + xx;
          for (int s = start[c]; s < start[c + 1]; ++s) {
            int j = order[s];
            if (j == i) continue;
            if (dist2(x, i, j) <= rsqr) row.push_back(j);
          }
        }
      }
    }
    std::sort(row.begin(), row.end());
    int cnt = (int)row.size();
    counts[i] = cnt;
    for (int k = 0; k < cnt && k < max_n; ++k) ids[(long)i * max_n + k] = row[k];
    if (cnt > maxc) maxc = cnt;
  }
  return maxc;
}

int oracle_neighbors_brute(int n, const double* x, double rc, int max_n,
                           int* counts, int* ids) {
  const double rsqr = rc * rc;
  int maxc = 0;
#pragma omp parallel for schedule(static) reduction(max : maxc)
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      if (dist2(x, i, j) <= rsqr) {
        if (cnt < max_n) ids[(long)i * max_n + cnt] = j;
        ++cnt;
      }
    }
    counts[i] = cnt;
    if (cnt > maxc) maxc = cnt;
  }
  return maxc;
}

// src/CabanaDEM_Force.hpp:39-47 (F = T = 0), extended with a body force:
// F_i starts at m_i * g (the reference has no gravity; g = 0 reproduces it).
void oracle_zero_forces(OracleParticles* p, const OracleParams* prm) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    for (int d = 0; d < 3; ++d) {
      p->force[3 * i + d] = 0.;
      p->torque[3 * i + d] = 0.;
    }
    if (prm && (prm->gravity[0] != 0. || prm->gravity[1] != 0. || prm->gravity[2] != 0.))
      for (int d = 0; d < 3; ++d) p->force[3 * i + d] = p->m[i] * prm->gravity[d];
  }
}

// src/CabanaDEM_Force.hpp:343-445 -- drop tracked pp contacts whose overlap
// is <= 0 by moving the last tracked slot into the hole (swap-remove).
void oracle_update_tangential(OracleParticles* p) {
  const int M = p->max_cnts;
  const double* x = p->x;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    int* idx = p->ss_idx + (long)i * M;
    double* sx = p->ss_x + (long)i * M;
    double* sy = p->ss_y + (long)i * M;
    double* sz = p->ss_z + (long)i * M;
    int count = 0, k = 0;
    const int total = p->ss_cnt[i];
    int last = total - 1;
    while (count < total) {                       // :372
      int s = idx[k];                             // :375
      if (s == -1) break;                         // :376-378
      double pij0 = x[3 * i + 0] - x[3 * s + 0];  // :388-390
      double pij1 = x[3 * i + 1] - x[3 * s + 1];
      double pij2 = x[3 * i + 2] - x[3 * s + 2];
      double r2 = pij0 * pij0 + pij1 * pij1 + pij2 * pij2;
      double rij = std::sqrt(r2);
      double overlap = p->rad[i] + p->rad[s] - rij; // :397
      if (overlap <= 0.) {                          // :399
        if (k == last) {                            // :402-407
          idx[k] = -1;
          sx[k] = 0.; sy[k] = 0.; sz[k] = 0.;
        } else {                                    // :408-429
          idx[k] = idx[last]; idx[last] = -1;
          sx[k] = sx[last]; sx[last] = 0.;
          sy[k] = sy[last]; sy[last] = 0.;
          sz[k] = sz[last]; sz[last] = 0.;
          last -= 1;
        }
        p->ss_cnt[i] -= 1;                          // :432
      } else {
        k = k + 1;                                  // :436
      }
      count += 1;                                   // :439
    }
  }
}

// src/CabanaDEM_Force.hpp:49-341 -- pp Hertz normal + tangential spring with
// tracking by neighbour id.  Accumulates into F_i (and, in corrected mode,
// T_i).  Returns 1 if any particle needed more than max_cnts tracked
// contacts (the reference writes out of bounds there; the oracle refuses).
static bool bonded_to(const OracleBonds* b, int i, int j) {
  if (!b) return false;
  for (int k = 0; k < b->count[i]; ++k)
    if (b->idx[(long)i * b->max_bonds + k] == j) return true;
  return false;
}

static int force_pp_impl(OracleParticles* p, const int* counts, const int* ids, int max_n,
                         const OracleParams* prm, const OracleBonds* skip);

int oracle_force_pp(OracleParticles* p, const int* counts, const int* ids,
                    int max_n, const OracleParams* prm) {
  return force_pp_impl(p, counts, ids, max_n, prm, nullptr);
}

int oracle_force_pp_unbonded(OracleParticles* p, const int* counts, const int* ids, int max_n,
                             const OracleParams* prm, const OracleBonds* b) {
  return force_pp_impl(p, counts, ids, max_n, prm, b);
}

void oracle_velocity_groups(OracleParticles* p, const int* group, const double* vel, int ngroups) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    const int g = group[i];
    if (g <= 0 || g > ngroups) continue;
    for (int d = 0; d < 3; ++d) {
      p->u[3 * i + d] = vel[3 * (g - 1) + d];
      p->omega[3 * i + d] = 0.;
    }
  }
}

static int force_pp_impl(OracleParticles* p, const int* counts, const int* ids, int max_n,
                         const OracleParams* prm, const OracleBonds* skip) {
  const int M = p->max_cnts;
  const bool ref = prm->mode == ORACLE_MODE_REFERENCE;
  const double dt = prm->dt;
  const double* x = p->x;
  const double* u = p->u;
  const double* w = p->omega;
  int overflow = 0;
#pragma omp parallel for schedule(static) reduction(| : overflow)
  for (int i = 0; i < p->n; ++i) {
    int* idx = p->ss_idx + (long)i * M;
    double* sx = p->ss_x + (long)i * M;
    double* sy = p->ss_y + (long)i * M;
    double* sz = p->ss_z + (long)i * M;
    for (int jj = 0; jj < counts[i]; ++jj) {
      const int j = ids[(long)i * max_n + jj];
      double pij0 = x[3 * i + 0] - x[3 * j + 0];               // :94-96
      double pij1 = x[3 * i + 1] - x[3 * j + 1];
      double pij2 = x[3 * i + 2] - x[3 * j + 2];
      double r2 = pij0 * pij0 + pij1 * pij1 + pij2 * pij2;     // :99
      double rij = std::sqrt(r2);                              // :101
      double overlap = p->rad[i] + p->rad[j] - rij;            // :107
      double a_i = p->rad[i] - overlap / 2.;                   // :109
      double a_j = p->rad[j] - overlap / 2.;                   // :110
      double nx = pij0 / rij, ny = pij1 / rij, nz = pij2 / rij; // :113-115
      double vi0, vi1, vi2, vj0, vj1, vj2;
      if (ref) {
        vi0 = u[3 * i + 0] + (w[3 * i + 1] * nz - w[3 * i + 2] * ny) * a_i; // :119-126
        vi1 = u[3 * i + 1] + (w[3 * i + 2] * nx - w[3 * i + 0] * nz) * a_i;
        vi2 = u[3 * i + 2] + (w[3 * i + 0] * ny - w[3 * i + 1] * nx) * a_i;
        vj0 = u[3 * j + 0] + (-w[3 * j + 1] * nz + w[3 * j + 2] * ny) * a_j; // :130-137 (Q1)
        vj1 = u[3 * i + 1] + (-w[3 * j + 2] * nx + w[3 * j + 0] * nz) * a_j;
        vj2 = u[3 * i + 2] + (-w[3 * j + 0] * ny + w[3 * j + 1] * nx) * a_j;
      } else {
        // contact point of i at -a_i n, of j at +a_j n; u_j in every component
        vi0 = u[3 * i + 0] - (w[3 * i + 1] * nz - w[3 * i + 2] * ny) * a_i;
        vi1 = u[3 * i + 1] - (w[3 * i + 2] * nx - w[3 * i + 0] * nz) * a_i;
        vi2 = u[3 * i + 2] - (w[3 * i + 0] * ny - w[3 * i + 1] * nx) * a_i;
        vj0 = u[3 * j + 0] + (w[3 * j + 1] * nz - w[3 * j + 2] * ny) * a_j;
        vj1 = u[3 * j + 1] + (w[3 * j + 2] * nx - w[3 * j + 0] * nz) * a_j;
        vj2 = u[3 * j + 2] + (w[3 * j + 0] * ny - w[3 * j + 1] * nx) * a_j;
      }
      double v0 = vi0 - vj0, v1 = vi1 - vj1, v2 = vi2 - vj2;   // :141-143
      double vdn = v0 * nx + v1 * ny + v2 * nz;                // :146
      double vn0 = vdn * nx, vn1 = vdn * ny, vn2 = vdn * nz;   // :147-149
      double vt0 = v0 - vn0, vt1 = v1 - vn1, vt2 = v2 - vn2;   // :152-154
      if (!(overlap > 0.)) continue;                           // :162
      if (bonded_to(skip, i, j)) continue;                     // bonded pairs: bond force only
      // :166-192 find j among the tracked ids (last match wins), else append
      int found_at = -1;
      for (int k = 0; k < p->ss_cnt[i]; ++k)
        if (j == idx[k]) found_at = k;
      if (found_at < 0) {
        if (p->ss_cnt[i] >= M) { overflow = 1; continue; }
        p->ss_cnt[i] += 1;
        found_at = p->ss_cnt[i] - 1;
        idx[found_at] = j;
        sx[found_at] = 0.; sy[found_at] = 0.; sz[found_at] = 0.;
      }
      double t1 = (1. - (p->nu[i] * p->nu[i])) / p->E[i];      // :196-203
      double t2 = (1. - (p->nu[j] * p->nu[j])) / p->E[j];
      double E_eff = 1. / (t1 + t2);
      double t3 = 1. / p->rad[i];
      double t4 = 1. / p->rad[j];
      double R_eff = 1. / (t3 + t4);
      double kn = 4. / 3. * E_eff * std::sqrt(R_eff);
      double beta = beta_of(prm->cor_pp);                      // :206-210
      double S_n = 2. * E_eff * std::sqrt(R_eff * overlap);    // :211
      t1 = 1. / p->m[i];                                       // :212-214
      t2 = 1. / p->m[j];
      double m_eff = 1. / (t1 + t2);
      double eta_n = eta_coef_normal() * beta * std::sqrt(S_n * m_eff); // :215
      t1 = (2. - p->nu[i]) / p->G[i];                          // :218-222
      t2 = (2. - p->nu[j]) / p->G[j];
      double G_eff = 1. / (t1 + t2);
      double kt = 8. * G_eff * std::sqrt(R_eff * overlap);
      double S_t = kt;
      double eta_t = (ref ? eta_coef_tangent_ref() : eta_coef_normal()) * beta *
                     std::sqrt(S_t * m_eff);                   // :224 (Q2)
      double fn = kn * std::pow(overlap, 1.5);                 // :227
      double fn0 = fn * nx - eta_n * vn0;                      // :228-230
      double fn1 = fn * ny - eta_n * vn1;
      double fn2 = fn * nz - eta_n * vn2;
      double tf0 = 0., tf1 = 0., tf2 = 0.;
      double vmag = std::pow(v0 * v0 + v1 * v1 + v2 * v2, 0.5); // :238
      if (vmag < 1e-12) {                                      // :239-243
        sx[found_at] = 0.; sy[found_at] = 0.; sz[found_at] = 0.;
      } else {
        double tx = v0 - nx * vdn, ty = v1 - ny * vdn, tz = v2 - nz * vdn; // :257-259
        double tmag = std::pow(tx * tx + ty * ty + tz * tz, 0.5);          // :261
        double ti0 = 0., ti1 = 0., ti2 = 0.;
        if (tmag > 1e-12) { ti0 = tx / tmag; ti1 = ty / tmag; ti2 = tz / tmag; } // :267-271
        double d0 = sx[found_at] + v0 * dt;                    // :274-276 (Q8)
        double d1 = sy[found_at] + v1 * dt;
        double d2 = sz[found_at] + v2 * dt;
        double ddt = (d0 * ti0 + d1 * ti1 + d2 * ti2);         // :278-280
        sx[found_at] = ddt * ti0;                              // :282-284
        sy[found_at] = ddt * ti1;
        sz[found_at] = ddt * ti2;
        double fs0 = -kt * sx[found_at] - eta_t * vt0;         // :286-288
        double fs1 = -kt * sy[found_at] - eta_t * vt1;
        double fs2 = -kt * sz[found_at] - eta_t * vt2;
        double ftm = std::pow(fs0 * fs0 + fs1 * fs1 + fs2 * fs2, 0.5); // :290
        double fnm = std::pow(fn0 * fn0 + fn1 * fn1 + fn2 * fn2, 0.5); // :291
        double ftms = ftm;                                     // :293-296
        if (ftms > prm->fr_pp * fnm) ftms = prm->fr_pp * fnm;
        tf0 = -ftms * ti0; tf1 = -ftms * ti1; tf2 = -ftms * ti2; // :299-301
        double m0 = -tf0 / kt, m1 = -tf1 / kt, m2 = -tf2 / kt; // :303-305
        if (ref) {
          double lm = std::pow(m0 * m0 + m1 * m1 + m2 * m2, 0.5); // :307-319 (Q3)
          if (lm > 1e-12) { sx[found_at] = m0 / lm; sy[found_at] = m1 / lm; sz[found_at] = m2 / lm; }
          else { sx[found_at] = 0.; sy[found_at] = 0.; sz[found_at] = 0.; }
        } else {
          // corrected: the spring keeps the elastic part of the applied
          // force, -(F_t + eta_t v_t) / k_t (k_t delta_t while sticking)
          sx[found_at] = -(tf0 + eta_t * vt0) / kt;
          sy[found_at] = -(tf1 + eta_t * vt1) / kt;
          sz[found_at] = -(tf2 + eta_t * vt2) / kt;
        }
      }
      p->force[3 * i + 0] += fn0 + tf0;                        // :324-326
      p->force[3 * i + 1] += fn1 + tf1;
      p->force[3 * i + 2] += fn2 + tf2;
      if (!ref) { // torque of the tangential force about i's centre, arm -a_i n
        p->torque[3 * i + 0] += (ny * tf2 - nz * tf1) * (-a_i);
        p->torque[3 * i + 1] += (nz * tf0 - nx * tf2) * (-a_i);
        p->torque[3 * i + 2] += (nx * tf1 - ny * tf0) * (-a_i);
      }
    }
  }
  return overflow;
}

// src/CabanaDEM_Force.hpp:447-678 -- particle vs infinite plane walls.
void oracle_force_pw(OracleParticles* p, const OracleWalls* wl, const OracleParams* prm) {
  const int MW = p->max_walls;
  const bool ref = prm->mode == ORACLE_MODE_REFERENCE;
  const double dt = prm->dt;
  const double* x = p->x;
  const double* u = p->u;
  const double* w = p->omega;
  const int nw = std::min(wl->n, MW);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    for (int j = 0; j < nw; ++j) {                             // :477
      double* swx = p->sw_x + (long)i * MW + j;
      double* swy = p->sw_y + (long)i * MW + j;
      double* swz = p->sw_z + (long)i * MW + j;
      // the wall point at the force pass: static in reference mode (the
      // reference never moves a wall); corrected mode translates it with u
      double xw0 = wl->x[3 * j + 0], xw1 = wl->x[3 * j + 1], xw2 = wl->x[3 * j + 2];
      if (!ref && wl->u) {
        xw0 += wl->u[3 * j + 0] * prm->wall_time;
        xw1 += wl->u[3 * j + 1] * prm->wall_time;
        xw2 += wl->u[3 * j + 2] * prm->wall_time;
      }
      double pij0 = x[3 * i + 0] - xw0;                        // :497-499
      double pij1 = x[3 * i + 1] - xw1;
      double pij2 = x[3 * i + 2] - xw2;
      double nx = wl->normal[3 * j + 0], ny = wl->normal[3 * j + 1], nz = wl->normal[3 * j + 2];
      double tmp = pij0 * nx + pij1 * ny + pij2 * nz;          // :514
      double overlap = p->rad[i] - tmp;                        // :515
      double a_i = p->rad[i] - overlap / 2.;                   // :517
      double vi0, vi1, vi2;
      if (ref) {                                               // :520-527
        vi0 = u[3 * i + 0] + (w[3 * i + 1] * nz - w[3 * i + 2] * ny) * a_i;
        vi1 = u[3 * i + 1] + (w[3 * i + 2] * nx - w[3 * i + 0] * nz) * a_i;
        vi2 = u[3 * i + 2] + (w[3 * i + 0] * ny - w[3 * i + 1] * nx) * a_i;
      } else {                                                 // contact point at -a_i n (Q5)
        vi0 = u[3 * i + 0] - (w[3 * i + 1] * nz - w[3 * i + 2] * ny) * a_i;
        vi1 = u[3 * i + 1] - (w[3 * i + 2] * nx - w[3 * i + 0] * nz) * a_i;
        vi2 = u[3 * i + 2] - (w[3 * i + 0] * ny - w[3 * i + 1] * nx) * a_i;
      }
      // :529-535: the reference sets the wall's velocity to 0 (Q7);
      // corrected mode takes the surface velocity at the contact point
      // c = x_i - a_i n: u_w + omega_w x (c - x_w)
      double vw0 = 0., vw1 = 0., vw2 = 0.;
      if (!ref && wl->u) {
        vw0 = wl->u[3 * j + 0]; vw1 = wl->u[3 * j + 1]; vw2 = wl->u[3 * j + 2];
      }
      if (!ref && wl->omega) {
        const double* ow = wl->omega + 3 * j;
        const double c0 = x[3 * i + 0] - a_i * nx - xw0;
        const double c1 = x[3 * i + 1] - a_i * ny - xw1;
        const double c2 = x[3 * i + 2] - a_i * nz - xw2;
        vw0 += ow[1] * c2 - ow[2] * c1;
        vw1 += ow[2] * c0 - ow[0] * c2;
        vw2 += ow[0] * c1 - ow[1] * c0;
      }
      double v0 = vi0 - vw0, v1 = vi1 - vw1, v2 = vi2 - vw2;
      double vdn = v0 * nx + v1 * ny + v2 * nz;                // :538
      double vn0 = vdn * nx, vn1 = vdn * ny, vn2 = vdn * nz;
      double vt0 = v0 - vn0, vt1 = v1 - vn1, vt2 = v2 - vn2;
      if (overlap > 0.) {                                      // :556
        double t1 = (1. - (p->nu[i] * p->nu[i])) / p->E[i];    // :561-568
        double t2 = (1. - (wl->nu[j] * wl->nu[j])) / wl->E[j];
        double E_eff = 1. / (t1 + t2);
        double t3 = 1. / p->rad[i];
        double t4 = 0.;
        double R_eff = 1. / (t3 + t4);
        double kn = 4. / 3. * E_eff * std::sqrt(R_eff);
        double beta = beta_of(prm->cor_pw);                    // :571-575
        double S_n = 2. * E_eff * std::sqrt(R_eff * overlap);
        t1 = 1. / p->m[i];
        t2 = 0.;
        double m_eff = 1. / (t1 + t2);
        double eta_n = eta_coef_normal() * beta * std::sqrt(S_n * m_eff); // :580
        double fn = kn * std::pow(overlap, 1.5);               // :583-586
        double fn0 = fn * nx - eta_n * vn0;
        double fn1 = fn * ny - eta_n * vn1;
        double fn2 = fn * nz - eta_n * vn2;
        double ft0 = 0., ft1 = 0., ft2 = 0.;
        double vmag = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);  // :597
        if (vmag < 1e-12) {                                    // :599-603
          *swx = 0.; *swy = 0.; *swz = 0.;
        } else {
          *swx += vt0 * dt; *swy += vt1 * dt; *swz += vt2 * dt; // :607-609
          t1 = (2. - p->nu[i]) / p->G[i];                      // :612-616
          t2 = (2. - wl->nu[j]) / wl->G[j];
          double G_eff = 1. / (t1 + t2);
          double kt = 8. * G_eff * std::sqrt(R_eff * overlap);
          double S_t = kt;
          double eta_t = (ref ? eta_coef_tangent_ref() : eta_coef_normal()) * beta *
                         std::sqrt(S_t * m_eff);               // :618 (Q2)
          double fs0 = -kt * *swx - eta_t * vt0;               // :620-622
          double fs1 = -kt * *swy - eta_t * vt1;
          double fs2 = -kt * *swz - eta_t * vt2;
          double ftm = ref ? std::sqrt(fs0 * fs0 + fs1 * fs1 + fs2 * fs1)   // :624 (Q6)
                           : std::sqrt(fs0 * fs0 + fs1 * fs1 + fs2 * fs2);
          double ti0 = 0., ti1 = 0., ti2 = 0.;
          if (ftm > 1e-12) { ti0 = fs0 / ftm; ti1 = fs1 / ftm; ti2 = fs2 / ftm; } // :630-634
          double fnm = std::sqrt(fn0 * fn0 + fn1 * fn1 + fn2 * fn2); // :636
          double ftms = std::min(prm->fr_pw * fnm, ftm);       // :638
          ft0 = ftms * ti0; ft1 = ftms * ti1; ft2 = ftms * ti2; // :641-643
          if (ref) {
            // :646-648 add +eta_t v_t (along the slip; quirk Q16, inert with
            // eta_t = 0), then reset the spring to the applied force (:651-653)
            ft0 += eta_t * vt0; ft1 += eta_t * vt1; ft2 += eta_t * vt2;
            *swx = -ft0 / kt; *swy = -ft1 / kt; *swz = -ft2 / kt;
          } else {
            // corrected: the limited force holds the damping already; the
            // spring keeps its elastic part
            *swx = -(ft0 + eta_t * vt0) / kt;
            *swy = -(ft1 + eta_t * vt1) / kt;
            *swz = -(ft2 + eta_t * vt2) / kt;
          }
        }
        p->force[3 * i + 0] += fn0 + ft0;                      // :657-659
        p->force[3 * i + 1] += fn1 + ft1;
        p->force[3 * i + 2] += fn2 + ft2;
        if (ref) {                                             // :662-664
          p->torque[3 * i + 0] += (ny * ft2 - nz * ft1) * a_i;
          p->torque[3 * i + 1] += (nz * ft0 - nx * ft2) * a_i;
          p->torque[3 * i + 2] += (nx * ft1 - ny * ft0) * a_i;
        } else {
          p->torque[3 * i + 0] += (ny * ft2 - nz * ft1) * (-a_i);
          p->torque[3 * i + 1] += (nz * ft0 - nx * ft2) * (-a_i);
          p->torque[3 * i + 2] += (nx * ft1 - ny * ft0) * (-a_i);
        }
      } else {                                                 // :666-670
        *swx = 0.; *swy = 0.; *swz = 0.;
      }
    }
  }
}

// src/CabanaDEM_Integrator.hpp:31-56 (stage1) and :77-103 (stage3) are the
// same half kick: u += (dt/2 F) / m, omega += (dt/2 T) / I.
static void half_kick(OracleParticles* p, double dt) {
  const double half_dt = 0.5 * dt;                             // :25
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    double m_inverse = 1. / p->m[i];                           // :42
    for (int d = 0; d < 3; ++d) p->u[3 * i + d] += half_dt * p->force[3 * i + d] * m_inverse;
    double I_inverse = 1. / p->I[i];                           // :48
    for (int d = 0; d < 3; ++d) p->omega[3 * i + d] += half_dt * p->torque[3 * i + d] * I_inverse;
  }
}
void oracle_stage1(OracleParticles* p, double dt) { half_kick(p, dt); }
void oracle_stage3(OracleParticles* p, double dt) { half_kick(p, dt); }
// src/CabanaDEM_Integrator.hpp:59-74 -- drift.
void oracle_stage2(OracleParticles* p, double dt) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i)
    for (int d = 0; d < 3; ++d) p->x[3 * i + d] += dt * p->u[3 * i + d];
}

// Kinetic energy (translational, rotational) and linear momentum, summed in
// ascending particle order.  Extension: the reference reports none.
void oracle_diagnostics(const OracleParticles* p, double out[5]) {
  double ket = 0., ker = 0., px = 0., py = 0., pz = 0.;
  for (int i = 0; i < p->n; ++i) {
    const double* u = p->u + 3 * i;
    const double* w = p->omega + 3 * i;
    ket += 0.5 * p->m[i] * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    ker += 0.5 * p->I[i] * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    px += p->m[i] * u[0];
    py += p->m[i] * u[1];
    pz += p->m[i] * u[2];
  }
  out[0] = ket; out[1] = ker; out[2] = px; out[3] = py; out[4] = pz;
}

// SolverDEM::run, src/CabanaDEM_Solver.hpp:79-97: per step stage1, stage2,
// neighbour build, forces (zero, update tangential, pp, pw), stage3.
// Extensions: the list is rebuilt when step % rebuild_every == 0 (the
// reference rebuilds every step, :87-88) and diagnostics are sampled after
// stage3 when step % diag_every == 0.  Returns 0, or a negative error.
int oracle_run(OracleParticles* p, const OracleWalls* w, const OracleParams* prm,
               const OracleRun* run, double* diag, int diag_cap, int* n_diag,
               double* seconds) {
#ifdef _OPENMP
  int saved = omp_get_max_threads();
  if (run->threads > 0) omp_set_num_threads(run->threads);
#endif
  const int max_n = run->max_n;
  std::vector<int> counts(p->n), ids((long)p->n * max_n);
  int nd = 0, err = 0;
  const int K = run->rebuild_every > 0 ? run->rebuild_every : 1;
  auto t0 = std::chrono::steady_clock::now();
  for (int step = 0; step < run->nsteps; ++step) {
    oracle_stage1(p, prm->dt);
    oracle_stage2(p, prm->dt);
    if (step % K == 0) {
      int mc = oracle_neighbors_cell(p->n, p->x, run->rc, run->lo, run->hi, max_n,
                                     counts.data(), ids.data());
      if (mc < 0) { err = -2; break; }
      if (mc > max_n) { err = -3; break; }
    }
    oracle_zero_forces(p, prm);
    oracle_update_tangential(p);
    if (oracle_force_pp(p, counts.data(), ids.data(), max_n, prm)) { err = -4; break; }
    if (w && w->n > 0) {
      // moving walls: the force pass of this run's step `step` sits at
      // wall_time + (step + 1) dt (after the drift)
      OracleParams ps = *prm;
      ps.wall_time = prm->wall_time + (step + 1) * prm->dt;
      oracle_force_pw(p, w, &ps);
    }
    oracle_stage3(p, prm->dt);
    if (run->diag_every > 0 && step % run->diag_every == 0 && diag && nd < diag_cap) {
      double d5[5];
      oracle_diagnostics(p, d5);
      diag[6 * nd + 0] = step;
      for (int q = 0; q < 5; ++q) diag[6 * nd + 1 + q] = d5[q];
      ++nd;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  if (n_diag) *n_diag = nd;
#ifdef _OPENMP
  omp_set_num_threads(saved);
#endif
  return err;
}

// src/CabanaBondedDEM_Particles.hpp:85-171 -- bond every list neighbour j of
// i with r_ij < factor * rad(i) (the reference hard-codes factor 3, :130),
// storing id and initial length in slot total_no_bonds(i) (:133-135).
int oracle_setup_bonds(int n, const double* x, const double* rad, double factor,
                       double rc, const double lo[3], const double hi[3],
                       OracleBonds* b) {
  const int MB = b->max_bonds;
  int max_n = 64;
  std::vector<int> counts(n), ids;
  for (;;) {
    ids.assign((long)n * max_n, 0);
    int mc = oracle_neighbors_cell(n, x, rc, lo, hi, max_n, counts.data(), ids.data());
    if (mc < 0) return -2;
    if (mc <= max_n) break;
    max_n = mc;
  }
  int overflow = 0;
  for (int i = 0; i < n; ++i) {
    b->count[i] = 0;
    for (int k = 0; k < MB; ++k) {
      b->idx[(long)i * MB + k] = -1;
      b->init_len[(long)i * MB + k] = 0.;
      for (int d = 0; d < 3; ++d) {
        b->ft[((long)i * MB + k) * 3 + d] = 0.;
        b->mn[((long)i * MB + k) * 3 + d] = 0.;
        b->mt[((long)i * MB + k) * 3 + d] = 0.;
      }
    }
    for (int jj = 0; jj < counts[i]; ++jj) {
      int j = ids[(long)i * max_n + jj];
      double p0 = x[3 * i + 0] - x[3 * j + 0];                 // :120-126
      double p1 = x[3 * i + 1] - x[3 * j + 1];
      double p2 = x[3 * i + 2] - x[3 * j + 2];
      double r2 = p0 * p0 + p1 * p1 + p2 * p2;
      double rij = std::sqrt(r2);
      double cond = factor * rad[i];                           // :130
      if (rij < cond) {                                        // :132-136
        if (b->count[i] >= MB) { overflow = 1; continue; }
        b->idx[(long)i * MB + b->count[i]] = j;
        b->init_len[(long)i * MB + b->count[i]] = rij;
        b->count[i] += 1;
      }
    }
  }
  return overflow ? -3 : 0;
}

// src/CabanaBondedDEM_Force.hpp:52-348 -- beam bond force and moments with
// non-viscous damping, accumulated into F_i, T_i.  History and rest length
// are indexed by bond slot k (the reference indexes them by the neighbour id
// j, quirk Q9, which is out of bounds for j >= MaxBonds).  Extensions: a bond
// whose tensile stress E_i (r - L0) / (r_i + r_j) exceeds tensile_strength
// breaks (its id becomes -1 and it is skipped from then on).
void oracle_bond_force(OracleParticles* p, OracleBonds* b, const OracleBondParams* bp,
                       double dt, int* n_broken) {
  const int MB = b->max_bonds;
  const bool ref = bp->mode == ORACLE_MODE_REFERENCE;
  const double* x = p->x;
  const double* u = p->u;
  const double* w = p->omega;
  int broken = 0;
#pragma omp parallel for schedule(static) reduction(+ : broken)
  for (int i = 0; i < p->n; ++i) {
    const int nb = b->count[i];
    for (int k = 0; k < nb; ++k) {                             // :88-89
      const long s = (long)i * MB + k;
      const int j = b->idx[s];                                 // :91
      if (j < 0) continue;
      double pij0 = x[3 * i + 0] - x[3 * j + 0];               // :106-108
      double pij1 = x[3 * i + 1] - x[3 * j + 1];
      double pij2 = x[3 * i + 2] - x[3 * j + 2];
      double r2 = pij0 * pij0 + pij1 * pij1 + pij2 * pij2;     // :111
      double rij = std::sqrt(r2);                              // :113
      double rad_sum = p->rad[i] + p->rad[j];                  // :182
      if (bp->tensile_strength > 0.) {
        double stretch = rij - b->init_len[s];
        if (stretch > 0. && p->E[i] * stretch / rad_sum > bp->tensile_strength) {
          b->idx[s] = -1;
          broken += 1;
          continue;
        }
      }
      double g_c = rij - p->rad[i] - p->rad[j];                // :115
      double a_i = p->rad[i] + g_c / 2.;                       // :116
      double a_j = p->rad[j] + g_c / 2.;                       // :117
      double n0 = pij0 / rij, n1 = pij1 / rij, n2 = pij2 / rij; // :122
      double vi0 = u[3 * i + 0] + (w[3 * i + 1] * n2 - w[3 * i + 2] * n1) * a_i; // :130-137
      double vi1 = u[3 * i + 1] + (w[3 * i + 2] * n0 - w[3 * i + 0] * n2) * a_i;
      double vi2 = u[3 * i + 2] + (w[3 * i + 0] * n1 - w[3 * i + 1] * n0) * a_i;
      double vj0 = u[3 * j + 0] + (-w[3 * j + 1] * n2 + w[3 * j + 2] * n1) * a_j; // :139-146 (Q1)
      double vj1 = (ref ? u[3 * i + 1] : u[3 * j + 1]) + (-w[3 * j + 2] * n0 + w[3 * j + 0] * n2) * a_j;
      double vj2 = (ref ? u[3 * i + 2] : u[3 * j + 2]) + (-w[3 * j + 0] * n1 + w[3 * j + 1] * n0) * a_j;
      double vr0 = vi0 - vj0, vr1 = vi1 - vj1, vr2 = vi2 - vj2; // :148-150
      double vdp = vr0 * n0 + vr1 * n1 + vr2 * n2;             // :153
      double vn0 = vdp * n0, vn1 = vdp * n1, vn2 = vdp * n2;   // :154-156
      double vt0 = vr0 - vn0, vt1 = vr1 - vn1, vt2 = vr2 - vn2; // :158-160
      double wr0 = w[3 * i + 0] - w[3 * j + 0];                // :163-165
      double wr1 = w[3 * i + 1] - w[3 * j + 1];
      double wr2 = w[3 * i + 2] - w[3 * j + 2];
      double wdp = wr0 * n0 + wr1 * n1 + wr2 * n2;             // :167
      double wn0 = wdp * n0, wn1 = wdp * n1, wn2 = wdp * n2;   // :168
      double wt0 = wr0 - wn0, wt1 = wr1 - wn1, wt2 = wr2 - wn2; // :169
      const double beta_bond = bp->beta_bond;                  // :186
      double bond_rad = std::min(p->rad[i], p->rad[j]) * bp->radius_factor; // :188
      double A_bond = M_PI * std::pow(bond_rad, 2.);           // :189
      double I_bond = M_PI * std::pow(bond_rad, 4.) / 4.0;     // :190
      double I_p_bond = I_bond * 2.0;                          // :191
      double k_n = p->E[i] * A_bond / rad_sum;                 // :195-198
      double k_t = p->G[i] * A_bond / rad_sum;
      double k_tor = p->G[i] * I_p_bond / rad_sum;
      double k_ben = p->E[i] * I_bond / rad_sum;
      double overlap_n = rij - b->init_len[s];                 // :207 (slot k, Q9)
      double fn0 = -k_n * overlap_n * n0;                      // :208-210
      double fn1 = -k_n * overlap_n * n1;
      double fn2 = -k_n * overlap_n * n2;
      double mx = (vn0 >= 0.0) ? 1.0 : -1.0;                   // :216-218
      double my = (vn1 >= 0.0) ? 1.0 : -1.0;
      double mz = (vn2 >= 0.0) ? 1.0 : -1.0;
      // :220-222 add +beta|fn| sign(v_n), feeding energy in (quirk Q15);
      // corrected mode subtracts it, as the tangential terms do (:258-260)
      const double bs = ref ? beta_bond : -beta_bond;
      double fd0 = fn0 + bs * std::fabs(fn0) * mx;
      double fd1 = fn1 + bs * std::fabs(fn1) * my;
      double fd2 = fn2 + bs * std::fabs(fn2) * mz;
      double* ft = b->ft + 3 * s;                              // :231-233
      double f0 = ft[0], f1 = ft[1], f2 = ft[2];
      double fdp = f0 * n0 + f1 * n1 + f2 * n2;                // :235-239
      f0 -= fdp * n0; f1 -= fdp * n1; f2 -= fdp * n2;
      double tmp = k_t * dt;                                   // :241-246
      f0 += -vt0 * tmp; f1 += -vt1 * tmp; f2 += -vt2 * tmp;
      mx = (vt0 >= 0.0) ? 1.0 : -1.0;                          // :251-256
      my = (vt1 >= 0.0) ? 1.0 : -1.0;
      mz = (vt2 >= 0.0) ? 1.0 : -1.0;
      double td0 = f0 - beta_bond * std::fabs(f0) * mx;
      double td1 = f1 - beta_bond * std::fabs(f1) * my;
      double td2 = f2 - beta_bond * std::fabs(f2) * mz;
      ft[0] = td0; ft[1] = td1; ft[2] = td2;                   // :258-260
      double* mn = b->mn + 3 * s;                              // :267-274
      double q0 = mn[0], q1 = mn[1], q2 = mn[2];
      double tdp = q0 * n0 + q1 * n1 + q2 * n2;
      q0 = tdp * n0; q1 = tdp * n1; q2 = tdp * n2;
      tmp = k_tor * dt;                                        // :276-280
      q0 += -wn0 * tmp; q1 += -wn1 * tmp; q2 += -wn2 * tmp;
      mx = (wn0 >= 0.0) ? 1.0 : -1.0;                          // :282-287
      my = (wn1 >= 0.0) ? 1.0 : -1.0;
      mz = (wn2 >= 0.0) ? 1.0 : -1.0;
      double nd0 = q0 - beta_bond * std::fabs(q0) * mx;
      double nd1 = q1 - beta_bond * std::fabs(q1) * my;
      double nd2 = q2 - beta_bond * std::fabs(q2) * mz;
      mn[0] = nd0; mn[1] = nd1; mn[2] = nd2;                   // :289-291
      double* mt = b->mt + 3 * s;                              // :298-304
      double t0 = mt[0], t1 = mt[1], t2 = mt[2];
      tdp = t0 * n0 + t1 * n1 + t2 * n2;
      t0 -= tdp * n0; t1 -= tdp * n1; t2 -= tdp * n2;
      tmp = k_ben * dt;                                        // :306-310
      t0 += -wt0 * tmp; t1 += -wt1 * tmp; t2 += -wt2 * tmp;
      mx = (wt0 >= 0.0) ? 1.0 : -1.0;                          // :313-318
      my = (wt1 >= 0.0) ? 1.0 : -1.0;
      mz = (wt2 >= 0.0) ? 1.0 : -1.0;
      double md0 = t0 - beta_bond * std::fabs(t0) * mx;
      double md1 = t1 - beta_bond * std::fabs(t1) * my;
      double md2 = t2 - beta_bond * std::fabs(t2) * mz;
      mt[0] = md0; mt[1] = md1; mt[2] = md2;                   // :320-322
      double cri = rij * 1. / p->rad[i] * 1. / (p->rad[i] + p->rad[j]); // :327
      p->force[3 * i + 0] += fd0 + td0;                        // :334-336
      p->force[3 * i + 1] += fd1 + td1;
      p->force[3 * i + 2] += fd2 + td2;
      p->torque[3 * i + 0] += nd0 + md0 - cri * 0.0;           // :338-340 (tor = 0)
      p->torque[3 * i + 1] += nd1 + md1 - cri * 0.0;
      p->torque[3 * i + 2] += nd2 + md2 - cri * 0.0;
    }
  }
  if (n_broken) *n_broken = broken;
}

} // extern "C"
