// dem_force_kernels.h -- the force kernels and their launch templates,
// shared by the translation units that instantiate them for one force-tile
// size each (dem_force.hip: 128 and the bonded kernels; dem_force_b64.hip,
// dem_force_b256.hip), so that they compile in parallel: contact forces and
// velocity-Verlet stages on the GPU.
//
// Reference paths are relative to the ExaPhysics/CabanaDEM root.
//   pp contact    src/CabanaDEM_Force.hpp:73-327  (Hertz normal + tangential spring)
//   tracking      src/CabanaDEM_Force.hpp:166-192, 343-445
//   pw contact    src/CabanaDEM_Force.hpp:477-671
//   stages        src/CabanaDEM_Integrator.hpp:31-103
// All contact arithmetic is FP64 in the reference's operation order, with
// the per-particle factors (1-nu^2)/E, (2-nu)/G, 1/r and 1/m precomputed
// (the same IEEE operations the reference performs per pair) and
// pow(x, 0.5) -> sqrt(x), pow(x, 1.5) -> x*sqrt(x), log(cor) hoisted.
//
// One thread owns particle i (full list, no atomics).  Phase 1 walks the
// neighbour row and keeps the j with positive overlap in LDS; phase 2 runs
// the contact arithmetic over that short list only, merging with the
// previous step's tracked contacts (both ascending in j): a contact found in
// the old list continues its tangential history, a new one starts from zero
// (src/CabanaDEM_Force.hpp:180-192), and an old one that no longer overlaps
// is dropped (src/CabanaDEM_Force.hpp:399-433).  The fused step kernel then
// applies stage3 of this step and stage1 + stage2 of the next one, writing
// the new state to the other buffer so neighbours still read the old one.
#pragma once

#include "dem_internal.h"

#include <climits>

namespace demhip {
namespace {

#ifndef DEM_PHASE1_BATCH
#define DEM_PHASE1_BATCH 6
#endif
#ifndef DEM_FORCE_MIN_WAVES
#define DEM_FORCE_MIN_WAVES 1
#endif
// The corrected arithmetic's force kernels would take about 150 VGPRs (3
// waves per SIMD); held to 128 (4 waves, a few spilled values) they are 4 %
// faster on contact-dense beds.  The bonded and sampling variants spill far
// more at 128 and stay unbounded.
#ifndef DEM_FORCE_MIN_WAVES_CORRECTED
#define DEM_FORCE_MIN_WAVES_CORRECTED 4
#endif
constexpr int kPhase1Batch = DEM_PHASE1_BATCH; // neighbour gathers issued together
#ifndef DEM_DETECT_BATCH
#define DEM_DETECT_BATCH 4
#endif
#ifndef DEM_DEFER_MIN_WAVES
#define DEM_DEFER_MIN_WAVES 1
#endif
#ifndef DEM_DEFER_EARLY_LOADS
#define DEM_DEFER_EARLY_LOADS 1 // k_defer loads u, omega, m, I before phase 1
#endif
#ifndef DEM_DEFER_BATCH
#define DEM_DEFER_BATCH 6 // neighbour gathers in flight in k_defer (109 VGPRs, 4 waves/SIMD)
#endif
#ifndef DEM_PRE_MIN_WAVES
#define DEM_PRE_MIN_WAVES 1
#endif
constexpr int kDetectBatch = DEM_DETECT_BATCH; // the same in k_detect

__device__ inline void load6(const double* p, double* v) {
  const double2* q = reinterpret_cast<const double2*>(p);
  double2 a = q[0], b = q[1], c = q[2];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
__device__ inline void store6(double* p, const double* v) {
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(v[0], v[1]);
  q[1] = make_double2(v[2], v[3]);
  q[2] = make_double2(v[4], v[5]);
}

// Particle-particle contact of i with j (overlap > 0 already established).
// h is the tangential history of (i, j), updated in place.  F (and, in
// corrected mode, T) accumulate.  Line numbers: src/CabanaDEM_Force.hpp.
template <bool REF>
__device__ inline void pp_contact(const double4& xi, const double* ui, const double4& mi,
                                  const double4& xj, const double* uj, const double4& mj,
                                  double* h, double* F, double* T, const StepArgs& a,
                                  const SysParams& P) {
  const double p0 = xi.x - xj.x, p1 = xi.y - xj.y, p2 = xi.z - xj.z; // :94-96
  const double r2 = p0 * p0 + p1 * p1 + p2 * p2;                      // :99
  const double rij = sqrt(r2);                                        // :101
  const double overlap = xi.w + xj.w - rij;                           // :107
  const double a_i = xi.w - overlap / 2.;                             // :109
  const double a_j = xj.w - overlap / 2.;                             // :110
  const double nx = p0 / rij, ny = p1 / rij, nz = p2 / rij;           // :113-115
  double vi0, vi1, vi2, vj0, vj1, vj2;
  if (REF) {
    vi0 = ui[0] + (ui[4] * nz - ui[5] * ny) * a_i;                    // :119-126
    vi1 = ui[1] + (ui[5] * nx - ui[3] * nz) * a_i;
    vi2 = ui[2] + (ui[3] * ny - ui[4] * nx) * a_i;
    vj0 = uj[0] + (-uj[4] * nz + uj[5] * ny) * a_j;                   // :130-137 (u_i in y, z: Q1)
    vj1 = ui[1] + (-uj[5] * nx + uj[3] * nz) * a_j;
    vj2 = ui[2] + (-uj[3] * ny + uj[4] * nx) * a_j;
  } else {
    vi0 = ui[0] - (ui[4] * nz - ui[5] * ny) * a_i;
    vi1 = ui[1] - (ui[5] * nx - ui[3] * nz) * a_i;
    vi2 = ui[2] - (ui[3] * ny - ui[4] * nx) * a_i;
    vj0 = uj[0] + (uj[4] * nz - uj[5] * ny) * a_j;
    vj1 = uj[1] + (uj[5] * nx - uj[3] * nz) * a_j;
    vj2 = uj[2] + (uj[3] * ny - uj[4] * nx) * a_j;
  }
  const double v0 = vi0 - vj0, v1 = vi1 - vj1, v2 = vi2 - vj2;      // :141-143
  const double vdn = v0 * nx + v1 * ny + v2 * nz;                   // :146
  const double vn0 = vdn * nx, vn1 = vdn * ny, vn2 = vdn * nz;      // :147-149
  const double vt0 = v0 - vn0, vt1 = v1 - vn1, vt2 = v2 - vn2;      // :152-154
  const double E_eff = 1. / (mi.x + mj.x);                          // :196-198
  const double R_eff = 1. / (mi.z + mj.z);                          // :199-201
  const double kn = 4. / 3. * E_eff * sqrt(R_eff);                  // :203
  const double sRo = sqrt(R_eff * overlap);
  const double S_n = 2. * E_eff * sRo;                              // :211
  const double m_eff = 1. / (mi.w + mj.w);                          // :212-214
  const double eta_n = P.c_eta_pp * sqrt(S_n * m_eff);              // :215
  const double G_eff = 1. / (mi.y + mj.y);                          // :218-220
  const double kt = 8. * G_eff * sRo;                               // :222
  const double eta_t = REF ? 0. : P.c_eta_pp * sqrt(kt * m_eff);    // :224 (5/6 == 0: Q2)
  const double fn = kn * (overlap * sqrt(overlap));                 // :227
  const double fn0 = fn * nx - eta_n * vn0;                         // :228-230
  const double fn1 = fn * ny - eta_n * vn1;
  const double fn2 = fn * nz - eta_n * vn2;
  double tf0 = 0., tf1 = 0., tf2 = 0.;
  const double vmag = sqrt(v0 * v0 + v1 * v1 + v2 * v2);            // :238
  if (vmag < 1e-12) {                                               // :239-243
    h[0] = 0.; h[1] = 0.; h[2] = 0.;
  } else {
    const double tx = v0 - nx * vdn, ty = v1 - ny * vdn, tz = v2 - nz * vdn; // :257-259
    const double tmag = sqrt(tx * tx + ty * ty + tz * tz);          // :261
    double ti0 = 0., ti1 = 0., ti2 = 0.;
    if (tmag > 1e-12) { ti0 = tx / tmag; ti1 = ty / tmag; ti2 = tz / tmag; } // :267-271
    const double d0 = h[0] + v0 * a.dt;                             // :274-276
    const double d1 = h[1] + v1 * a.dt;
    const double d2 = h[2] + v2 * a.dt;
    const double ddt = (d0 * ti0 + d1 * ti1 + d2 * ti2);            // :278-280
    h[0] = ddt * ti0; h[1] = ddt * ti1; h[2] = ddt * ti2;           // :282-284
    const double fs0 = -kt * h[0] - eta_t * vt0;                    // :286-288
    const double fs1 = -kt * h[1] - eta_t * vt1;
    const double fs2 = -kt * h[2] - eta_t * vt2;
    const double ftm = sqrt(fs0 * fs0 + fs1 * fs1 + fs2 * fs2);     // :290
    const double fnm = sqrt(fn0 * fn0 + fn1 * fn1 + fn2 * fn2);     // :291
    double ftms = ftm;                                              // :293-296
    if (ftms > P.fr_pp * fnm) ftms = P.fr_pp * fnm;
    tf0 = -ftms * ti0; tf1 = -ftms * ti1; tf2 = -ftms * ti2;        // :299-301
    const double m0 = -tf0 / kt, m1 = -tf1 / kt, m2 = -tf2 / kt;    // :303-305
    if (REF) {                                                      // :307-319 (Q3)
      const double lm = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
      if (lm > 1e-12) { h[0] = m0 / lm; h[1] = m1 / lm; h[2] = m2 / lm; }
      else { h[0] = 0.; h[1] = 0.; h[2] = 0.; }
    } else {
      // the spring keeps the elastic part of the applied force
      h[0] = -(tf0 + eta_t * vt0) / kt;
      h[1] = -(tf1 + eta_t * vt1) / kt;
      h[2] = -(tf2 + eta_t * vt2) / kt;
    }
  }
  F[0] += fn0 + tf0;                                                // :324-326
  F[1] += fn1 + tf1;
  F[2] += fn2 + tf2;
  if (!REF) {
    T[0] += (ny * tf2 - nz * tf1) * (-a_i);
    T[1] += (nz * tf0 - nx * tf2) * (-a_i);
    T[2] += (nx * tf1 - ny * tf0) * (-a_i);
  }
}

// Particle-wall contact (overlap > 0 already established).  sw is the
// wall tangential history, updated in place.  src/CabanaDEM_Force.hpp.
// wx: the wall point at this force pass (moved in corrected mode).
template <bool REF>
__device__ inline void pw_contact(const double4& xi, const double* ui, const double4& mi,
                                  const WallConst& W, const double* wx, double overlap, double* sw,
                                  double* F, double* T, const StepArgs& a, const SysParams& P) {
  const double nx = W.n[0], ny = W.n[1], nz = W.n[2];               // :509-511
  const double a_i = xi.w - overlap / 2.;                           // :517
  double v0, v1, v2;
  if (REF) {
    v0 = ui[0] + (ui[4] * nz - ui[5] * ny) * a_i;                   // :520-527
    v1 = ui[1] + (ui[5] * nx - ui[3] * nz) * a_i;
    v2 = ui[2] + (ui[3] * ny - ui[4] * nx) * a_i;
    // :529-535: the wall's velocity is taken as 0 (Q7)
  } else {
    v0 = ui[0] - (ui[4] * nz - ui[5] * ny) * a_i;
    v1 = ui[1] - (ui[5] * nx - ui[3] * nz) * a_i;
    v2 = ui[2] - (ui[3] * ny - ui[4] * nx) * a_i;
    // minus the wall surface's velocity at the contact point c = x_i - a_i n
    const double c0 = xi.x - a_i * nx - wx[0];
    const double c1 = xi.y - a_i * ny - wx[1];
    const double c2 = xi.z - a_i * nz - wx[2];
    v0 -= W.u[0] + (W.w[1] * c2 - W.w[2] * c1);
    v1 -= W.u[1] + (W.w[2] * c0 - W.w[0] * c2);
    v2 -= W.u[2] + (W.w[0] * c1 - W.w[1] * c0);
  }
  const double vdn = v0 * nx + v1 * ny + v2 * nz;                   // :538
  const double vn0 = vdn * nx, vn1 = vdn * ny, vn2 = vdn * nz;      // :539-541
  const double vt0 = v0 - vn0, vt1 = v1 - vn1, vt2 = v2 - vn2;      // :544-546
  const double E_eff = 1. / (mi.x + W.a1);                          // :561-563
  const double R_eff = 1. / (mi.z + 0.);                            // :564-566
  const double kn = 4. / 3. * E_eff * sqrt(R_eff);                  // :568
  const double sRo = sqrt(R_eff * overlap);
  const double S_n = 2. * E_eff * sRo;                              // :576
  const double m_eff = 1. / (mi.w + 0.);                            // :577-579
  const double eta_n = P.c_eta_pw * sqrt(S_n * m_eff);              // :580
  const double fn = kn * (overlap * sqrt(overlap));                 // :583
  const double fn0 = fn * nx - eta_n * vn0;                         // :584-586
  const double fn1 = fn * ny - eta_n * vn1;
  const double fn2 = fn * nz - eta_n * vn2;
  double ft0 = 0., ft1 = 0., ft2 = 0.;
  const double vmag = sqrt(v0 * v0 + v1 * v1 + v2 * v2);            // :597
  if (vmag < 1e-12) {                                               // :599-603
    sw[0] = 0.; sw[1] = 0.; sw[2] = 0.;
  } else {
    sw[0] += vt0 * a.dt; sw[1] += vt1 * a.dt; sw[2] += vt2 * a.dt;  // :607-609
    const double G_eff = 1. / (mi.y + W.a2);                        // :612-614
    const double kt = 8. * G_eff * sRo;                             // :616
    const double eta_t = REF ? 0. : P.c_eta_pw * sqrt(kt * m_eff);  // :618 (Q2)
    const double fs0 = -kt * sw[0] - eta_t * vt0;                   // :620-622
    const double fs1 = -kt * sw[1] - eta_t * vt1;
    const double fs2 = -kt * sw[2] - eta_t * vt2;
    const double ftm = REF ? sqrt(fs0 * fs0 + fs1 * fs1 + fs2 * fs1) // :624 (Q6)
                           : sqrt(fs0 * fs0 + fs1 * fs1 + fs2 * fs2);
    double ti0 = 0., ti1 = 0., ti2 = 0.;
    if (ftm > 1e-12) { ti0 = fs0 / ftm; ti1 = fs1 / ftm; ti2 = fs2 / ftm; } // :630-634
    const double fnm = sqrt(fn0 * fn0 + fn1 * fn1 + fn2 * fn2);     // :636
    const double lim = P.fr_pw * fnm;                               // :638 std::min(lim, ftm)
    const double ftms = (ftm < lim) ? ftm : lim;
    ft0 = ftms * ti0; ft1 = ftms * ti1; ft2 = ftms * ti2;           // :641-643
    if (REF) {
      // :646-648 add +eta_t v_t to the applied force (along the slip, quirk
      // Q16; inert here since eta_t = 0), then reset the spring to it
      ft0 += eta_t * vt0; ft1 += eta_t * vt1; ft2 += eta_t * vt2;
      sw[0] = -ft0 / kt; sw[1] = -ft1 / kt; sw[2] = -ft2 / kt;      // :651-653
    } else {
      // the limited force already holds the damping; the spring keeps the
      // elastic part of it
      sw[0] = -(ft0 + eta_t * vt0) / kt;
      sw[1] = -(ft1 + eta_t * vt1) / kt;
      sw[2] = -(ft2 + eta_t * vt2) / kt;
    }
  }
  F[0] += fn0 + ft0;                                                // :657-659
  F[1] += fn1 + ft1;
  F[2] += fn2 + ft2;
  const double arm = REF ? a_i : -a_i;                              // :662-664 (Q5)
  T[0] += (ny * ft2 - nz * ft1) * arm;
  T[1] += (nz * ft0 - nx * ft2) * arm;
  T[2] += (nx * ft1 - ny * ft0) * arm;
}

// Beam bond between i and j (src/CabanaBondedDEM_Force.hpp:86-341; line
// numbers below), with the history h = (ft, mn, mt) and the rest length L0
// of the bond's slot (the reference indexes both by the partner id j, quirk
// Q9).  Returns true, contributing nothing, when the bond's tensile stress
// E_i (r - L0) / (r_i + r_j) exceeds the strength (breakage, an extension).
template <bool REF>
__device__ inline bool bond_force(const double4& xi, const double* ui, double2 EGi,
                                  const double4& xj, const double* uj, double L0, double* h,
                                  double* F, double* T, const StepArgs& a) {
  const double p0 = xi.x - xj.x, p1 = xi.y - xj.y, p2 = xi.z - xj.z; // :106-108
  const double r2 = p0 * p0 + p1 * p1 + p2 * p2;                      // :111
  const double rij = sqrt(r2);                                        // :113
  const double rad_sum = xi.w + xj.w;                                 // :182
  if (a.tensile_strength > 0.) {
    const double stretch = rij - L0;
    if (stretch > 0. && EGi.x * stretch / rad_sum > a.tensile_strength) return true;
  }
  const double g_c = rij - xi.w - xj.w;                               // :115
  const double a_i = xi.w + g_c / 2.;                                 // :116
  const double a_j = xj.w + g_c / 2.;                                 // :117
  const double n0 = p0 / rij, n1 = p1 / rij, n2 = p2 / rij;           // :122
  const double vi0 = ui[0] + (ui[4] * n2 - ui[5] * n1) * a_i;         // :130-137
  const double vi1 = ui[1] + (ui[5] * n0 - ui[3] * n2) * a_i;
  const double vi2 = ui[2] + (ui[3] * n1 - ui[4] * n0) * a_i;
  const double vj0 = uj[0] + (-uj[4] * n2 + uj[5] * n1) * a_j;        // :139-146 (Q1)
  const double vj1 = (REF ? ui[1] : uj[1]) + (-uj[5] * n0 + uj[3] * n2) * a_j;
  const double vj2 = (REF ? ui[2] : uj[2]) + (-uj[3] * n1 + uj[4] * n0) * a_j;
  const double vr0 = vi0 - vj0, vr1 = vi1 - vj1, vr2 = vi2 - vj2;    // :148-150
  const double vdp = vr0 * n0 + vr1 * n1 + vr2 * n2;                  // :153
  const double vn0 = vdp * n0, vn1 = vdp * n1, vn2 = vdp * n2;        // :154-156
  const double vt0 = vr0 - vn0, vt1 = vr1 - vn1, vt2 = vr2 - vn2;    // :158-160
  const double wr0 = ui[3] - uj[3], wr1 = ui[4] - uj[4], wr2 = ui[5] - uj[5]; // :163-165
  const double wdp = wr0 * n0 + wr1 * n1 + wr2 * n2;                  // :167
  const double wn0 = wdp * n0, wn1 = wdp * n1, wn2 = wdp * n2;        // :168
  const double wt0 = wr0 - wn0, wt1 = wr1 - wn1, wt2 = wr2 - wn2;    // :169
  const double beta = a.beta_bond;                                    // :186
  const double rb = (xi.w < xj.w ? xi.w : xj.w) * a.bond_radius_factor; // :188
  const double A_bond = M_PI * (rb * rb);                             // :189
  const double I_bond = M_PI * ((rb * rb) * (rb * rb)) / 4.0;         // :190
  const double I_p_bond = I_bond * 2.0;                               // :191
  const double k_n = EGi.x * A_bond / rad_sum;                        // :195-198
  const double k_t = EGi.y * A_bond / rad_sum;
  const double k_tor = EGi.y * I_p_bond / rad_sum;
  const double k_ben = EGi.x * I_bond / rad_sum;
  const double overlap_n = rij - L0;                                  // :207
  const double fn0 = -k_n * overlap_n * n0;                           // :208-210
  const double fn1 = -k_n * overlap_n * n1;
  const double fn2 = -k_n * overlap_n * n2;
  double mx = (vn0 >= 0.0) ? 1.0 : -1.0;                              // :216-218
  double my = (vn1 >= 0.0) ? 1.0 : -1.0;
  double mz = (vn2 >= 0.0) ? 1.0 : -1.0;
  // :220-222 add +beta|fn| sign(v_n), which feeds energy in (quirk Q15);
  // corrected mode damps, as the tangential and moment terms do (:258-260)
  const double bs = REF ? beta : -beta;
  const double fd0 = fn0 + bs * fabs(fn0) * mx;
  const double fd1 = fn1 + bs * fabs(fn1) * my;
  const double fd2 = fn2 + bs * fabs(fn2) * mz;
  double f0 = h[0], f1 = h[1], f2 = h[2];                             // :231-233
  const double fdp = f0 * n0 + f1 * n1 + f2 * n2;                     // :235-239
  f0 -= fdp * n0; f1 -= fdp * n1; f2 -= fdp * n2;
  double tmp = k_t * a.dt;                                            // :241-246
  f0 += -vt0 * tmp; f1 += -vt1 * tmp; f2 += -vt2 * tmp;
  mx = (vt0 >= 0.0) ? 1.0 : -1.0;                                     // :251-256
  my = (vt1 >= 0.0) ? 1.0 : -1.0;
  mz = (vt2 >= 0.0) ? 1.0 : -1.0;
  h[0] = f0 - beta * fabs(f0) * mx;                                   // :258-260
  h[1] = f1 - beta * fabs(f1) * my;
  h[2] = f2 - beta * fabs(f2) * mz;
  double q0 = h[3], q1 = h[4], q2 = h[5];                             // :267-274
  double tdp = q0 * n0 + q1 * n1 + q2 * n2;
  q0 = tdp * n0; q1 = tdp * n1; q2 = tdp * n2;
  tmp = k_tor * a.dt;                                                 // :276-280
  q0 += -wn0 * tmp; q1 += -wn1 * tmp; q2 += -wn2 * tmp;
  mx = (wn0 >= 0.0) ? 1.0 : -1.0;                                     // :282-287
  my = (wn1 >= 0.0) ? 1.0 : -1.0;
  mz = (wn2 >= 0.0) ? 1.0 : -1.0;
  h[3] = q0 - beta * fabs(q0) * mx;                                   // :289-291
  h[4] = q1 - beta * fabs(q1) * my;
  h[5] = q2 - beta * fabs(q2) * mz;
  double t0 = h[6], t1 = h[7], t2 = h[8];                             // :298-304
  tdp = t0 * n0 + t1 * n1 + t2 * n2;
  t0 -= tdp * n0; t1 -= tdp * n1; t2 -= tdp * n2;
  tmp = k_ben * a.dt;                                                 // :306-310
  t0 += -wt0 * tmp; t1 += -wt1 * tmp; t2 += -wt2 * tmp;
  mx = (wt0 >= 0.0) ? 1.0 : -1.0;                                     // :313-318
  my = (wt1 >= 0.0) ? 1.0 : -1.0;
  mz = (wt2 >= 0.0) ? 1.0 : -1.0;
  h[6] = t0 - beta * fabs(t0) * mx;                                   // :320-322
  h[7] = t1 - beta * fabs(t1) * my;
  h[8] = t2 - beta * fabs(t2) * mz;
  F[0] += fd0 + h[0];                                                 // :334-336
  F[1] += fd1 + h[1];
  F[2] += fd2 + h[2];
  T[0] += h[3] + h[6];                                                // :338-340 (the cross
  T[1] += h[4] + h[7];                                                //  term is cri * 0)
  T[2] += h[5] + h[8];
  return false;
}

__device__ inline bool bonded_to(const StepArgs& a, int i, int nb, int j) {
  for (int b = 0; b < nb; ++b)
    if (a.bj[(long)b * a.ld + i] == j) return true;
  return false;
}

// u = v_g, omega = 0 for particles of a prescribed-velocity group.
__device__ inline void apply_vgroup(const StepArgs& a, int i, double* s) {
  if (!a.vgroup) return;
  const int g = a.vgroup[i];
  if (g <= 0 || g > a.n_vgroups || a.vgkind[g - 1] != DEM_GROUP_VELOCITY) return;
  s[0] = a.vgv[g - 1][0]; s[1] = a.vgv[g - 1][1]; s[2] = a.vgv[g - 1][2];
  s[3] = 0.; s[4] = 0.; s[5] = 0.;
}

// F = 0 for particles of a fixed-force group, after the whole force pass
// (examples/tst_01_BDEM_two_particles_tensile_3D.cpp:57-59, applied between
// the bond forces and stage3, :306-309).
__device__ inline void apply_force_bc(const StepArgs& a, int i, double* F) {
  if (!a.vgroup) return;
  const int g = a.vgroup[i];
  if (g <= 0 || g > a.n_vgroups || a.vgkind[g - 1] != DEM_GROUP_FIXED_FORCE) return;
  F[0] = 0.; F[1] = 0.; F[2] = 0.;
}

// Initial force of the force pass: m g (the body-force extension), then the
// external force (examples/tst_01_BDEM_two_particles_tensile_3D.cpp:110-112).
__device__ inline void initial_force(const StepArgs& a, int i, double m, double* F) {
  if (a.gravity_on) { F[0] = m * a.g[0]; F[1] = m * a.g[1]; F[2] = m * a.g[2]; }
  else { F[0] = 0.; F[1] = 0.; F[2] = 0.; }
  if (a.fext) {
    F[0] += a.fext[3l * i + 0];
    F[1] += a.fext[3l * i + 1];
    F[2] += a.fext[3l * i + 2];
  }
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Fixed-order block reduction of 8 doubles into part[blk][8] (B threads).
template <int B>
__device__ inline void block_reduce8(const double* v, double* part, int blk) {
  __shared__ double s_red[B / 64][8];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double r[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 8; ++k) s_red[wid][k] = r[k];
  __syncthreads();
  if (threadIdx.x < 8) {
    double s = 0.;
    for (int w = 0; w < B / 64; ++w) s += s_red[w][threadIdx.x];
    part[(long)blk * 8 + threadIdx.x] = s;
  }
}

// The displacement check's terms of one written particle (automatic
// cadence): |x - xb|^2 and |u|^2 of the state this pass writes.
__device__ __forceinline__ void disp_terms(const StepArgs& a, int i, const double4& p, const double* u,
                                           double& d2, double& v2) {
  const double d0 = p.x - a.xb[3 * i + 0], d1 = p.y - a.xb[3 * i + 1], dz = p.z - a.xb[3 * i + 2];
  d2 = max_nan(0., d0 * d0 + d1 * d1 + dz * dz);
  v2 = max_nan(0., u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
}

// The tile's maxima into out[0], out[1] (bits); the atomic is skipped when
// the value already there is at least as large (the maxima only grow).
// Every lane of the workgroup must call this (barriers).
template <int B>
__device__ inline void disp_block_max(double d2, double v2, unsigned long long* out) {
  __shared__ double s_d[B / 64], s_v[B / 64];
  for (int off = 32; off > 0; off >>= 1) {
    d2 = max_nan(d2, __shfl_xor(d2, off));
    v2 = max_nan(v2, __shfl_xor(v2, off));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_d[w] = d2;
    s_v[w] = v2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < B / 64; ++k) {
      d2 = max_nan(d2, s_d[k]);
      v2 = max_nan(v2, s_v[k]);
    }
    const unsigned long long bd = (unsigned long long)__double_as_longlong(d2);
    const unsigned long long bv = (unsigned long long)__double_as_longlong(v2);
    if (bd > __hip_atomic_load(&out[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&out[0], bd);
    if (bv > __hip_atomic_load(&out[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&out[1], bv);
  }
  __syncthreads(); // the next tile of a persistent grid rewrites s_d, s_v
}

// A wave's own LDS columns: its writes before its later reads, and back.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bytes of dynamic LDS per workgroup of B threads: the contact ids found in
// phase 1 (int [max_c][B]), for the wave/tile pass also the old slot of each
// (byte [max_c][B], padded to 8), the contact shares (double [B][6]) and the
// lanes' first contact numbers (int [B]).
__host__ __device__ inline size_t force_lds_bytes(int max_c, bool wave, int B) {
  const size_t mc = (size_t)(max_c > 0 ? max_c : 1);
  size_t b = mc * B * sizeof(int);
  if (wave) b += (mc * B + 7) / 8 * 8 + (size_t)B * 6 * sizeof(double) + (size_t)B * sizeof(int);
  return b;
}

// Particle-particle contact arithmetic spread over the lanes of a wave.
// Lane l owns nc_l contacts: ids ascending in its s_cj column (with PRE in
// its cj_out row), old history slot + 1 in its s_op column (0: a new
// contact).  The wave numbers its contacts owner-major (an exclusive scan of
// nc) and, in rounds of 64, every lane takes one: it gathers both particles,
// runs the contact arithmetic into zero accumulators of negative sign (so
// each share is exactly the term the owner would add), writes the contact's
// id and history into the owner's rows, and leaves the owner's share of F
// and T in LDS.  Owners add their shares in contact order: the same sums in
// the same order as one lane walking its own list, so the result is bitwise
// the same, while the FP64 block no longer waits on the lane with the most
// contacts.  Every lane of the wave must call this (shuffles).
template <bool REF, bool PRE, int B>
__device__ __forceinline__ void wave_contacts(const StepArgs& a, const SysParams& P, const int first,
                                              const int nc, const int* s_cj, const unsigned char* s_op,
                                              double* s_con, double* F, double* T) {
  const int tid = threadIdx.x, lane = tid & 63, wbase = tid - lane;
  const long ld = a.ld;
  int incl = nc;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  const int excl = incl - nc;
  const int total = __shfl(incl, 63);
  double* con = s_con + (long)wbase * 6; // [64][6] shares of this wave's round
  for (int base = 0; base < total; base += 64) {
    const int k = base + lane;
    // owner: the last lane whose first contact number is <= k
    int L = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
      const int e = __shfl(excl, L + st);
      if (e <= k) L += st;
    }
    const int eL = __shfl(excl, L);
    if (k < total) {
      const int m = k - eL;
      const int o = wbase + L; // the owner's thread in the workgroup
      const int io = first + o;
      const int j = PRE ? a.cj_out[m * ld + io] : s_cj[m * B + o];
      const int p1 = s_op[m * B + o];
      double h[3] = {0., 0., 0.};
      if (p1 > 0) { // continued contact: src/CabanaDEM_Force.hpp:166-178
        h[0] = a.ch_in[(3 * (p1 - 1) + 0) * ld + io];
        h[1] = a.ch_in[(3 * (p1 - 1) + 1) * ld + io];
        h[2] = a.ch_in[(3 * (p1 - 1) + 2) * ld + io];
      }
      const double4 xo = a.xr_in[io];
      double uo[6];
      load6(a.uw_in + 6l * io, uo);
      const double4 mo = a.mat[io];
      const double4 xj = a.xr_in[j];
      double uj[6];
      load6(a.uw_in + 6l * j, uj);
      const double4 mj = a.mat[j];
      double Fc[3] = {-0., -0., -0.}, Tc[3] = {-0., -0., -0.};
      pp_contact<REF>(xo, uo, mo, xj, uj, mj, h, Fc, Tc, a, P);
      if (!PRE) a.cj_out[m * ld + io] = j;
      a.ch_out[(3 * m + 0) * ld + io] = h[0];
      a.ch_out[(3 * m + 1) * ld + io] = h[1];
      a.ch_out[(3 * m + 2) * ld + io] = h[2];
      double* c = con + lane * 6;
      c[0] = Fc[0]; c[1] = Fc[1]; c[2] = Fc[2];
      c[3] = Tc[0]; c[4] = Tc[1]; c[5] = Tc[2];
    }
    wave_lds_sync();
    const int q0 = excl > base ? excl : base;
    const int q1 = incl < base + 64 ? incl : base + 64;
    for (int q = q0; q < q1; ++q) {
      const double* c = con + (q - base) * 6;
      F[0] += c[0]; F[1] += c[1]; F[2] += c[2];
      if (!REF) { T[0] += c[3]; T[1] += c[4]; T[2] += c[5]; }
    }
    wave_lds_sync(); // the next round rewrites the shares
  }
}

// The same over a whole tile of B > 64 lanes: the tile numbers its contacts
// owner-major (a wave scan, then the waves' totals through LDS) and runs them
// in rounds of B, every lane taking one; owners add their shares in contact
// order, as above.  With few contacts per tile (a sparse bed) one wave of
// the tile runs them all and the others skip the FP64 block, where a lane
// walk or a wave pass has every wave holding a contact run it.  Every lane
// of the workgroup must call this (barriers).
template <bool REF, bool PRE, int B>
__device__ __forceinline__ void tile_contacts(const StepArgs& a, const SysParams& P, const int first,
                                              const int nc, const int* s_cj, const unsigned char* s_op,
                                              double* s_con, int* s_ex, double* F, double* T) {
  if constexpr (B == 64) {
    wave_contacts<REF, PRE, B>(a, P, first, nc, s_cj, s_op, s_con, F, T);
  } else {
    constexpr int NW = B / 64;
    __shared__ int s_wt[NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long ld = a.ld;
    int incl = nc;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off);
      if (lane >= off) incl += t;
    }
    if (lane == 63) s_wt[wv] = incl;
    __syncthreads(); // also orders every lane's s_cj / s_op columns before the rounds
    int wb = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int t = s_wt[w];
      wb += w < wv ? t : 0;
      total += t;
    }
    incl += wb;
    const int excl = incl - nc;
    s_ex[tid] = excl;
    __syncthreads();
    for (int base = 0; base < total; base += B) {
      const int k = base + tid;
      if (k < total) {
        // owner: the last lane whose first contact number is <= k
        int L = 0;
#pragma unroll
        for (int st = B / 2; st > 0; st >>= 1)
          if (s_ex[L + st] <= k) L += st;
        const int m = k - s_ex[L];
        const int io = first + L;
        const int j = PRE ? a.cj_out[m * ld + io] : s_cj[m * B + L];
        const int p1 = s_op[m * B + L];
        double h[3] = {0., 0., 0.};
        if (p1 > 0) { // continued contact: src/CabanaDEM_Force.hpp:166-178
          h[0] = a.ch_in[(3 * (p1 - 1) + 0) * ld + io];
          h[1] = a.ch_in[(3 * (p1 - 1) + 1) * ld + io];
          h[2] = a.ch_in[(3 * (p1 - 1) + 2) * ld + io];
        }
        const double4 xo = a.xr_in[io];
        double uo[6];
        load6(a.uw_in + 6l * io, uo);
        const double4 mo = a.mat[io];
        const double4 xj = a.xr_in[j];
        double uj[6];
        load6(a.uw_in + 6l * j, uj);
        const double4 mj = a.mat[j];
        double Fc[3] = {-0., -0., -0.}, Tc[3] = {-0., -0., -0.};
        pp_contact<REF>(xo, uo, mo, xj, uj, mj, h, Fc, Tc, a, P);
        if (!PRE) a.cj_out[m * ld + io] = j;
        a.ch_out[(3 * m + 0) * ld + io] = h[0];
        a.ch_out[(3 * m + 1) * ld + io] = h[1];
        a.ch_out[(3 * m + 2) * ld + io] = h[2];
        double* c = s_con + (long)tid * 6;
        c[0] = Fc[0]; c[1] = Fc[1]; c[2] = Fc[2];
        c[3] = Tc[0]; c[4] = Tc[1]; c[5] = Tc[2];
      }
      __syncthreads();
      const int q0 = excl > base ? excl : base;
      const int q1 = incl < base + B ? incl : base + B;
      for (int q = q0; q < q1; ++q) {
        const double* c = s_con + (long)(q - base) * 6;
        F[0] += c[0]; F[1] += c[1]; F[2] += c[2];
        if (!REF) { T[0] += c[3]; T[1] += c[4]; T[2] += c[5]; }
      }
      __syncthreads(); // the next round rewrites the shares
    }
  }
}

// KICK: 0 = forces only (write F,T); 1 = + stage3 (write state, F,T);
//       2 = + stage3 + stage1 + stage2 of the next step (write state).
// One tile: B consecutive slots of one system.  ENS: an ensemble (tile
// table and per-system parameters from the device); otherwise one system
// whose parameters travel in the kernel arguments.  Each lane walks its own
// contact list, or (a.wave, systems without bonds) the tile runs the contact
// arithmetic of all its lanes spread over the lanes (tile_contacts).
// PASS (systems without bonds, whole force passes, no sampling):
//   0  everything in one pass;
//   1  deferral: phase 1 only; a particle without particle or wall contacts
//      finishes here (its force is the start value), the others write their
//      contact ids to cj_out and their count to cmeta_out and join their
//      tile's work list (slots [t B, t B + work_tile[t]) of a.work, in lane
//      order: no atomics); the FP64 contact arithmetic is not compiled in, so
//      this pass runs at the occupancy of a gather kernel;
//   2  the work lists: tile t runs the particles of its list, as with PRE
//      (ids in cj_out); a tile with an empty list returns at once.
// Every particle's arithmetic is that of PASS 0, so the results are bitwise
// the same.
template <bool REF, bool PP, bool PW, int KICK, bool ACCUM, bool DIAG, bool BOND, bool ENS, bool PRE, int B,
          int PASS = 0>
__device__ __forceinline__ void force_tile(const StepArgs& a, const int tile, int* s_cj) {
  constexpr bool DEFER = PASS == 1, LIST = PASS == 2;
  static_assert(PASS == 0 || (PP && PW && !ACCUM && !DIAG && !BOND), "deferral: whole passes only");
  static_assert(!LIST || PRE, "the work-list pass reads its contact ids from cj_out");
  constexpr bool WAVE = PP && !BOND && PASS == 0;
  const int tid = threadIdx.x;
  int sys = 0, first = tile * B, end = a.n;
  int i;
  bool act;
  if (LIST) {
    const int cnt = a.work_tile[tile];
    if (cnt == 0) return; // uniform over the workgroup
    act = tid < cnt;
    i = act ? a.work[(long)tile * B + tid] : 0;
    if (ENS) sys = a.tiles[tile].x;
  } else {
    if (ENS) {
      const int4 t = a.tiles[tile];
      sys = t.x; first = t.y; end = t.z;
    }
    i = first + tid;
    act = i < end;
  }
  const SysParams& P = ENS ? a.sys[sys] : a.sys0;
  const long ld = a.ld;
  unsigned char* s_op = reinterpret_cast<unsigned char*>(s_cj + (long)(a.max_c > 0 ? a.max_c : 1) * B);
  double* s_con = reinterpret_cast<double*>(
      reinterpret_cast<char*>(s_cj) + force_lds_bytes(a.max_c, false, B) +
      ((size_t)(a.max_c > 0 ? a.max_c : 1) * B + 7) / 8 * 8);
  double dg[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double cd2 = 0., cv2 = 0.; // displacement check terms (a.disp)
  double4 xi = make_double4(0., 0., 0., 0.);
  double ui[6] = {0., 0., 0., 0., 0., 0.};
  double2 mI = make_double2(1., 1.); // m, I
  // (1-nu^2)/E, (2-nu)/G, 1/r, 1/m of i: read once, on the first contact
  double4 mti = make_double4(0., 0., 0., 0.);
  bool have_mti = false;
  double F[3] = {0., 0., 0.}, T[3] = {0., 0., 0.};
  int meta = 0, nnb = 0, nc = 0, nwc = 0;
  // in a bonded run, contacts only between unbonded pairs (or none at all)
  const bool pp_on = !a.nbond || a.unbonded_contacts;
  if (act) {
    xi = a.xr_in[i];
    // loaded here, in flight over phase 1 (the deferral pass too, unless
    // DEM_DEFER_EARLY_LOADS is 0: fewer registers, one more round trip)
    if (!DEFER || DEM_DEFER_EARLY_LOADS) {
      load6(a.uw_in + 6l * i, ui);
      mI = a.mI[i];
    }
    if (DEFER) {
    } else if (ACCUM) {
      double ft[6];
      load6(a.ft + 6l * i, ft);
      F[0] = ft[0]; F[1] = ft[1]; F[2] = ft[2];
      T[0] = ft[3]; T[1] = ft[4]; T[2] = ft[5];
    } else {
      initial_force(a, i, mI.x, F);
    }
    int nb = 0;
    if (BOND) {
      // bond forces first, in slot order (the contacts follow)
      nb = a.nbond[i];
      const double2 EGi = a.bEG[i];
      for (int b = 0; b < nb; ++b) {
        const int j = a.bj[(long)b * ld + i];
        if (j < 0) continue;
        double h[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) h[c] = a.bh[(9 * b + c) * ld + i];
        const double4 xj = a.xr_in[j];
        double uj[6];
        load6(a.uw_in + 6l * j, uj);
        if (bond_force<REF>(xi, ui, EGi, xj, uj, a.bl[(long)b * ld + i], h, F, T, a)) {
          a.bj[(long)b * ld + i] = -1;
          atomicAdd(&a.flags->bonds_broken, 1);
          continue;
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) a.bh[(9 * b + c) * ld + i] = h[c];
      }
    }
    meta = a.cmeta_in[i];
    nc = meta_count(meta);
    const int nbk = a.nbond ? (BOND ? nb : a.nbond[i]) : 0;
    if (PP && !pp_on) nc = 0;
    if (PP && pp_on && PRE) {
      // Phase 1 ran in k_detect: the count is in cmeta_out, the ids (ascending)
      // already in cj_out.
      if (DIAG) nnb = a.nn[i];
      nc = meta_count(a.cmeta_out[i]);
    } else if (PP && pp_on) {
      // Phase 1: contact detection over the neighbour row.
      constexpr int NB = DEFER ? DEM_DEFER_BATCH : kPhase1Batch;
      // The first batch's row entries are loaded together with the count
      // instead of after it: entries below the row capacity max_n are in
      // bounds whatever the count, and the count masks them afterwards.
      int jr0[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) jr0[q] = (q < a.max_n) ? a.nbr[q * ld + i] : i;
      nnb = a.nn[i];
      nc = 0;
      // Batches of NB independent row loads, then gathers, then
      // tests: several gathers in flight per lane instead of one chain.
      for (int k0 = 0; k0 < nnb; k0 += NB) {
        int jb[NB];
        double4 xb[NB];
        if (k0 == 0) {
#pragma unroll
          for (int q = 0; q < NB; ++q) jb[q] = (q < nnb) ? jr0[q] : i;
        } else {
#pragma unroll
          for (int q = 0; q < NB; ++q) jb[q] = (k0 + q < nnb) ? a.nbr[(k0 + q) * ld + i] : i;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) xb[q] = a.xr_in[jb[q]];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          if (k0 + q >= nnb) break;
          const double4 xj = xb[q];
          const double d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          const double s = xi.w + xj.w;
          // r2 < s^2 is necessary for sqrt(r2) < s; the 1e-12 margin covers
          // the rounding of s*s, so no contact is ever filtered out here.
          if (r2 < s * s * (1. + 1e-12)) {
            if (s - sqrt(r2) > 0. && !(nbk > 0 && bonded_to(a, i, nbk, jb[q]))) {
              if (nc < a.max_c) {
                if (DEFER) a.cj_out[nc * ld + i] = jb[q];
                else s_cj[nc * B + tid] = jb[q];
              }
              ++nc;
            }
          }
        }
      }
      if (nc > a.max_c) {
        atomicMax(&a.flags->contact_overflow, nc);
        nc = a.max_c;
      }
    }
    if (WAVE && pp_on && a.wave) {
      // merge with last step's tracked list (both ascending in j): the old
      // slot of each contact, or 0 for a new one
      // (src/CabanaDEM_Force.hpp:166-192, 399-433)
      const int oc = meta_count(meta);
      int p = 0;
      int jo = oc > 0 ? a.cj_in[i] : INT_MAX;
      for (int m = 0; m < nc; ++m) {
        const int j = PRE ? a.cj_out[m * ld + i] : s_cj[m * B + tid];
        while (jo < j) {
          ++p;
          jo = (p < oc) ? a.cj_in[p * ld + i] : INT_MAX;
        }
        s_op[m * B + tid] = (unsigned char)(jo == j ? p + 1 : 0);
      }
    }
  }
  const bool wave = WAVE && pp_on && a.wave; // uniform over the launch
  if (wave) {
    // every lane of the tile, active or not: the contacts of the tile
    wave_lds_sync();
    tile_contacts<REF, PRE, B>(a, P, first, act ? nc : 0, s_cj, s_op, s_con,
                               reinterpret_cast<int*>(s_con + (long)B * 6), F, T);
  } else if (!DEFER && PP && pp_on && act) {
    // Phase 2 walking each lane's own list:
    // contact arithmetic, merged with last step's tracked list.
    const int oc = meta_count(meta);
    int p = 0;
    int jo = oc > 0 ? a.cj_in[i] : INT_MAX;
    if (nc > 0) { mti = a.mat[i]; have_mti = true; }
    for (int m = 0; m < nc; ++m) {
      const int j = PRE ? a.cj_out[m * ld + i] : s_cj[m * B + tid];
      while (jo < j) {
        ++p;
        jo = (p < oc) ? a.cj_in[p * ld + i] : INT_MAX;
      }
      double h[3] = {0., 0., 0.};
      if (jo == j) {
        h[0] = a.ch_in[(3 * p + 0) * ld + i];
        h[1] = a.ch_in[(3 * p + 1) * ld + i];
        h[2] = a.ch_in[(3 * p + 2) * ld + i];
      }
      const double4 xj = a.xr_in[j];
      double uj[6];
      load6(a.uw_in + 6l * j, uj);
      const double4 mj = a.mat[j];
      pp_contact<REF>(xi, ui, mti, xj, uj, mj, h, F, T, a, P);
      if (!PRE) a.cj_out[m * ld + i] = j;
      a.ch_out[(3 * m + 0) * ld + i] = h[0];
      a.ch_out[(3 * m + 1) * ld + i] = h[1];
      a.ch_out[(3 * m + 2) * ld + i] = h[2];
    }
  }
  bool defer = false;
  if (DEFER) {
    if (act) {
      defer = nc > 0;
      for (int w = 0; w < P.n_walls && !defer; ++w) {
        const WallConst& W = P.walls[w];
        double wx[3] = {W.x[0], W.x[1], W.x[2]};
        if (!REF)
          for (int d = 0; d < 3; ++d) wx[d] += W.u[d] * a.tw;
        const double q0 = xi.x - wx[0], q1 = xi.y - wx[1], q2 = xi.z - wx[2];
        const double tmp = q0 * W.n[0] + q1 * W.n[1] + q2 * W.n[2];
        defer = xi.w - tmp > 0.;
      }
      if (defer) a.cmeta_out[i] = make_meta(nc, 0);
    }
    // the tile's deferred particles into its work list, in lane order
    __shared__ int s_wc[B / 64];
    const unsigned long long mask = __ballot(defer);
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) s_wc[wv] = __popcll(mask);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < B / 64; ++k) {
      base += k < wv ? s_wc[k] : 0;
      total += s_wc[k];
    }
    if (defer) a.work[(long)tile * B + base + __popcll(mask & ((1ull << lane) - 1ull))] = i;
    if (tid == 0) a.work_tile[tile] = total;
  }
  if (DEFER && act && !defer) {
    if (!DEM_DEFER_EARLY_LOADS) {
      load6(a.uw_in + 6l * i, ui);
      mI = a.mI[i];
    }
    initial_force(a, i, mI.x, F);
  }
  if (act && !defer) {
    int wn = meta_walls(meta);
    if (PW && !DEFER) {
      const int wm = wn;
      wn = 0;
      for (int w = 0; w < P.n_walls; ++w) {
        const WallConst& W = P.walls[w];
        // the reference never moves a wall; corrected mode translates it
        double wx[3] = {W.x[0], W.x[1], W.x[2]};
        if (!REF)
          for (int d = 0; d < 3; ++d) wx[d] += W.u[d] * a.tw;
        const double q0 = xi.x - wx[0], q1 = xi.y - wx[1], q2 = xi.z - wx[2];    // :497-499
        const double tmp = q0 * W.n[0] + q1 * W.n[1] + q2 * W.n[2];              // :514
        const double overlap = xi.w - tmp;                                        // :515
        if (overlap > 0.) {                                                       // :556
          if (!have_mti) { mti = a.mat[i]; have_mti = true; }
          double sw[3] = {0., 0., 0.};
          if (wm & (1 << w)) {
            sw[0] = a.sw[(3 * w + 0) * ld + i];
            sw[1] = a.sw[(3 * w + 1) * ld + i];
            sw[2] = a.sw[(3 * w + 2) * ld + i];
          }
          pw_contact<REF>(xi, ui, mti, W, wx, overlap, sw, F, T, a, P);
          a.sw[(3 * w + 0) * ld + i] = sw[0];
          a.sw[(3 * w + 1) * ld + i] = sw[1];
          a.sw[(3 * w + 2) * ld + i] = sw[2];
          wn |= 1 << w;
          ++nwc;
        }
        // no overlap: the history is zero (:666-670); the cleared bit says so
      }
    }
    if (DEFER) wn = 0; // no wall overlaps: any old history is dropped
    if (PP) a.cmeta_out[i] = make_meta(nc, wn);
    else if (PW) a.cmeta_in[i] = make_meta(nc, wn); // walls only: in place
    if (!ACCUM) apply_force_bc(a, i, F);
    if (KICK == 0) {
      double ft[6] = {F[0], F[1], F[2], T[0], T[1], T[2]};
      store6(a.ft + 6l * i, ft);
      if (DIAG) {
        dg[0] = 0.5 * mI.x * (ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2]);
        dg[1] = 0.5 * mI.y * (ui[3] * ui[3] + ui[4] * ui[4] + ui[5] * ui[5]);
        dg[2] = mI.x * ui[0]; dg[3] = mI.x * ui[1]; dg[4] = mI.x * ui[2];
      }
    } else {
      // 1/m and 1/I per use, src/CabanaDEM_Integrator.hpp:42,48
      const double minv = 1. / mI.x, Iinv = 1. / mI.y;
      double s[6];
      // stage3: src/CabanaDEM_Integrator.hpp:89-98
      s[0] = ui[0] + a.hdt * F[0] * minv;
      s[1] = ui[1] + a.hdt * F[1] * minv;
      s[2] = ui[2] + a.hdt * F[2] * minv;
      s[3] = ui[3] + a.hdt * T[0] * Iinv;
      s[4] = ui[4] + a.hdt * T[1] * Iinv;
      s[5] = ui[5] + a.hdt * T[2] * Iinv;
      apply_vgroup(a, i, s);
      if (DIAG) {
        dg[0] = 0.5 * mI.x * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        dg[1] = 0.5 * mI.y * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
        dg[2] = mI.x * s[0]; dg[3] = mI.x * s[1]; dg[4] = mI.x * s[2];
      }
      if (KICK == 2) {
        // stage1 of the next step (:42-51), then stage2 (:67-69)
        for (int d = 0; d < 3; ++d) s[d] = s[d] + a.hdt * F[d] * minv;
        for (int d = 0; d < 3; ++d) s[3 + d] = s[3 + d] + a.hdt * T[d] * Iinv;
        apply_vgroup(a, i, s);
        double4 xo;
        xo.x = xi.x + a.dt * s[0];
        xo.y = xi.y + a.dt * s[1];
        xo.z = xi.z + a.dt * s[2];
        xo.w = xi.w;
        a.xr_out[i] = xo;
        store6(a.uw_out + 6l * i, s);
        if (a.disp) disp_terms(a, i, xo, s, cd2, cv2);
      } else {
        a.xr_out[i] = xi;
        store6(a.uw_out + 6l * i, s);
        double ft[6] = {F[0], F[1], F[2], T[0], T[1], T[2]};
        store6(a.ft + 6l * i, ft);
        if (a.disp) disp_terms(a, i, xi, s, cd2, cv2);
      }
    }
    dg[5] = (double)nnb;
    dg[6] = (double)nc;
    dg[7] = (double)nwc;
  }
  if (DIAG) block_reduce8<B>(dg, a.diag_part, tile);
  if (KICK != 0 && a.disp) disp_block_max<B>(cd2, cv2, a.disp);
}

// The fused force / step kernel.  A grid of one workgroup per tile, or a
// persistent grid that strides over the tiles; either way tile v is mapped
// XCD-aware (contiguous tiles per XCD when the grid is a multiple of 8).
template <bool REF, bool PP, bool PW, int KICK, bool ACCUM, bool DIAG, bool BOND, bool ENS, bool PRE, int B>
__global__ __launch_bounds__(B, PRE ? DEM_PRE_MIN_WAVES
                                 : (!REF && !BOND && !DIAG) ? DEM_FORCE_MIN_WAVES_CORRECTED : DEM_FORCE_MIN_WAVES)
void k_force(StepArgs a) {
  extern __shared__ int s_cj[]; // [max_c][B] contacts found in phase 1
  if ((int)gridDim.x >= a.ntiles) {
    // one workgroup per tile: no loop, so the compiler keeps the kernel
    // arguments where they are used instead of live across a loop
    force_tile<REF, PP, PW, KICK, ACCUM, DIAG, BOND, ENS, PRE, B>(a, xcd_block(blockIdx.x, a.ntiles), s_cj);
    return;
  }
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x) {
    force_tile<REF, PP, PW, KICK, ACCUM, DIAG, BOND, ENS, PRE, B>(a, xcd_block(v, a.ntiles), s_cj);
    if (DIAG) __syncthreads(); // the reduction's LDS is reused by the next tile
  }
}

// The split force pass without sampling: k_defer (PASS 1) over all tiles,
// then k_force_list (PASS 2) over the same tiles, each running only its
// particles with contacts (on the same launch grid).
template <bool REF, int KICK, bool ENS, int B>
__global__ __launch_bounds__(B, DEM_DEFER_MIN_WAVES) void k_defer(StepArgs a) {
  if ((int)gridDim.x >= a.ntiles) {
    force_tile<REF, true, true, KICK, false, false, false, ENS, false, B, 1>(
        a, xcd_block(blockIdx.x, a.ntiles), nullptr);
    return;
  }
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x) {
    force_tile<REF, true, true, KICK, false, false, false, ENS, false, B, 1>(a, xcd_block(v, a.ntiles), nullptr);
    __syncthreads(); // the list compaction's LDS is reused by the next tile
  }
}

template <bool REF, int KICK, bool ENS, int B>
__global__ __launch_bounds__(B, DEM_PRE_MIN_WAVES) void k_force_list(StepArgs a) {
  if ((int)gridDim.x >= a.ntiles) {
    force_tile<REF, true, true, KICK, false, false, false, ENS, true, B, 2>(
        a, xcd_block(blockIdx.x, a.ntiles), nullptr);
    return;
  }
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x)
    force_tile<REF, true, true, KICK, false, false, false, ENS, true, B, 2>(a, xcd_block(v, a.ntiles), nullptr);
}

// Phase 1 of the force pass on its own (the split force pass): for each
// particle, the neighbours j of its row that overlap it
// (src/CabanaDEM_Force.hpp:107,162), ascending in j, into cj_out and their
// count into cmeta_out (wall bits 0: the force kernel writes the final
// word).  The same test and order as phase 1 of force_tile, so the lists
// are identical; this kernel holds no FP64 contact arithmetic, so it runs at
// a much higher occupancy than the fused kernel, and the force kernel that
// follows (k_force with PRE) only walks the short lists.  Systems without
// bonds only.
template <bool ENS, int B>
__device__ __forceinline__ void detect_tile(const StepArgs& a, const int tile) {
  const int tid = threadIdx.x;
  int first = tile * B, end = a.n;
  if (ENS) {
    const int4 t = a.tiles[tile];
    first = t.y; end = t.z;
  }
  const int i = first + tid;
  if (i >= end) return;
  const long ld = a.ld;
  const double4 xi = a.xr_in[i];
  const int nnb = a.nn[i];
  int nc = 0;
  for (int k0 = 0; k0 < nnb; k0 += kDetectBatch) {
    int jb[kDetectBatch];
    double4 xb[kDetectBatch];
#pragma unroll
    for (int q = 0; q < kDetectBatch; ++q)
      jb[q] = (k0 + q < nnb) ? a.nbr[(k0 + q) * ld + i] : i;
#pragma unroll
    for (int q = 0; q < kDetectBatch; ++q) xb[q] = a.xr_in[jb[q]];
#pragma unroll
    for (int q = 0; q < kDetectBatch; ++q) {
      if (k0 + q >= nnb) break;
      const double4 xj = xb[q];
      const double d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double s = xi.w + xj.w;
      // as in force_tile: the margin never filters out a contact
      if (r2 < s * s * (1. + 1e-12) && s - sqrt(r2) > 0.) {
        if (nc < a.max_c) a.cj_out[nc * ld + i] = jb[q];
        ++nc;
      }
    }
  }
  if (nc > a.max_c) {
    atomicMax(&a.flags->contact_overflow, nc);
    nc = a.max_c;
  }
  a.cmeta_out[i] = make_meta(nc, 0);
}

template <bool ENS, int B>
__global__ __launch_bounds__(B, 8) void k_detect(StepArgs a) {
  if ((int)gridDim.x >= a.ntiles) {
    detect_tile<ENS, B>(a, xcd_block(blockIdx.x, a.ntiles));
    return;
  }
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x) detect_tile<ENS, B>(a, xcd_block(v, a.ntiles));
}

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// The split force pass: k_detect, then the force kernel over its lists.
// Whole force passes of systems without bonds only; the result is bitwise
// that of the fused kernel.
template <bool REF, bool ENS, int B>
hipError_t launch_split(ForceKind kind, const StepArgs& a, bool sample, const dim3& g, hipStream_t s) {
  const dim3 b(B);
  const size_t lds = force_lds_bytes(a.max_c, true, B);
  hipLaunchKernelGGL((k_detect<ENS, B>), g, b, 0, s, a);
  switch (kind) {
    case FK_PP_PW_FORCES:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 0, false, true, false, ENS, true, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 0, false, false, false, ENS, true, B>), g, b, lds, s, a);
      break;
    case FK_STEP_FUSED:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 2, false, true, false, ENS, true, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 2, false, false, false, ENS, true, B>), g, b, lds, s, a);
      break;
    case FK_STEP_LAST:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 1, false, true, false, ENS, true, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 1, false, false, false, ENS, true, B>), g, b, lds, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The split force pass without sampling: k_defer, then k_force_list.
template <bool REF, bool ENS, int B>
hipError_t launch_deferred(ForceKind kind, const StepArgs& a, const dim3& g, hipStream_t s) {
  const dim3 gl = g, b(B);
  switch (kind) {
    case FK_PP_PW_FORCES:
      hipLaunchKernelGGL((k_defer<REF, 0, ENS, B>), g, b, 0, s, a);
      hipLaunchKernelGGL((k_force_list<REF, 0, ENS, B>), gl, b, 0, s, a);
      break;
    case FK_STEP_FUSED:
      hipLaunchKernelGGL((k_defer<REF, 2, ENS, B>), g, b, 0, s, a);
      hipLaunchKernelGGL((k_force_list<REF, 2, ENS, B>), gl, b, 0, s, a);
      break;
    case FK_STEP_LAST:
      hipLaunchKernelGGL((k_defer<REF, 1, ENS, B>), g, b, 0, s, a);
      hipLaunchKernelGGL((k_force_list<REF, 1, ENS, B>), gl, b, 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// B: the workgroup size and tile of the launch (kBlock for bonded systems).
template <bool REF, bool BOND, bool ENS, int B>
hipError_t launch_force_mode(ForceKind kind, const StepArgs& a, bool sample, int grid_blocks,
                             hipStream_t s) {
  // systems without bonds run the wave pass, which needs its LDS columns
  const size_t lds = force_lds_bytes(a.max_c, !BOND, B);
  const int nb = (grid_blocks > 0 && grid_blocks < a.ntiles) ? grid_blocks : a.ntiles;
  const dim3 g(nb), b(B);
  if constexpr (!BOND) {
    if (a.split && (kind == FK_PP_PW_FORCES || kind == FK_STEP_FUSED || kind == FK_STEP_LAST)) {
      if (sample) return launch_split<REF, ENS, B>(kind, a, sample, g, s);
      return launch_deferred<REF, ENS, B>(kind, a, g, s);
    }
  }
  switch (kind) {
    case FK_PP_PW_FORCES:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 0, false, true, BOND, ENS, false, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 0, false, false, BOND, ENS, false, B>), g, b, lds, s, a);
      break;
    case FK_PP_ACCUM:
      hipLaunchKernelGGL((k_force<REF, true, false, 0, true, false, false, ENS, false, B>), g, b, lds, s, a);
      break;
    case FK_PW_ACCUM:
      hipLaunchKernelGGL((k_force<REF, false, true, 0, true, false, false, ENS, false, B>), g, b, lds, s, a);
      break;
    case FK_BOND_ACCUM:
      if constexpr (BOND && !ENS)
        hipLaunchKernelGGL((k_force<REF, false, false, 0, true, false, true, false, false, B>), g, b, lds, s, a);
      else
        return hipErrorInvalidValue;
      break;
    case FK_STEP_FUSED:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 2, false, true, BOND, ENS, false, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 2, false, false, BOND, ENS, false, B>), g, b, lds, s, a);
      break;
    case FK_STEP_LAST:
      if (sample) hipLaunchKernelGGL((k_force<REF, true, true, 1, false, true, BOND, ENS, false, B>), g, b, lds, s, a);
      else hipLaunchKernelGGL((k_force<REF, true, true, 1, false, false, BOND, ENS, false, B>), g, b, lds, s, a);
      break;
  }
  return hipGetLastError();
}

} // namespace
} // namespace demhip
