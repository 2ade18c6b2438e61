// dem_gen.cpp -- synthetic jittered-lattice packings (host only).
//
// Counter-based stream: draw k of particle i under seed s is
//   splitmix64(splitmix64(s) ^ (16 i + k)) >> 11, scaled to [0, 1).
// Normal draws use Box-Muller on two uniforms.  The same definition is
// restated in numpy by the tests, so the workloads are reproducible without
// calling this library.
#include <cmath>
#include <cstdint>

#include "../../include/cabanadem.h"

namespace {

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

inline double uniform01(uint64_t skey, uint64_t i, uint64_t k) {
  return (double)(splitmix64(skey ^ (16ull * i + k)) >> 11) * (1.0 / 9007199254740992.0);
}

inline double normal01(uint64_t skey, uint64_t i, uint64_t k) {
  const double u1 = 1.0 - uniform01(skey, i, k); // (0, 1]
  const double u2 = uniform01(skey, i, k + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

} // namespace

extern "C" int dem_generate_lattice(int nx, int ny, int nz, double spacing, const double lo[3],
                                    double jitter, double rmin, double rmax, double vsigma,
                                    double rho, double E, double nu, uint64_t seed, double* x,
                                    double* u, double* rad, double* m, double* I, double* Ey,
                                    double* nuv, double* G) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !lo) return DEM_ERR_ARG;
  const uint64_t skey = splitmix64(seed);
  const double Gval = E / (2. * (1. + nu));
  const long n = (long)nx * ny * nz;
  for (long i = 0; i < n; ++i) {
    const long ix = i % nx, iy = (i / nx) % ny, iz = i / ((long)nx * ny);
    const double base[3] = {lo[0] + spacing * (double)ix, lo[1] + spacing * (double)iy,
                            lo[2] + spacing * (double)iz};
    for (int d = 0; d < 3; ++d) {
      const double r = uniform01(skey, (uint64_t)i, (uint64_t)d);
      if (x) x[3 * i + d] = base[d] + jitter * (2.0 * r - 1.0);
    }
    const double rr = rmin + (rmax - rmin) * uniform01(skey, (uint64_t)i, 3);
    if (rad) rad[i] = rr;
    // examples/01_elastic_normal_impact_of_two_identical_particles.cpp:106-107
    const double mm = 4. / 3. * M_PI * rr * rr * rr * rho;
    if (m) m[i] = mm;
    if (I) I[i] = 2. / 5. * mm * rr * rr;
    if (u)
      for (int d = 0; d < 3; ++d) u[3 * i + d] = vsigma * normal01(skey, (uint64_t)i, 4 + 2 * d);
    if (Ey) Ey[i] = E;
    if (nuv) nuv[i] = nu;
    if (G) G[i] = Gval;
  }
  return DEM_OK;
}
