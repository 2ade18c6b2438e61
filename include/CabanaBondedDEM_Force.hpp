// CabanaBondedDEM_Force.hpp -- bonded force stages with the reference's API
// (src/CabanaBondedDEM_Force.hpp:13-350): resetForcesAndTorques and
// computeBondedForce, each one HIP kernel through the C ABI on the
// particles' device copy.
#pragma once

#include "CabanaBondedDEM_Particles.hpp"

namespace CabanaBondedDEM {

// src/CabanaBondedDEM_Force.hpp:14-50: F = 0 and T = 0 (plus the external
// force set with setExternalForce, the driver's applyTensileForce).
template <class ExecutionSpace, class ParticleType>
void resetForcesAndTorques(ParticleType& particles) {
  auto& dev = particles.prepare(-1.);
  check(dem_zero_forces(dev.ctx), dev.ctx);
  dev.after_compute();
}

// src/CabanaBondedDEM_Force.hpp:52-348: beam bond force, torsion and
// bending moments with non-viscous damping, accumulated into F and T; the
// bond histories are updated.  With unbonded_contacts set, contacts between
// unbonded pairs are not part of this call (dem_step includes them).
template <class ExecutionSpace, class ParticleType>
void computeBondedForce(ParticleType& particles, double dt) {
  auto& dev = particles.prepare(dt);
  check(dem_bond_forces(dev.ctx), dev.ctx);
  dev.after_compute();
}

} // namespace CabanaBondedDEM
