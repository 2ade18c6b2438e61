// CabanaDEM_Force.hpp -- the contact force model with the reference's API
// (src/CabanaDEM_Force.hpp:13-711): Force<ExecSpace>(cor_pp, cor_pw,
// friction_pp, friction_pw) with its four stage methods, and the free
// functions computeForceParticleParticle[InfiniteWall].  Every method runs
// the matching HIP kernel through the C ABI on the particles' device copy.
#pragma once

#include "CabanaDEM_Core.hpp"
#include "CabanaDEM_Neighbor.hpp"

namespace CabanaDEM {

template <class ExecutionSpace>
class Force {
  using exec_space = ExecutionSpace;

 public:
  double cor_pp, cor_pw, friction_pp, friction_pw;
  // Extensions (not in the reference): body force per unit mass and the
  // arithmetic mode (DEM_MODE_REFERENCE reproduces the reference exactly).
  double gravity[3] = {0., 0., 0.};
  int physics_mode = DEM_MODE_REFERENCE;

  Force(double cor_pp, double cor_pw, double friction_pp, double friction_pw)
      : cor_pp(cor_pp), cor_pw(cor_pw), friction_pp(friction_pp), friction_pw(friction_pw) {}
  ~Force() {}

  // Fills the force-model part of a context configuration.
  void configure(dem_config& c) const {
    c.cor_pp = cor_pp;
    c.cor_pw = cor_pw;
    c.friction_pp = friction_pp;
    c.friction_pw = friction_pw;
    for (int d = 0; d < 3; ++d) c.gravity[d] = gravity[d];
    c.physics_mode = physics_mode;
  }

  // src/CabanaDEM_Force.hpp:39-47
  template <class ParticleType>
  void makeForceTorqueZeroOnParticle(ParticleType& particles) {
    auto& dev = prepare(particles, -1.);
    check(dem_zero_forces(dev.ctx), dev.ctx);
    dev.after_compute();
  }

  // src/CabanaDEM_Force.hpp:49-341
  template <class ParticleType, class NeighListType, class ParallelType>
  void computeForceFullParticleParticle(ParticleType& particles, const NeighListType& neigh_list,
                                        ParallelType&, double dt) {
    auto& dev = prepare(particles, dt);
    neigh_list.refresh();
    check(dem_force_pp(dev.ctx), dev.ctx);
    dev.after_compute();
  }

  // src/CabanaDEM_Force.hpp:343-445
  template <class ParticleType, class NeighListType>
  void update_tangential_contacts(ParticleType& particles, const NeighListType&) {
    auto& dev = prepare(particles, -1.);
    check(dem_update_tangential(dev.ctx), dev.ctx);
    dev.after_compute();
  }

  // src/CabanaDEM_Force.hpp:447-678
  template <class ParticleType, class InfiniteWallType>
  void computeForceFullParticleInfiniteWall(ParticleType& particles, InfiniteWallType& wall, double dt) {
    auto& dev = prepare(particles, dt);
    if ((int)wall.size() > ParticleType::max_walls)
      throw Error(DEM_ERR_ARG, "more walls than the particles' MaxNoWalls");
    dev.bind_walls(&wall, wall.version(), wall.view());
    check(dem_force_pw(dev.ctx), dev.ctx);
    dev.after_compute();
  }

 private:
  template <class ParticleType>
  detail::DeviceBinding& prepare(ParticleType& particles, double dt) {
    auto& dev = particles.device();
    dem_config want = dev.cfg;
    configure(want);
    if (dt > 0.) want.dt = dt;
    dev.ensure(want);
    return dev;
  }
};

// src/CabanaDEM_Force.hpp:686-696
template <class ForceType, class ParticleType, class NeighListType, class ParallelType>
void computeForceParticleParticle(ForceType& force, ParticleType& particles,
                                  const NeighListType& neigh_list, const ParallelType& neigh_op_tag,
                                  double dt) {
  force.makeForceTorqueZeroOnParticle(particles);
  force.update_tangential_contacts(particles, neigh_list);
  ParallelType tag = neigh_op_tag;
  force.computeForceFullParticleParticle(particles, neigh_list, tag, dt);
}

// src/CabanaDEM_Force.hpp:698-710
template <class ForceType, class ParticleType, class InfiniteWallType, class NeighListType,
          class ParallelType>
void computeForceParticleParticleInfiniteWall(ForceType& force, ParticleType& particles,
                                              InfiniteWallType& wall, const NeighListType& neigh_list,
                                              const ParallelType& neigh_op_tag, double dt) {
  force.makeForceTorqueZeroOnParticle(particles);
  force.update_tangential_contacts(particles, neigh_list);
  ParallelType tag = neigh_op_tag;
  force.computeForceFullParticleParticle(particles, neigh_list, tag, dt);
  force.computeForceFullParticleInfiniteWall(particles, wall, dt);
}

} // namespace CabanaDEM
