// CabanaBondedDEM_Solver.hpp -- SolverBondedDEM.  The reference's class is an
// empty stub whose commented-out body (src/CabanaBondedDEM_Solver.hpp:10-104)
// describes the loop stage1, stage2, computeBondedForce, stage3 with output
// every output_frequency steps, for num_steps + 1 steps.  This is that loop,
// with the bonded driver's reset, external force and boundary condition
// (examples/tst_01_BDEM_two_particles_tensile_3D.cpp:295-314), run by the
// fused step kernel: stage1, stage2, reset + external force, bond forces,
// [contacts between unbonded pairs], fixed-force groups, stage3.
#pragma once

#include <iostream>
#include <memory>

#include "CabanaBondedDEM_Force.hpp"
#include "CabanaBondedDEM_Integrator.hpp"
#include "CabanaBondedDEM_Particles.hpp"

namespace CabanaBondedDEM {

template <class MemorySpace, class ParticleType>
class SolverBondedDEM {
 public:
  using memory_space = MemorySpace;
  using exec_space = typename memory_space::execution_space;
  using particle_type = ParticleType;
  using integrator_type = BondedDEMIntegrator<exec_space>;

  int num_steps;
  int output_frequency;
  double dt;
  int physics_mode = DEM_MODE_REFERENCE;
  bool verbose = true;

  SolverBondedDEM(std::shared_ptr<particle_type> particles, double dt, int num_steps,
                  int output_frequency)
      : num_steps(num_steps), output_frequency(output_frequency), dt(dt), particles(particles) {
    integrator = std::make_shared<integrator_type>(dt);
  }

  // Steps 0..num_steps; the particles are written at every multiple of
  // output_frequency, after that step's stage3.
  void run() {
    auto& dev = particles->prepare(dt, physics_mode);
    int step = 0; // next step to take
    while (step <= num_steps) {
      // last step of this call: the next output step, or the end
      int last = num_steps;
      if (output_frequency > 0)
        last = std::min(num_steps, (step + output_frequency - 1) / output_frequency * output_frequency);
      check(dem_step(dev.ctx, last - step + 1), dev.ctx);
      dev.after_compute();
      step = last + 1;
      output(last);
    }
  }

  void output(const int step) {
    if (output_frequency > 0 && step % output_frequency == 0) {
      if (verbose) std::cout << "We are at " << step << " / " << num_steps << std::endl;
      particles->output(step / output_frequency, step * dt);
    }
  }

 protected:
  std::shared_ptr<particle_type> particles;
  std::shared_ptr<integrator_type> integrator;
};

} // namespace CabanaBondedDEM
