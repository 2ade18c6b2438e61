// CabanaDEM_Solver.hpp -- the DEM solver with the reference's API
// (src/CabanaDEM_Solver.hpp:10-138): SolverDEM(inputs, particles, wall,
// force, delta), run() and output(step), num_steps / output_frequency / dt,
// and the createSolverDEM factory.
//
// run() takes the reference's num_steps + 1 steps (its loop runs
// step = 0 .. num_steps, :79) in the reference's stage order (stage1,
// stage2, neighbour build, zero / update tangential / pp / pw forces,
// stage3, output).  On the GPU consecutive steps are fused into one kernel
// each, and the list is rebuilt every `rebuild_every` steps (1, as the
// reference, unless the input file says otherwise).
//
// Optional input keys (extensions, same {"value": ...} form):
//   gravity        [gx, gy, gz]          body force per unit mass (default 0)
//   rebuild_every  K                      neighbour rebuild cadence (default 1)
//   neighbour_skin s                      flag motion beyond s/2 between rebuilds
//   physics_mode   "reference"|"corrected" (default "reference")
#pragma once

#include <iostream>
#include <memory>
#include <string>

#include "CabanaDEM_Force.hpp"
#include "CabanaDEM_Input.hpp"
#include "CabanaDEM_Integrator.hpp"
#include "CabanaDEM_Neighbor.hpp"
#include "CabanaDEM_Particles.hpp"

namespace CabanaDEM {

template <class MemorySpace, class InputType, class ParticleType, class WallType, class ForceModel>
class SolverDEM {
 public:
  using memory_space = MemorySpace;
  using exec_space = typename memory_space::execution_space;
  using particle_type = ParticleType;
  using wall_type = WallType;
  using integrator_type = Integrator<exec_space>;
  using force_type = ForceModel;
  using neighbor_type = VerletList<memory_space, FullNeighborTag, VerletLayout2D, TeamOpTag>;
  using neigh_iter_tag = SerialOpTag;
  using input_type = InputType;

  SolverDEM(input_type _inputs, std::shared_ptr<particle_type> _particles,
            std::shared_ptr<wall_type> _wall, std::shared_ptr<force_type> _force, double _delta)
      : inputs(_inputs), particles(_particles), wall(_wall), force(_force), delta(_delta) {
    num_steps = inputs["num_steps"];
    output_frequency = inputs["output_frequency"];
    dt = inputs["timestep"];
    integrator = std::make_shared<integrator_type>(dt);
    if (inputs.contains("gravity")) {
      auto& g = inputs["gravity"];
      for (int d = 0; d < 3; ++d) force->gravity[d] = g.at(d).as_double();
    }
    if (inputs.contains("rebuild_every")) rebuild_every = inputs["rebuild_every"];
    if (inputs.contains("neighbour_skin")) skin = inputs["neighbour_skin"];
    if (inputs.contains("physics_mode"))
      force->physics_mode = inputs["physics_mode"].as_string() == "corrected" ? DEM_MODE_CORRECTED
                                                                              : DEM_MODE_REFERENCE;
    // the initial list (src/CabanaDEM_Solver.hpp:60-63), cell_ratio 1.0
    double mesh_min[3] = {particles->mesh_lo[0], particles->mesh_lo[1], particles->mesh_lo[2]};
    double mesh_max[3] = {particles->mesh_hi[0], particles->mesh_hi[1], particles->mesh_hi[2]};
    configure();
    neighbors = std::make_shared<neighbor_type>(particles->slicePosition(), 0, particles->size(),
                                                delta, 1.0, mesh_min, mesh_max);
  }

  // src/CabanaDEM_Solver.hpp:66-98
  void run() {
    configure();
    auto& dev = particles->device();
    dev.bind_walls(wall.get(), wall->version(), wall->view());
    for (int step = 0; step <= num_steps;) {
      // run up to (and including) the next output step in one fused call
      int next = step;
      if (output_frequency > 0) next = ((step + output_frequency - 1) / output_frequency) * output_frequency;
      if (next > num_steps) next = num_steps;
      const int chunk = next - step + 1;
      check(dem_step(dev.ctx, chunk), dev.ctx);
      dev.after_compute();
      step += chunk;
      output(step - 1);
    }
  }

  // src/CabanaDEM_Solver.hpp:100-108
  void output(const int step) {
    if (output_frequency > 0 && step % output_frequency == 0) {
      std::cout << "We are at " << step << " " << "/ " << num_steps;
      std::cout << std::endl;
      particles->output(step / output_frequency, step * dt);
    }
  }

  int num_steps;
  int output_frequency;
  double dt;
  int rebuild_every = 1;
  double skin = 0.;

 protected:
  void configure() {
    auto& dev = particles->device();
    dem_config want = dev.cfg;
    force->configure(want);
    want.dt = dt;
    want.cutoff = delta;
    want.rebuild_every = rebuild_every;
    want.skin = skin;
    for (int d = 0; d < 3; ++d) {
      want.lo[d] = particles->mesh_lo[d];
      want.hi[d] = particles->mesh_hi[d];
    }
    dev.ensure(want);
    if (neighbors) neighbors->refresh();
  }

  input_type inputs;
  std::shared_ptr<particle_type> particles;
  std::shared_ptr<wall_type> wall;
  std::shared_ptr<integrator_type> integrator;
  std::shared_ptr<force_type> force;
  std::shared_ptr<neighbor_type> neighbors;
  double delta;
};

// src/CabanaDEM_Solver.hpp:127-138
template <class MemorySpace, class InputsType, class ParticleType, class WallType, class ForceModel>
auto createSolverDEM(InputsType inputs, std::shared_ptr<ParticleType> particles,
                     std::shared_ptr<WallType> wall, std::shared_ptr<ForceModel> force, double delta) {
  return std::make_shared<SolverDEM<MemorySpace, InputsType, ParticleType, WallType, ForceModel>>(
      inputs, particles, wall, force, delta);
}

} // namespace CabanaDEM
