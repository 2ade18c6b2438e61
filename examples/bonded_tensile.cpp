// bonded_tensile.cpp -- the reference's two-particle bonded tensile driver
// (examples/tst_01_BDEM_two_particles_tensile_3D.cpp) through the drop-in
// C++ API on the GPU.
//
//   bonded_tensile <staged|solver> <final_time> <output_dir>
//
// staged: the reference's loop as written (:295-314): integrator stages,
//         resetForcesAndTorques, the external force and the boundary
//         condition applied through host slices, computeBondedForce.
// solver: the same physics through SolverBondedDEM, with the external
//         force and the boundary condition set once (setExternalForce,
//         setGroups) and every step run by the fused kernel.
// Prints one JSON line with the final state of both particles.
#include <array>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include <CabanaDEM.hpp>

#define MAX_BONDS 27

using exec_space = CabanaDEM::DefaultExecutionSpace;
using memory_space = typename exec_space::memory_space;
using particles_type = CabanaBondedDEM::ParticlesBondedDEM<memory_space, 3, MAX_BONDS>;

// examples/tst_01_...cpp:18-66: force of particle 0 set to zero.
template <class ParticleType>
void applyBoundaryConditions(ParticleType& particles) {
  auto force = particles.sliceForce();
  force(0, 0) = 0.;
  force(0, 1) = 0.;
  force(0, 2) = 0.;
}

// examples/tst_01_...cpp:69-117: F_i += ext_frc_i.
template <class ParticleType>
void applyTensileForce(ParticleType& particles, const std::vector<double>& ext_frc) {
  auto force = particles.sliceForce();
  for (std::size_t i = 0; i < particles.size(); ++i)
    for (int d = 0; d < 3; ++d) force(i, d) += ext_frc[3 * i + d];
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <staged|solver> <final_time> <output_dir>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  const double final_time = std::atof(argv[2]);
  const std::string out_dir = argv[3];
  try {
    // examples/tst_01_...cpp:128-232
    const double rho_p_inp = 2000.;
    const double radius_p_inp = 0.5;
    const double E_p_inp = 1e4;
    const double G_p_inp = 1e3;
    const double nu_p_inp = 0.25;
    auto particles = std::make_shared<particles_type>(exec_space(), 2);
    particles->setOutputFolder(out_dir);
    particles->update_mesh_limits({-4. * radius_p_inp, -4. * radius_p_inp, -4. * radius_p_inp},
                                  {4. * radius_p_inp, 4. * radius_p_inp, 4. * radius_p_inp});
    auto x = particles->slicePosition();
    auto u = particles->sliceVelocity();
    auto au = particles->sliceAcceleration();
    auto force = particles->sliceForce();
    auto torque = particles->sliceTorque();
    auto omega = particles->sliceOmega();
    auto m = particles->sliceMass();
    auto rho = particles->sliceDensity();
    auto rad = particles->sliceRadius();
    auto E = particles->sliceYoungsMod();
    auto nu = particles->slicePoissonsRatio();
    auto G = particles->sliceShearMod();
    auto I = particles->sliceMomentOfInertia();
    auto bond_idx = particles->sliceBondIdx();
    auto bond_init_len = particles->sliceBondInitLen();
    auto total_no_bonds = particles->sliceTotalNoBonds();
    auto init = [&](const int i) {
      double m_p_i = 4. / 3. * M_PI * radius_p_inp * radius_p_inp * radius_p_inp * rho_p_inp;
      double I_p_i = 2. / 5. * m_p_i * radius_p_inp * radius_p_inp;
      x(i, 0) = -radius_p_inp + i * (2. * radius_p_inp);
      x(i, 1) = 0.;
      x(i, 2) = 0.;
      for (int j = 0; j < 3; ++j) {
        u(i, j) = 0.;
        au(i, j) = 0.;
        force(i, j) = 0.;
        torque(i, j) = 0.;
        omega(i, j) = 0.;
      }
      m(i) = m_p_i;
      rho(i) = rho_p_inp;
      rad(i) = radius_p_inp;
      E(i) = E_p_inp;
      nu(i) = nu_p_inp;
      G(i) = G_p_inp;
      I(i) = I_p_i;
      for (int j = 0; j < MAX_BONDS; ++j) {
        bond_idx(i, j) = -1;
        bond_init_len(i, j) = 0.;
      }
      total_no_bonds(i) = 0;
    };
    particles->updateParticles(exec_space{}, init);
    particles->setupBonds<exec_space>();

    const double dt = 1e-4;
    const int num_steps = final_time / dt;
    const int output_frequency = 1000;
    std::vector<double> ext_frc(6, 0.);
    ext_frc[0] = -1e3;  // :288-289
    ext_frc[3] = 1e3;

    if (mode == "staged") {
      CabanaBondedDEM::BondedDEMIntegrator<exec_space> integrator(dt);
      for (int step = 0; step <= num_steps; step++) {
        integrator.stage1(*particles);
        integrator.stage2(*particles);
        CabanaBondedDEM::resetForcesAndTorques<exec_space>(*particles);
        applyTensileForce(*particles, ext_frc);
        CabanaBondedDEM::computeBondedForce<exec_space>(*particles, dt);
        applyBoundaryConditions(*particles);
        integrator.stage3(*particles);
        if (step % output_frequency == 0) particles->output(step / output_frequency, step * dt);
      }
    } else if (mode == "solver") {
      particles->setExternalForce(ext_frc);
      particles->setGroups({1, 0}, {DEM_GROUP_FIXED_FORCE});
      CabanaBondedDEM::SolverBondedDEM<memory_space, particles_type> solver(particles, dt, num_steps,
                                                                          output_frequency);
      solver.verbose = false;
      solver.run();
    } else {
      std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
      return 2;
    }

    auto xs = particles->slicePosition();
    auto us = particles->sliceVelocity();
    auto ws = particles->sliceOmega();
    auto nb = particles->sliceTotalNoBonds();
    std::printf("{\"mode\": \"%s\", \"num_steps\": %d, \"particles\": [", mode.c_str(), num_steps);
    for (int i = 0; i < 2; ++i) {
      std::printf("%s{\"x\": [%.17g, %.17g, %.17g], \"u\": [%.17g, %.17g, %.17g], "
                  "\"omega\": [%.17g, %.17g, %.17g], \"bonds\": %d}",
                  i ? ", " : "", xs(i, 0), xs(i, 1), xs(i, 2), us(i, 0), us(i, 1), us(i, 2),
                  ws(i, 0), ws(i, 1), ws(i, 2), nb(i));
    }
    std::printf("]}\n");
  } catch (const CabanaDEM::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
