// Particle insertion with the solver running on the GPU: the container flow
// of examples/tst_01_particle_insertion.cpp:100-190 (fill, write, resize(20),
// refill the new entries through fresh slices, write, resize(15), write),
// with SolverDEM::run between the resizes so every segment steps on the
// device and the state (including the tangential histories) crosses each
// resize.  Spheres fall onto a floor under gravity, so pp and wall contacts
// are live when the container grows and shrinks.
//
// After the first segment the program writes the particles (output) and
// then edits a velocity through a slice: that edit must reach the device
// on the next run.
//
//   test_insertion_gpu <output_dir>
// Prints one JSON line per segment: {"segment": k, "n": n, "x": [...], "u":
// [...], "omega": [...], "force": [...]} in particle order.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>

#include <CabanaDEM.hpp>

using exec_space = CabanaDEM::DefaultExecutionSpace;
using memory_space = typename exec_space::memory_space;
using particles_type = CabanaDEM::Particles<memory_space, 3, 6, 1>;
using wall_type = CabanaDEM::Wall<memory_space, 3>;

constexpr double kR = 1e-3, kRho = 2600., kE = 1e7, kNu = 0.3;

// Sphere pid of the falling column: on one vertical line, with small
// sideways velocities so that the contacts have tangential components.  (A
// column set slightly off the line is an unstable stack: a one-ulp change
// of the start moves its end state by ~1e-5, which no parity bar survives.)
static void init_sphere(particles_type& p, int pid) {
  auto x = p.slicePosition();
  auto u = p.sliceVelocity();
  auto w = p.sliceOmega();
  auto m = p.sliceMass();
  auto I = p.sliceMomentOfInertia();
  auto rho = p.sliceDensity();
  auto rad = p.sliceRadius();
  auto E = p.sliceYoungsMod();
  auto G = p.sliceShearMod();
  auto nu = p.slicePoissonsRatio();
  auto idx = p.sliceTangentialDispIdx();
  auto cnt = p.sliceTotalNoTangentialContacts();
  const double mass = 4. / 3. * M_PI * kR * kR * kR * kRho;
  x(pid, 0) = 0.;
  x(pid, 1) = kR + 2.1 * kR * pid;
  x(pid, 2) = 0.;
  u(pid, 0) = 0.02 * ((pid % 3) - 1);
  u(pid, 1) = -0.5;
  u(pid, 2) = 0.01 * (pid % 2);
  for (int d = 0; d < 3; ++d) w(pid, d) = 0.;
  m(pid) = mass;
  I(pid) = 2. / 5. * mass * kR * kR;
  rho(pid) = kRho;
  rad(pid) = kR;
  E(pid) = kE;
  nu(pid) = kNu;
  G(pid) = kE / (2. * (1. + kNu));
  for (int k = 0; k < particles_type::max_contacts; ++k) idx(pid, k) = -1;
  cnt(pid) = 0;
}

static void print_segment(int seg, particles_type& p) {
  auto x = p.slicePosition();
  auto u = p.sliceVelocity();
  auto w = p.sliceOmega();
  auto f = p.sliceForce();
  std::printf("{\"segment\": %d, \"n\": %zu", seg, p.size());
  const char* names[4] = {"x", "u", "omega", "force"};
  CabanaDEM::Slice<double, 3> s[4] = {x, u, w, f};
  for (int k = 0; k < 4; ++k) {
    std::printf(", \"%s\": [", names[k]);
    for (std::size_t i = 0; i < p.size(); ++i)
      std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", s[k](i, 0), s[k](i, 1), s[k](i, 2));
    std::printf("]");
  }
  std::printf("}\n");
  std::fflush(stdout);
}

int main(int argc, char* argv[]) {
  if (argc < 2) return 2;
  const std::string out = argv[1];
  CabanaDEM::detail::make_output_dir(out);
  const std::string json = out + "/insertion.json";
  {
    std::ofstream os(json);
    os << "{\"final_time\": {\"value\": 2e-3}, \"timestep\": {\"value\": 1e-6},\n"
          " \"output_frequency\": {\"value\": 1000}, \"gravity\": {\"value\": [0.0, -9.81, 0.0]}}\n";
  }
  CabanaDEM::Inputs inputs(json);
  auto force = std::make_shared<CabanaDEM::Force<exec_space>>(0.8, 0.8, 0.4, 0.4);
  auto particles = std::make_shared<particles_type>(exec_space(), 10, out);
  particles->updateParticles(exec_space{}, [&](const int pid) { init_sphere(*particles, pid); });
  particles->mesh_lo = {-10. * kR, -2. * kR, -10. * kR};
  particles->mesh_hi = {10. * kR, 50. * kR, 10. * kR};
  auto wall = std::make_shared<wall_type>(exec_space(), 1, out);
  {
    auto x = wall->slicePosition();
    auto n = wall->sliceNormal();
    auto E = wall->sliceYoungsMod();
    auto nu = wall->slicePoissonsRatio();
    auto G = wall->sliceShearMod();
    for (int d = 0; d < 3; ++d) { x(0, d) = 0.; n(0, d) = d == 1 ? 1. : 0.; }
    E(0) = kE; nu(0) = kNu; G(0) = kE / (2. * (1. + kNu));
  }
  auto solver = CabanaDEM::createSolverDEM<memory_space>(inputs, particles, wall, force, 3. * kR);

  solver->run();
  print_segment(0, *particles);
  particles->output(100, 0.1);
  {
    auto u = particles->sliceVelocity(); // edited after the output: must reach the device
    u(0, 0) += 0.1;
  }

  particles->resize(20);
  for (int pid = 10; pid < 20; ++pid) init_sphere(*particles, pid);
  solver->run();
  print_segment(1, *particles);

  particles->resize(15);
  solver->run();
  print_segment(2, *particles);
  particles->output(300, 0.3);
  return 0;
}
