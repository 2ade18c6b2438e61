// dem_validation.cpp -- the reference's validation set-ups run through the
// drop-in C++ API on the GPU.
//
//   dem_validation <case> <input.json> <output_dir> [--angle=A | --cor-pw=E]
//
// case: head_on      two identical spheres head on        (examples/01_...cpp)
//       sphere_wall  sphere dropped normally on a wall     (examples/02_...cpp)
//       cor          sphere_wall with cor_pw from --cor-pw (examples/03_...cpp)
//       oblique      oblique impact at --angle degrees     (examples/04_...cpp)
//       one_hits_two one sphere hitting two                (examples/tst_03_...cpp)
// Each case sets up the particles, the wall, the force model and the solver
// exactly as the reference driver of that name does and calls run(); the
// solver writes particles_<k>.csv every output_frequency steps, and the
// program prints one JSON line with the final state of every particle.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include <CabanaDEM.hpp>

using exec_space = CabanaDEM::DefaultExecutionSpace;
using memory_space = typename exec_space::memory_space;
using particles_type = CabanaDEM::Particles<memory_space, 3, 6, 1>;
using wall_type = CabanaDEM::Wall<memory_space, 3>;

struct Material {
  double rho, r, E, G, nu;
};

static Material particle_material(CabanaDEM::Inputs& in) {
  return {in["particle_density"], in["particle_radius"], in["particle_youngs_modulus"],
          in["particle_shear_modulus"], in["particle_poissons_ratio"]};
}

// The drivers' per-particle initialisation: m = 4/3 pi r^3 rho,
// I = 2/5 m r^2, material constants, empty tracking rows.
template <class P>
static void set_sphere(P& p, int pid, const Material& mat) {
  auto m = p.sliceMass();
  auto I = p.sliceMomentOfInertia();
  auto rho = p.sliceDensity();
  auto rad = p.sliceRadius();
  auto E = p.sliceYoungsMod();
  auto G = p.sliceShearMod();
  auto nuv = p.slicePoissonsRatio();
  auto idx = p.sliceTangentialDispIdx();
  auto cnt = p.sliceTotalNoTangentialContacts();
  const double mass = 4. / 3. * M_PI * mat.r * mat.r * mat.r * mat.rho;
  m(pid) = mass;
  I(pid) = 2. / 5. * mass * mat.r * mat.r;
  rho(pid) = mat.rho;
  rad(pid) = mat.r;
  E(pid) = mat.E;
  G(pid) = mat.G;
  nuv(pid) = mat.nu;
  for (int k = 0; k < P::max_contacts; ++k) idx(pid, k) = -1;
  cnt(pid) = 0;
}

static std::shared_ptr<wall_type> one_wall(CabanaDEM::Inputs& in, const std::string& out,
                                           double wy, double ny) {
  auto wall = std::make_shared<wall_type>(exec_space(), 1, out);
  auto x = wall->slicePosition();
  auto u = wall->sliceVelocity();
  auto n = wall->sliceNormal();
  auto E = wall->sliceYoungsMod();
  auto nu = wall->slicePoissonsRatio();
  auto G = wall->sliceShearMod();
  const double Ew = in["wall_youngs_modulus"], Gw = in["wall_shear_modulus"], nuw = in["wall_poissons_ratio"];
  wall->updateParticles(exec_space{}, [&](const int w) {
    x(w, 0) = 0.; x(w, 1) = wy; x(w, 2) = 0.;
    for (int d = 0; d < 3; ++d) u(w, d) = 0.;
    n(w, 0) = 0.; n(w, 1) = ny; n(w, 2) = 0.;
    E(w) = Ew; G(w) = Gw; nu(w) = nuw;
  });
  return wall;
}

static void print_state(const std::string& name, particles_type& p) {
  auto x = p.slicePosition();
  auto u = p.sliceVelocity();
  auto w = p.sliceOmega();
  auto f = p.sliceForce();
  auto c = p.sliceTotalNoTangentialContacts();
  std::printf("{\"case\": \"%s\", \"particles\": [", name.c_str());
  for (std::size_t i = 0; i < p.size(); ++i) {
    std::printf("%s{\"x\": [%.17g, %.17g, %.17g], \"u\": [%.17g, %.17g, %.17g], "
                "\"omega\": [%.17g, %.17g, %.17g], \"force\": [%.17g, %.17g, %.17g], \"contacts\": %d}",
                i ? ", " : "", x(i, 0), x(i, 1), x(i, 2), u(i, 0), u(i, 1), u(i, 2), w(i, 0), w(i, 1),
                w(i, 2), f(i, 0), f(i, 1), f(i, 2), c(i));
  }
  std::printf("]}\n");
}

int main(int argc, char* argv[]) {
  if (argc < 4) {
    std::cerr << "usage: dem_validation <head_on|sphere_wall|cor|oblique|one_hits_two> "
                 "<input.json> <output_dir> [--angle=A|--cor-pw=E]\n";
    return 2;
  }
  const std::string which = argv[1], out = argv[3];
  CabanaDEM::Inputs inputs(argv[2]);
  const Material mat = particle_material(inputs);
  const double v = inputs["velocity_p"];
  double cor_pp = inputs["coefficient_of_restitution_pp"];
  double cor_pw = inputs.contains("coefficient_of_restitution_pw")
                      ? (double)inputs["coefficient_of_restitution_pw"] : 1.0;
  double arg = 0.;
  if (argc > 4) arg = std::atof(CabanaDEM::splitString(argv[4], '=')[1].c_str());
  if (which == "cor") cor_pw = arg;
  auto force = std::make_shared<CabanaDEM::Force<exec_space>>(
      cor_pp, cor_pw, (double)inputs["friction_pp"], (double)inputs["friction_pw"]);

  const double r = mat.r;
  int np = (which == "head_on") ? 2 : (which == "one_hits_two") ? 3 : 1;
  auto particles = std::make_shared<particles_type>(exec_space(), np, out);
  auto x = particles->slicePosition();
  auto u = particles->sliceVelocity();
  std::shared_ptr<wall_type> wall;
  if (which == "head_on" || which == "one_hits_two") {
    // examples/01_...cpp:109-115 and examples/tst_03_...cpp:109-129
    particles->updateParticles(exec_space{}, [&](const int pid) {
      set_sphere(*particles, pid, mat);
      x(pid, 0) = -r + pid * (2. * r) + pid * r / 100.;
      x(pid, 1) = 0.;
      x(pid, 2) = 0.;
      u(pid, 0) = pid == 0 ? v : -v;
      u(pid, 1) = 0.;
      u(pid, 2) = 0.;
      if (which == "one_hits_two") {
        if (pid > 0) {
          x(pid, 0) = -r + 1.5 * r + r / 100.;
          x(pid, 1) = pid == 1 ? r + r / 2. : -r - r / 2.;
          u(pid, 0) = 0.;
        }
      }
    });
    wall = one_wall(inputs, out, -3. * r, which == "head_on" ? 1. : 0.);
    for (int d = 0; d < 3; ++d) {
      particles->mesh_lo[d] = -4. * 0.01;
      particles->mesh_hi[d] = 4. * 0.01;
    }
  } else {
    // examples/02_...cpp:108-114, 03_...cpp:110-116, 04_...cpp:117-136
    double ux = 0., uy = -v;
    if (which == "oblique") {
      const double ang = arg / 180. * M_PI;
      ux = v * std::sin(ang);
      uy = -v * std::cos(ang);
    }
    particles->updateParticles(exec_space{}, [&](const int pid) {
      set_sphere(*particles, pid, mat);
      x(pid, 0) = 0.;
      x(pid, 1) = r + r / 1000000.;
      x(pid, 2) = 0.;
      u(pid, 0) = ux;
      u(pid, 1) = uy;
      u(pid, 2) = 0.;
    });
    wall = one_wall(inputs, out, 0., 1.);
    const double span = which == "sphere_wall" ? 40. : 10.;
    particles->mesh_lo = {-span * r, -span * r, -4. * r};
    particles->mesh_hi = {span * r, span * r, 4. * r};
  }

  auto solver = CabanaDEM::createSolverDEM<memory_space>(inputs, particles, wall, force, 3. * r);
  solver->run();
  print_state(which, *particles);
  return 0;
}
