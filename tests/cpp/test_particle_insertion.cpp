// The container flow of the reference's particle-insertion test
// (examples/tst_01_particle_insertion.cpp:100-190): fill 10 particles through
// updateParticles, write, resize to 20 and refill through the slices, move
// them, write, shrink to 15, refill and write.  Host only (no solver).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include <CabanaDEM.hpp>

static int rows(const std::string& fn) {
  std::ifstream is(fn);
  std::string line;
  int n = 0;
  while (std::getline(is, line)) ++n;
  return n - 2; // time and column headers
}

int main(int argc, char* argv[]) {
  if (argc < 2) return 2;
  const std::string out = argv[1];
  using exec_space = CabanaDEM::DefaultExecutionSpace;
  using memory_space = typename exec_space::memory_space;
  const double r = 0.01, rho = 2000.;
  auto particles = std::make_shared<CabanaDEM::Particles<memory_space, 3, 6, 1>>(exec_space(), 10, out);
  auto x_p = particles->slicePosition();
  auto u_p = particles->sliceVelocity();
  auto m_p = particles->sliceMass();
  auto rad_p = particles->sliceRadius();
  auto init = [&](const int pid) {
    x_p(pid, 0) = 0.;
    x_p(pid, 1) = 3. * pid * r;
    x_p(pid, 2) = 0.;
    u_p(pid, 0) = 0.;
    u_p(pid, 1) = 1.;
    u_p(pid, 2) = 0.;
    m_p(pid) = rho * 4. / 3. * M_PI * r * r * r;
    rad_p(pid) = r;
  };
  particles->updateParticles(exec_space{}, init);
  particles->output(0, 0.);
  int fail = 0;
  auto check = [&](bool ok, const char* what) {
    if (!ok) {
      std::cout << "FAIL " << what << "\n";
      ++fail;
    }
  };
  check(rows(out + "/particles_0.csv") == 10, "10 rows");

  particles->resize(20);
  auto x20 = particles->slicePosition();
  auto u20 = particles->sliceVelocity();
  check(x20.size() == 20 && particles->size() == 20, "resized to 20");
  // the first 10 keep their values, the new ones start at zero
  check(x20(9, 1) == 27. * r && u20(9, 1) == 1. && x20(15, 1) == 0. && u20(15, 1) == 0., "grow keeps and zeroes");
  for (std::size_t i = 0; i < x20.size(); i++) {
    x20(i, 0) = 0.;
    x20(i, 1) = 3 * i * r + r / 1000000.;
    x20(i, 2) = 0.;
    u20(i, 0) = 0.;
    u20(i, 1) = 1.;
    u20(i, 2) = 0.;
  }
  particles->output(10, 0.1);
  for (std::size_t i = 0; i < x20.size(); i++)
    for (int d = 0; d < 3; ++d) x20(i, d) += 2. * u20(i, d);
  particles->output(20, 0.2);
  check(rows(out + "/particles_20.csv") == 20, "20 rows");
  {
    std::ifstream is(out + "/particles_20.csv");
    std::string line;
    std::getline(is, line);
    check(line.rfind("# Time=0.2", 0) == 0, "time header");
    std::getline(is, line);
    for (int k = 0; k < 20; ++k) std::getline(is, line);
    double px, py;
    check(std::sscanf(line.c_str(), "%lf,%lf", &px, &py) == 2 && px == 0. && py == 3 * 19 * r + r / 1000000. + 2.,
          "last row moved");
  }

  particles->resize(15);
  auto x15 = particles->slicePosition();
  check(x15.size() == 15 && particles->size() == 15, "shrunk to 15");
  check(x15(14, 1) == 3 * 14 * r + r / 1000000. + 2., "shrink keeps the first 15");
  for (std::size_t i = 0; i < x15.size(); i++) {
    x15(i, 0) = 0.;
    x15(i, 1) = 5 * i * r + r / 1000000.;
    x15(i, 2) = 0.;
  }
  particles->output(30, 0.3);
  check(rows(out + "/particles_30.csv") == 15, "15 rows");
  if (fail) return 1;
  std::cout << "OK\n";
  return 0;
}
