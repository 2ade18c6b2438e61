// CPU check of the drop-in Inputs / splitString (src/CabanaDEM_Input.hpp:24-82).
#include <cmath>
#include <cstdio>
#include <string>

#include <CabanaDEM_Input.hpp>

static int fails = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++fails; } \
  } while (0)

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden/inputs";
  CabanaDEM::Inputs in(dir + "/01_elastic_normal_impact_of_two_identical_particles.json");
  int num_steps = in["num_steps"];
  EXPECT(num_steps == 800);  // int(80e-6 / 1e-7), truncating
  double rho = in["particle_density"];
  EXPECT(rho == 2800.0);
  int of = in["output_frequency"];
  EXPECT(of == 30);
  EXPECT(in.contains("friction_pw"));
  EXPECT(!in.contains("gravity"));
  EXPECT(in.units("particle_density") == "");  // the files use "unit", the reader "units"
  bool threw = false;
  try {
    double missing = in["no_such_key"];
    (void)missing;
  } catch (const cabanadem_json::Error&) {
    threw = true;
  }
  EXPECT(threw);
  CabanaDEM::Inputs in4(dir + "/04_oblique_particle_wall_different_angles.json");
  int n4 = in4["num_steps"];
  EXPECT(n4 == 600);
  auto parts = CabanaDEM::splitString("--angle=30.", '=');
  EXPECT(parts.size() == 2 && parts[0] == "--angle" && parts[1] == "30.");
  auto p3 = CabanaDEM::splitString("a==b", '=');
  EXPECT(p3.size() == 3 && p3[1].empty());
  auto v = cabanadem_json::parse("{\"g\": {\"value\": [0, 0, -9.81]}, \"s\": \"x\\\"y\", \"t\": true}");
  EXPECT(v["g"]["value"].at(2).as_double() == -9.81);
  EXPECT(v["s"].as_string() == "x\"y");
  EXPECT(v["t"].as_bool());
  std::printf(fails ? "FAILED\n" : "OK\n");
  return fails ? 1 : 0;
}
