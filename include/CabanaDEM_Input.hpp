// CabanaDEM_Input.hpp -- JSON inputs with the reference's API
// (src/CabanaDEM_Input.hpp:24-82): values are read as inputs["key"], which
// is the "value" field of {"key": {"value": v, "unit": ""}}, and the
// constructor adds num_steps = int(final_time / timestep) (truncating).
#pragma once

#include <string>
#include <vector>

#include "cabanadem_json.hpp"

namespace CabanaDEM {

// src/CabanaDEM_Input.hpp:24-38: split on a delimiter, keeping empty pieces.
inline std::vector<std::string> splitString(const std::string& str, char delimiter) {
  std::vector<std::string> out;
  std::size_t start = 0;
  for (std::size_t pos = str.find(delimiter); pos != std::string::npos;
       pos = str.find(delimiter, start)) {
    out.push_back(str.substr(start, pos - start));
    start = pos + 1;
  }
  out.push_back(str.substr(start));
  return out;
}

class Inputs {
 public:
  Inputs(const std::string filename) {
    inputs = cabanadem_json::parse_file(filename);
    const double tf = inputs["final_time"]["value"];
    const double dt = inputs["timestep"]["value"];
    const int num_steps = (int)(tf / dt);
    inputs["num_steps"]["value"] = num_steps;
  }
  ~Inputs() {}

  // A missing key reads as null and throws when converted.
  cabanadem_json::Value& operator[](const std::string& label) { return inputs[label]["value"]; }

  // The reference looks for "units"; the shipped input files use "unit".
  std::string units(const std::string& label) {
    if (inputs[label].contains("units")) return inputs[label]["units"];
    return "";
  }

  bool contains(const std::string& label) const { return inputs.contains(label); }

 protected:
  cabanadem_json::Value inputs;
};

} // namespace CabanaDEM
