// CabanaDEM_Integrator.hpp -- velocity-Verlet stages with the reference's
// API (src/CabanaDEM_Integrator.hpp:15-106): stage1 and stage3 are the half
// kick u += (dt/2 F) / m, omega += (dt/2 T) / I; stage2 the drift x += dt u.
#pragma once

#include "CabanaDEM_Core.hpp"

namespace CabanaDEM {

template <class ExecutionSpace>
class Integrator {
  using exec_space = ExecutionSpace;
  double _dt, _half_dt;

 public:
  Integrator(double dt) : _dt(dt) { _half_dt = 0.5 * dt; }
  ~Integrator() {}

  double dt() const { return _dt; }

  template <class ParticlesType>
  void stage1(ParticlesType& p) { run(p, 1); }
  template <class ParticlesType>
  void stage2(ParticlesType& p) { run(p, 2); }
  template <class ParticlesType>
  void stage3(ParticlesType& p) { run(p, 3); }

 private:
  template <class ParticlesType>
  void run(ParticlesType& p, int which) {
    auto& dev = p.device();
    dem_config want = dev.cfg;
    want.dt = _dt;
    dev.ensure(want);
    int rc = which == 1 ? dem_stage1(dev.ctx) : which == 2 ? dem_stage2(dev.ctx) : dem_stage3(dev.ctx);
    check(rc, dev.ctx);
    dev.after_compute();
  }
};

} // namespace CabanaDEM
