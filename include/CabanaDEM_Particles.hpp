// CabanaDEM_Particles.hpp -- particle container with the reference's API
// (src/CabanaDEM_Particles.hpp:22-378): the 21 per-particle fields, the slice
// getters and labels, updateParticles, resize and output.  Storage is a host
// mirror in the reference layout; the GPU copy lives behind the C ABI.
#pragma once

#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "CabanaDEM_Core.hpp"
#include "CabanaDEM_HDF5.hpp"

#ifndef DIM
#define DIM 3
#endif

namespace CabanaDEM {

template <class MemorySpace, int Dimension, int MaxCnts, int MaxNoWalls>
class Particles {
 public:
  using memory_space = MemorySpace;
  using execution_space = typename memory_space::execution_space;
  static constexpr int dim = Dimension;
  static constexpr int max_contacts = MaxCnts;
  static constexpr int max_walls = MaxNoWalls;
  static_assert(Dimension == 3, "three-dimensional particles only");
  static_assert(MaxCnts >= 1 && MaxCnts <= 32, "MaxCnts must be in [1, 32]");
  static_assert(MaxNoWalls >= 1 && MaxNoWalls <= DEM_MAX_WALLS, "MaxNoWalls must be in [1, 8]");

  std::array<double, DIM> mesh_lo;
  std::array<double, DIM> mesh_hi;

  // src/CabanaDEM_Particles.hpp:52-71: creates the output folder, sizes the
  // arrays, zeroes the mesh limits.
  template <class ExecSpace>
  Particles(const ExecSpace&, std::size_t no_of_particles, std::string output_folder_name)
      : output_folder_name_(std::move(output_folder_name)),
        mirror_(std::make_unique<detail::Mirror>()),
        dev_(std::make_unique<detail::DeviceBinding>()) {
    detail::make_output_dir(output_folder_name_);
    resize(no_of_particles);
    mesh_lo = {0.0, 0.0, 0.0};
    mesh_hi = {0.0, 0.0, 0.0};
    mirror_->pull = [this] { dev_->download(); };
    dev_->mirror = mirror_.get();
    mirror_->binding = dev_.get();
    dev_->fill_view = [this](dem_particles_view& v) { fill_view(v); };
    dev_->count = [this] { return (int)n_; };
    dev_->max_cnts = [] { return MaxCnts; };
    dev_->max_walls = [] { return MaxNoWalls; };
  }

  // Runs the user's functor for every particle on the host mirror
  // (src/CabanaDEM_Particles.hpp:90-97); the next device call uploads it.
  template <class ExecSpace, class FunctorType>
  void updateParticles(const ExecSpace, const FunctorType init_functor) {
    mirror_->before_host_access();
    for (std::size_t pid = 0; pid < n_; ++pid) init_functor((int)pid);
  }

  auto slicePosition() const { return vec(x_, "positions"); }
  auto sliceVelocity() const { return vec(u_, "velocities"); }
  auto sliceAcceleration() const { return vec(au_, "accelerations"); }
  auto sliceForce() const { return vec(force_, "forces"); }
  auto sliceTorque() const { return vec(torque_, "torques"); }
  auto sliceOmega() const { return vec(omega_, "omega"); }
  auto sliceMass() const { return sca(m_, "mass"); }
  auto sliceDensity() const { return sca(rho_, "density"); }
  auto sliceRadius() const { return sca(rad_, "radius"); }
  auto sliceYoungsMod() const { return sca(E_, "youngs_mod"); }
  auto slicePoissonsRatio() const { return sca(nu_, "poissons_ratio"); }
  auto sliceShearMod() const { return sca(G_, "shear_mod"); }
  auto sliceMomentOfInertia() const { return sca(I_, "moment_of_inertia"); }
  auto sliceTangentialDispSWx() const { return wall(sw_x_, "tangential_disp_sw_x"); }
  auto sliceTangentialDispSWy() const { return wall(sw_y_, "tangential_disp_sw_y"); }
  auto sliceTangentialDispSWz() const { return wall(sw_z_, "tangential_disp_sw_z"); }
  auto sliceTangentialDispSSx() const { return cnt(ss_x_, "tangential_disp_ss_x"); }
  auto sliceTangentialDispSSy() const { return cnt(ss_y_, "tangential_disp_ss_y"); }
  auto sliceTangentialDispSSz() const { return cnt(ss_z_, "tangential_disp_ss_z"); }
  auto sliceTangentialDispIdx() const {
    return Slice<int, MaxCnts>(const_cast<int*>(ss_idx_.data()), n_, "tangential_disp_idx", mirror_.get());
  }
  auto sliceTotalNoTangentialContacts() const {
    return Slice<int, 1>(const_cast<int*>(ss_cnt_.data()), n_, "total_no_tangential_contacts", mirror_.get());
  }

  // src/CabanaDEM_Particles.hpp:273-297.  New entries are zero (Kokkos
  // zero-initialises; the drivers set ids to -1 themselves).
  void resize(const std::size_t n) {
    if (mirror_) mirror_->before_host_access();
    n_ = n;
    for (auto* v : {&x_, &u_, &au_, &force_, &torque_, &omega_}) v->resize(3 * n, 0.);
    for (auto* v : {&m_, &rho_, &rad_, &E_, &nu_, &G_, &I_}) v->resize(n, 0.);
    for (auto* v : {&sw_x_, &sw_y_, &sw_z_}) v->resize(MaxNoWalls * n, 0.);
    for (auto* v : {&ss_x_, &ss_y_, &ss_z_}) v->resize(MaxCnts * n, 0.);
    ss_idx_.resize(MaxCnts * n, 0);
    ss_cnt_.resize(n, 0);
    if (dev_) dev_->reset();
  }

  // src/CabanaDEM_Particles.hpp:299-346.  With CABANADEM_ENABLE_HDF5 this
  // writes the reference's HDF5 step file: <output_folder>/particles_<step>.h5
  // with the fields it passes to Cabana's writer (:309-326), each under its
  // slice label, a "Time" attribute, and the matching .xmf.  Without HDF5 it
  // writes those fields (the per-contact histories aside) plus the torques
  // as one CSV file, <output_folder>/particles_<step>.csv, the time in the
  // header.
  void output(const int output_step, const double output_time, const bool use_reference = true) {
    (void)use_reference;
    // a read-only access: host edits made before this call must still be
    // uploaded by the next device call, so only this call's mark is undone
    const bool touched = mirror_->host_touched;
    mirror_->before_host_access();
    mirror_->host_touched = touched;
#ifdef CABANADEM_ENABLE_HDF5
    using detail::H5Field;
    const std::vector<H5Field> fields = {
        {"positions", x_.data(), 3, false},       {"velocities", u_.data(), 3, false},
        {"omega", omega_.data(), 3, false},       {"accelerations", au_.data(), 3, false},
        {"forces", force_.data(), 3, false},      {"mass", m_.data(), 1, false},
        {"density", rho_.data(), 1, false},       {"radius", rad_.data(), 1, false},
        {"tangential_disp_ss_x", ss_x_.data(), MaxCnts, false},
        {"tangential_disp_ss_y", ss_y_.data(), MaxCnts, false},
        {"tangential_disp_ss_z", ss_z_.data(), MaxCnts, false},
        {"tangential_disp_idx", ss_idx_.data(), MaxCnts, true},
        {"total_no_tangential_contacts", ss_cnt_.data(), 1, true}};
    if (!detail::write_h5_step(output_folder_name_ + "/particles", output_step, output_time, n_, fields))
      std::cout << "particle output: writing " << output_folder_name_ << "/particles_" << output_step
                << ".h5 failed" << std::endl;
    return;
#endif
    std::string fn = output_folder_name_ + "/particles_" + std::to_string(output_step) + ".csv";
    std::ofstream os(fn);
    if (!os) {
      std::cout << "No particle output enabled.";
      return;
    }
    os << "# Time=" << std::setprecision(17) << output_time << "\n";
    os << "positions_0,positions_1,positions_2,velocities_0,velocities_1,velocities_2,"
          "omega_0,omega_1,omega_2,forces_0,forces_1,forces_2,torques_0,torques_1,torques_2,"
          "mass,density,radius,total_no_tangential_contacts\n";
    os << std::setprecision(17);
    for (std::size_t i = 0; i < n_; ++i) {
      for (auto* v : {&x_, &u_, &omega_, &force_, &torque_})
        for (int d = 0; d < 3; ++d) os << (*v)[3 * i + d] << ",";
      os << m_[i] << "," << rho_[i] << "," << rad_[i] << "," << ss_cnt_[i] << "\n";
    }
  }

  std::size_t size() const { return n_; }
  const std::string& outputFolder() const { return output_folder_name_; }

  // Device binding used by the solver, force and integrator classes.
  detail::DeviceBinding& device() { return *dev_; }
  detail::Mirror& mirror() { return *mirror_; }

 private:
  Slice<double, 3> vec(const std::vector<double>& v, const char* label) const {
    return Slice<double, 3>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }
  Slice<double, 1> sca(const std::vector<double>& v, const char* label) const {
    return Slice<double, 1>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }
  Slice<double, MaxNoWalls> wall(const std::vector<double>& v, const char* label) const {
    return Slice<double, MaxNoWalls>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }
  Slice<double, MaxCnts> cnt(const std::vector<double>& v, const char* label) const {
    return Slice<double, MaxCnts>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }

  void fill_view(dem_particles_view& v) {
    v.n = (int)n_;
    v.max_cnts = MaxCnts;
    v.max_walls = MaxNoWalls;
    v.x = x_.data(); v.u = u_.data(); v.omega = omega_.data();
    v.force = force_.data(); v.torque = torque_.data();
    v.m = m_.data(); v.rad = rad_.data(); v.E = E_.data(); v.nu = nu_.data();
    v.G = G_.data(); v.I = I_.data();
    v.sw_x = sw_x_.data(); v.sw_y = sw_y_.data(); v.sw_z = sw_z_.data();
    v.ss_x = ss_x_.data(); v.ss_y = ss_y_.data(); v.ss_z = ss_z_.data();
    v.ss_idx = ss_idx_.data(); v.ss_cnt = ss_cnt_.data();
  }

  std::size_t n_ = 0;
  std::string output_folder_name_;
  std::vector<double> x_, u_, au_, force_, torque_, omega_;
  std::vector<double> m_, rho_, rad_, E_, nu_, G_, I_;
  std::vector<double> sw_x_, sw_y_, sw_z_;
  std::vector<double> ss_x_, ss_y_, ss_z_;
  std::vector<int> ss_idx_, ss_cnt_;
  std::unique_ptr<detail::Mirror> mirror_;
  std::unique_ptr<detail::DeviceBinding> dev_;
};

} // namespace CabanaDEM
