// CabanaDEM_Wall.hpp -- infinite-plane walls with the reference's API
// (src/CabanaDEM_Wall.hpp:22-209): position, velocity, omega, normal and
// material per wall.  In reference mode, as in the reference, wall velocity
// and omega are carried but not used by the contact force
// (src/CabanaDEM_Force.hpp:468-469,529).  In corrected mode a wall
// translates with its velocity from the position it has when the solver
// uploads it, and its surface moves with u + omega x (c - x_w) at a contact
// point c; the positions in this container stay those of the upload.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "CabanaDEM_Core.hpp"

namespace CabanaDEM {

template <class MemorySpace, int Dimension>
class Wall {
 public:
  using memory_space = MemorySpace;
  using execution_space = typename memory_space::execution_space;
  static constexpr int dim = Dimension;
  static_assert(Dimension == 3, "three-dimensional walls only");

  template <class ExecSpace>
  Wall(const ExecSpace&, std::size_t no_of_walls, std::string output_folder_name)
      : output_folder_name_(std::move(output_folder_name)), mirror_(std::make_unique<detail::Mirror>()) {
    detail::make_output_dir(output_folder_name_);
    resize(no_of_walls);
  }

  template <class ExecSpace, class FunctorType>
  void updateParticles(const ExecSpace, const FunctorType init_functor) {
    ++version_;
    for (std::size_t pid = 0; pid < n_; ++pid) init_functor((int)pid);
  }

  auto slicePosition() { return vec(x_, "positions"); }
  auto sliceVelocity() { return vec(u_, "velocities"); }
  auto sliceOmega() { return vec(omega_, "omega"); }
  auto sliceNormal() { return vec(normal_, "normal"); }
  auto sliceYoungsMod() { return sca(E_, "youngs_mod"); }
  auto slicePoissonsRatio() { return sca(nu_, "poissons_ratio"); }
  auto sliceShearMod() { return sca(G_, "shear_mod"); }

  void resize(const std::size_t n) {
    ++version_;
    n_ = n;
    for (auto* v : {&x_, &u_, &omega_, &normal_}) v->resize(3 * n, 0.);
    for (auto* v : {&E_, &nu_, &G_, &I_}) v->resize(n, 0.);
  }

  void output(const int, const double, const bool = true) {}

  std::size_t size() const { return n_; }

  // Changes on every host access through a slice, so the solver re-uploads.
  long version() const { return version_; }
  dem_walls_view view() const {
    dem_walls_view w{};
    w.n = (int)n_;
    w.x = x_.data();
    w.normal = normal_.data();
    w.E = E_.data();
    w.nu = nu_.data();
    w.G = G_.data();
    w.u = u_.data();
    w.omega = omega_.data();
    return w;
  }

 private:
  Slice<double, 3> vec(std::vector<double>& v, const char* label) {
    ++version_;
    return Slice<double, 3>(v.data(), n_, label, nullptr);
  }
  Slice<double, 1> sca(std::vector<double>& v, const char* label) {
    ++version_;
    return Slice<double, 1>(v.data(), n_, label, nullptr);
  }

  std::size_t n_ = 0;
  long version_ = 0;
  std::string output_folder_name_;
  std::vector<double> x_, u_, omega_, normal_, E_, nu_, G_, I_;
  std::unique_ptr<detail::Mirror> mirror_;
};

} // namespace CabanaDEM
