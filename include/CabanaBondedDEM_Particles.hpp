// CabanaBondedDEM_Particles.hpp -- bonded particle container with the
// reference's API (src/CabanaBondedDEM_Particles.hpp:29-294): the 28 fields,
// their slice getters and labels, update_mesh_limits, updateParticles,
// setupBonds, resize and output.  Storage is a host mirror in the reference
// layout (bond rows [n][MaxBonds]); the GPU copy lives behind the C ABI.
//
// Differences from the reference, all deliberate:
//  - bond history and rest length are read at bond slot k, not at the
//    partner id j (src/CabanaBondedDEM_Force.hpp:207,231-233);
//  - bond_fn_* is kept for the API but, as in the reference, never written;
//  - bond parameters (beta 0.1 and radius factor 1, hard-coded at
//    src/CabanaBondedDEM_Force.hpp:186-188), tensile breakage and contacts
//    between unbonded pairs are set through bond_parameters;
//  - setExternalForce / setGroups run the tst_01 driver's external force and
//    boundary condition on the GPU instead of through host slices.
#pragma once

#include <algorithm>
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

namespace CabanaBondedDEM {

using CabanaDEM::Error;
using CabanaDEM::Slice;
using CabanaDEM::check;

// Bond model scalars (extension; defaults reproduce the reference).
struct BondParameters {
  double beta_bond = 0.1;          // src/CabanaBondedDEM_Force.hpp:186
  double radius_factor = 1.0;      // :187-188
  double tensile_strength = 0.0;   // <= 0: unbreakable
  bool unbonded_contacts = false;  // Hertz contacts between unbonded pairs
  double bond_factor = 3.0;        // setupBonds: r_ij < factor r_i (:130)
  bool operator==(const BondParameters& o) const {
    return beta_bond == o.beta_bond && radius_factor == o.radius_factor &&
           tensile_strength == o.tensile_strength && unbonded_contacts == o.unbonded_contacts;
  }
};

template <class MemorySpace, int Dimension, int MaxBonds>
class ParticlesBondedDEM {
 public:
  using memory_space = MemorySpace;
  using execution_space = typename memory_space::execution_space;
  static constexpr int dim = Dimension;
  static constexpr int max_bonds = MaxBonds;
  static constexpr int max_contacts = 8; // unbonded-pair contacts (extension)
  static constexpr int max_walls = 1;
  static_assert(Dimension == 3, "three-dimensional particles only");
  static_assert(MaxBonds >= 1 && MaxBonds <= DEM_MAX_BONDS, "MaxBonds must be in [1, 32]");

  double delta = 0.;
  BondParameters bond_parameters;

  // src/CabanaBondedDEM_Particles.hpp:52-61
  template <class ExecSpace>
  ParticlesBondedDEM(const ExecSpace&, std::size_t no_of_particles)
      : mirror_(std::make_unique<CabanaDEM::detail::Mirror>()),
        dev_(std::make_unique<CabanaDEM::detail::DeviceBinding>()) {
    mesh_lo_ = {0., 0., 0.};
    mesh_hi_ = {0., 0., 0.};
    resize(no_of_particles);
    mirror_->pull = [this] { dev_->download(); };
    dev_->mirror = mirror_.get();
    mirror_->binding = dev_.get();
    dev_->fill_view = [this](dem_particles_view& v) { fill_view(v); };
    dev_->count = [this] { return (int)n_; };
    dev_->max_cnts = [] { return max_contacts; };
    dev_->max_walls = [] { return max_walls; };
    dev_->after_create = [this] { on_create(); };
    dev_->upload_extra = [this] {
      dem_bonds_view b = bonds_view();
      check(dem_upload_bonds(dev_->ctx, &b), dev_->ctx);
    };
    dev_->download_extra = [this] {
      dem_bonds_view b = bonds_view();
      check(dem_download_bonds(dev_->ctx, &b), dev_->ctx);
    };
  }

  // src/CabanaBondedDEM_Particles.hpp:62-66
  void update_mesh_limits(const std::array<double, 3>& lo, const std::array<double, 3>& hi) {
    mesh_lo_ = lo;
    mesh_hi_ = hi;
  }

  // src/CabanaBondedDEM_Particles.hpp:79-84: the functor runs for every
  // particle on the host mirror; the next device call uploads it.
  template <class ExecSpace, class FunctorType>
  void updateParticles(const ExecSpace&, const FunctorType& functor) {
    mirror_->before_host_access();
    for (std::size_t pid = 0; pid < n_; ++pid) functor((int)pid);
  }

  // src/CabanaBondedDEM_Particles.hpp:85-171: bond every neighbour with
  // r_ij < 3 r_i (bond_parameters.bond_factor), rest length r_ij, in the
  // next free slot.  The reference searches a 4 r_0 Verlet list
  // (:155-157); here the list radius covers factor * max r, so every
  // partner the criterion admits is found for any radius distribution.
  template <class ExecSpace>
  void setupBonds() {
    auto& dev = prepare(-1.);
    check(dem_setup_bonds(dev.ctx, bond_parameters.bond_factor), dev.ctx);
    dev.after_compute();
  }

  auto slicePosition() const { return vec(x_, "x"); }
  auto sliceVelocity() const { return vec(u_, "u"); }
  auto sliceAcceleration() const { return vec(au_, "au"); }
  auto sliceForce() const { return vec(force_, "force"); }
  auto sliceTorque() const { return vec(torque_, "torque"); }
  auto sliceOmega() const { return vec(omega_, "omega"); }
  auto sliceMass() const { return sca(m_, "m"); }
  auto sliceDensity() const { return sca(rho_, "rho"); }
  auto sliceRadius() const { return sca(rad_, "rad"); }
  auto sliceYoungsMod() const { return sca(E_, "E"); }
  auto slicePoissonsRatio() const { return sca(nu_, "nu"); }
  auto sliceShearMod() const { return sca(G_, "G"); }
  auto sliceMomentOfInertia() const { return sca(I_, "I"); }
  auto sliceBondFnX() const { return bond(fn_x_, "bond_fn_x"); }
  auto sliceBondFnY() const { return bond(fn_y_, "bond_fn_y"); }
  auto sliceBondFnZ() const { return bond(fn_z_, "bond_fn_z"); }
  auto sliceBondFtX() const { return bond(ft_x_, "bond_ft_x"); }
  auto sliceBondFtY() const { return bond(ft_y_, "bond_ft_y"); }
  auto sliceBondFtZ() const { return bond(ft_z_, "bond_ft_z"); }
  auto sliceBondMnX() const { return bond(mn_x_, "bond_mn_x"); }
  auto sliceBondMnY() const { return bond(mn_y_, "bond_mn_y"); }
  auto sliceBondMnZ() const { return bond(mn_z_, "bond_mn_z"); }
  auto sliceBondMtX() const { return bond(mt_x_, "bond_mt_x"); }
  auto sliceBondMtY() const { return bond(mt_y_, "bond_mt_y"); }
  auto sliceBondMtZ() const { return bond(mt_z_, "bond_mt_z"); }
  auto sliceBondIdx() const {
    return Slice<int, MaxBonds>(const_cast<int*>(idx_.data()), n_, "bond_idx", mirror_.get());
  }
  auto sliceBondInitLen() const { return bond(init_len_, "bond_init_len"); }
  auto sliceTotalNoBonds() const {
    return Slice<int, 1>(const_cast<int*>(count_.data()), n_, "total_no_bonds", mirror_.get());
  }

  // src/CabanaBondedDEM_Particles.hpp:203-233.  New entries are zero, bond
  // ids -1.
  void resize(std::size_t n) {
    if (mirror_) mirror_->before_host_access();
    n_ = n;
    for (auto* v : {&x_, &u_, &au_, &force_, &torque_, &omega_}) v->resize(3 * n, 0.);
    for (auto* v : {&m_, &rho_, &rad_, &E_, &nu_, &G_, &I_}) v->resize(n, 0.);
    for (auto* v : {&fn_x_, &fn_y_, &fn_z_, &ft_x_, &ft_y_, &ft_z_, &mn_x_, &mn_y_, &mn_z_,
                    &mt_x_, &mt_y_, &mt_z_, &init_len_})
      v->resize((std::size_t)MaxBonds * n, 0.);
    idx_.resize((std::size_t)MaxBonds * n, -1);
    count_.resize(n, 0);
    fext_.clear();
    group_.clear();
    if (dev_) dev_->reset();
  }

  // src/CabanaBondedDEM_Particles.hpp:235-271.  With CABANADEM_ENABLE_HDF5
  // this writes the reference's HDF5 step file, <folder>/particles_<step>.h5
  // (+ .xmf), with the fields it passes to Cabana's writer (:244-267) under
  // their slice labels and a "Time" attribute; without HDF5, the same fields
  // by their labels as <folder>/particles_<step>.csv (bond rows as MaxBonds
  // columns each).
  void output(int step, double time, bool use_reference = true) {
    (void)use_reference;
    // a read-only access: host edits made before this call must still be
    // uploaded by the next device call, so only this call's mark is undone
    const bool touched = mirror_->host_touched;
    mirror_->before_host_access();
    mirror_->host_touched = touched;
#ifdef CABANADEM_ENABLE_HDF5
    using CabanaDEM::detail::H5Field;
    const std::vector<H5Field> fields = {
        {"x", x_.data(), 3, false},          {"u", u_.data(), 3, false},
        {"omega", omega_.data(), 3, false},  {"au", au_.data(), 3, false},
        {"force", force_.data(), 3, false},  {"m", m_.data(), 1, false},
        {"rho", rho_.data(), 1, false},      {"rad", rad_.data(), 1, false},
        {"bond_fn_x", fn_x_.data(), MaxBonds, false}, {"bond_fn_y", fn_y_.data(), MaxBonds, false},
        {"bond_fn_z", fn_z_.data(), MaxBonds, false}, {"bond_ft_x", ft_x_.data(), MaxBonds, false},
        {"bond_ft_y", ft_y_.data(), MaxBonds, false}, {"bond_ft_z", ft_z_.data(), MaxBonds, false},
        {"bond_mn_x", mn_x_.data(), MaxBonds, false}, {"bond_mn_y", mn_y_.data(), MaxBonds, false},
        {"bond_mn_z", mn_z_.data(), MaxBonds, false}, {"bond_mt_x", mt_x_.data(), MaxBonds, false},
        {"bond_mt_y", mt_y_.data(), MaxBonds, false}, {"bond_mt_z", mt_z_.data(), MaxBonds, false},
        {"bond_idx", idx_.data(), MaxBonds, true},
        {"bond_init_len", init_len_.data(), MaxBonds, false},
        {"total_no_bonds", count_.data(), 1, true}};
    if (!CabanaDEM::detail::write_h5_step(output_folder_ + "/particles", step, time, n_, fields))
      std::cout << "particle output: writing " << output_folder_ << "/particles_" << step
                << ".h5 failed" << std::endl;
    return;
#endif
    const std::string fn = output_folder_ + "/particles_" + std::to_string(step) + ".csv";
    std::ofstream os(fn);
    if (!os) {
      std::cout << "No particle output enabled." << std::endl;
      return;
    }
    os << "# Time=" << std::setprecision(17) << time << "\n";
    os << "x_0,x_1,x_2,u_0,u_1,u_2,omega_0,omega_1,omega_2,au_0,au_1,au_2,force_0,force_1,"
          "force_2,m,rho,rad,total_no_bonds";
    for (const char* f : {"bond_idx", "bond_init_len", "bond_ft_x", "bond_ft_y", "bond_ft_z",
                          "bond_mn_x", "bond_mn_y", "bond_mn_z", "bond_mt_x", "bond_mt_y",
                          "bond_mt_z"})
      for (int k = 0; k < MaxBonds; ++k) os << "," << f << "_" << k;
    os << "\n" << std::setprecision(17);
    for (std::size_t i = 0; i < n_; ++i) {
      for (auto* v : {&x_, &u_, &omega_, &au_, &force_})
        for (int d = 0; d < 3; ++d) os << (*v)[3 * i + d] << ",";
      os << m_[i] << "," << rho_[i] << "," << rad_[i] << "," << count_[i];
      for (int k = 0; k < MaxBonds; ++k) os << "," << idx_[i * MaxBonds + k];
      for (auto* v : {&init_len_, &ft_x_, &ft_y_, &ft_z_, &mn_x_, &mn_y_, &mn_z_, &mt_x_, &mt_y_, &mt_z_})
        for (int k = 0; k < MaxBonds; ++k) os << "," << (*v)[i * MaxBonds + k];
      os << "\n";
    }
  }
  void setOutputFolder(const std::string& dir) {
    output_folder_ = dir;
    CabanaDEM::detail::make_output_dir(dir);
  }

  // ---- extensions ----------------------------------------------------------
  // Constant external force per particle, [n][3] (applyTensileForce of
  // examples/tst_01_BDEM_two_particles_tensile_3D.cpp:69-117, on the GPU).
  void setExternalForce(const std::vector<double>& f) {
    if (!f.empty() && f.size() != 3 * n_) throw Error(DEM_ERR_ARG, "external force must be [n][3]");
    fext_ = f;
    if (dev_->ctx) check(dem_set_external_force(dev_->ctx, fext_.empty() ? nullptr : fext_.data()), dev_->ctx);
  }
  // group[i] in 0..kinds.size(); kinds[g] = DEM_GROUP_*; vel[g] = prescribed
  // velocity of a DEM_GROUP_VELOCITY group.  DEM_GROUP_FIXED_FORCE is the
  // driver's applyBoundaryConditions (:18-66).
  void setGroups(const std::vector<int>& group, const std::vector<int>& kinds,
                 const std::vector<std::array<double, 3>>& vel = {}) {
    if (!group.empty() && group.size() != n_) throw Error(DEM_ERR_ARG, "group must have n entries");
    group_ = group;
    kinds_ = kinds;
    vel_.assign(3 * kinds.size(), 0.);
    for (std::size_t g = 0; g < vel.size() && g < kinds.size(); ++g)
      for (int d = 0; d < 3; ++d) vel_[3 * g + d] = vel[g][d];
    if (dev_->ctx) apply_groups();
  }
  int64_t bondsBroken() {
    if (!dev_->ctx) return 0;
    dem_counters c{};
    check(dem_get_counters(dev_->ctx, &c), dev_->ctx);
    return c.bonds_broken;
  }

  std::size_t size() const { return n_; }
  const std::array<double, 3>& meshLo() const { return mesh_lo_; }
  const std::array<double, 3>& meshHi() const { return mesh_hi_; }

  CabanaDEM::detail::DeviceBinding& device() { return *dev_; }
  CabanaDEM::detail::Mirror& mirror() { return *mirror_; }

  // Makes the device context current: mesh box, a list radius that covers
  // the bond criterion, dt (when > 0) and the bond parameters.
  CabanaDEM::detail::DeviceBinding& prepare(double dt, int physics_mode = -1) {
    auto& dev = *dev_;
    dem_config want = dev.cfg;
    if (dt > 0.) want.dt = dt;
    if (physics_mode >= 0) want.physics_mode = physics_mode;
    double rmax = 0.;
    for (double r : rad_) rmax = std::max(rmax, r);
    const double need = bond_parameters.bond_factor * rmax * (1. + 1e-12);
    const double contact = 2. * rmax;
    want.cutoff = std::max({need, contact, delta});
    bool box = true;
    for (int d = 0; d < 3; ++d) box = box && mesh_hi_[d] > mesh_lo_[d];
    for (int d = 0; d < 3; ++d) {
      want.lo[d] = box ? mesh_lo_[d] : 0.;
      want.hi[d] = box ? mesh_hi_[d] : 0.;
    }
    // no neighbour list is needed between set-up and the end of the run
    // unless unbonded pairs may touch
    want.rebuild_every = bond_parameters.unbonded_contacts ? 1 : (1 << 30);
    dev.ensure(want);
    if (!(applied_ == bond_parameters)) enable_bonds();
    return dev;
  }

 private:
  void enable_bonds() {
    dem_bond_config b{};
    b.max_bonds = MaxBonds;
    b.unbonded_contacts = bond_parameters.unbonded_contacts ? 1 : 0;
    b.beta_bond = bond_parameters.beta_bond;
    b.radius_factor = bond_parameters.radius_factor;
    b.tensile_strength = bond_parameters.tensile_strength;
    check(dem_enable_bonds(dev_->ctx, &b), dev_->ctx);
    applied_ = bond_parameters;
  }
  void apply_groups() {
    if (group_.empty()) {
      check(dem_set_groups(dev_->ctx, nullptr, 0, nullptr, nullptr), dev_->ctx);
      return;
    }
    check(dem_set_groups(dev_->ctx, group_.data(), (int)kinds_.size(), kinds_.data(), vel_.data()),
          dev_->ctx);
  }
  // A new context: bonds on, then the extension state.
  void on_create() {
    enable_bonds();
    if (!fext_.empty()) check(dem_set_external_force(dev_->ctx, fext_.data()), dev_->ctx);
    if (!group_.empty()) apply_groups();
  }

  Slice<double, 3> vec(const std::vector<double>& v, const char* label) const {
    return Slice<double, 3>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }
  Slice<double, 1> sca(const std::vector<double>& v, const char* label) const {
    return Slice<double, 1>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }
  Slice<double, MaxBonds> bond(const std::vector<double>& v, const char* label) const {
    return Slice<double, MaxBonds>(const_cast<double*>(v.data()), n_, label, mirror_.get());
  }

  void fill_view(dem_particles_view& v) {
    v.n = (int)n_;
    v.max_cnts = max_contacts;
    v.max_walls = max_walls;
    v.x = x_.data(); v.u = u_.data(); v.omega = omega_.data();
    v.force = force_.data(); v.torque = torque_.data();
    v.m = m_.data(); v.rad = rad_.data(); v.E = E_.data(); v.nu = nu_.data();
    v.G = G_.data(); v.I = I_.data();
  }
  dem_bonds_view bonds_view() {
    dem_bonds_view b{};
    b.n = (int)n_;
    b.max_bonds = MaxBonds;
    b.idx = idx_.data(); b.count = count_.data(); b.init_len = init_len_.data();
    b.ft_x = ft_x_.data(); b.ft_y = ft_y_.data(); b.ft_z = ft_z_.data();
    b.mn_x = mn_x_.data(); b.mn_y = mn_y_.data(); b.mn_z = mn_z_.data();
    b.mt_x = mt_x_.data(); b.mt_y = mt_y_.data(); b.mt_z = mt_z_.data();
    return b;
  }

  std::size_t n_ = 0;
  std::string output_folder_ = ".";
  std::array<double, 3> mesh_lo_{}, mesh_hi_{};
  std::vector<double> x_, u_, au_, force_, torque_, omega_;
  std::vector<double> m_, rho_, rad_, E_, nu_, G_, I_;
  std::vector<double> fn_x_, fn_y_, fn_z_, ft_x_, ft_y_, ft_z_, mn_x_, mn_y_, mn_z_;
  std::vector<double> mt_x_, mt_y_, mt_z_, init_len_;
  std::vector<int> idx_, count_;
  std::vector<double> fext_;
  std::vector<int> group_, kinds_;
  std::vector<double> vel_;
  BondParameters applied_{-1., -1., -1., false, 0.};
  std::unique_ptr<CabanaDEM::detail::Mirror> mirror_;
  std::unique_ptr<CabanaDEM::detail::DeviceBinding> dev_;
};

} // namespace CabanaBondedDEM
