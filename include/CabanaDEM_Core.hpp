// CabanaDEM_Core.hpp -- spaces, tags, host slices and the device binding
// behind the reference-shaped C++ API.
//
// The reference runs its containers on Kokkos/Cabana (AoSoA + slices,
// src/CabanaDEM_Particles.hpp:30-46).  Here the containers keep host
// mirrors in the reference layout; a Slice is a host view of one field
// with operator()(i) / operator()(i, d) like Cabana::slice.  Compute runs
// on the GPU through the C ABI (cabanadem.h): the first device call uploads
// the mirror, every later host access through a slice first downloads
// whatever the device has newer.
#pragma once

#include <array>
#include <cstddef>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cabanadem.h"

namespace CabanaDEM {

// Memory / execution spaces.  HIPSpace runs the kernels on the GPU; the
// drivers' `typename exec_space::memory_space` works as with Kokkos.
struct HIPSpace {
  using memory_space = HIPSpace;
  using execution_space = HIPSpace;
  static const char* name() { return "HIP"; }
};
struct HostSpace {
  using memory_space = HostSpace;
  using execution_space = HostSpace;
  static const char* name() { return "Host"; }
};
using DefaultExecutionSpace = HIPSpace;

// Neighbour-list tags with the names Cabana uses (src/CabanaDEM_Solver.hpp:25-28).
struct FullNeighborTag {};
struct HalfNeighborTag {};
struct VerletLayout2D {};
struct TeamOpTag {};
struct SerialOpTag {};
struct FirstNeighborsTag {};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const dem_ctx* ctx) {
  if (rc != DEM_OK) throw Error(rc, dem_last_error(ctx));
}

namespace detail {

struct DeviceBinding;

// Host/device coherence of one container's mirror.
struct Mirror {
  DeviceBinding* binding = nullptr; // device side of the owning container
  bool device_newer = false;  // the device holds results not yet downloaded
  bool host_touched = true;   // the host arrays may differ from the device
  std::function<void()> pull; // download into the host arrays
  void before_host_access() {
    if (device_newer && pull) {
      device_newer = false;
      pull();
    }
    host_touched = true;
  }
};

} // namespace detail

// Host view of one field: Width components per particle (1 = scalar).
template <class T, int Width>
class Slice {
 public:
  Slice() = default;
  Slice(T* data, std::size_t n, std::string label, detail::Mirror* m)
      : data_(data), n_(n), label_(std::move(label)), m_(m) {}

  T& operator()(std::size_t i) const {
    touch();
    return data_[i * Width];
  }
  T& operator()(std::size_t i, std::size_t d) const {
    touch();
    return data_[i * Width + d];
  }
  std::size_t size() const { return n_; }
  std::string label() const { return label_; }
  T* data() const {
    touch();
    return data_;
  }
  detail::Mirror* mirror() const { return m_; }

 private:
  void touch() const {
    if (m_) m_->before_host_access();
  }
  T* data_ = nullptr;
  std::size_t n_ = 0;
  std::string label_;
  detail::Mirror* m_ = nullptr;
};

namespace detail {

inline void make_output_dir(const std::string& dir) {
  namespace fs = std::filesystem;
  if (!dir.empty() && (!fs::is_directory(dir) || !fs::exists(dir))) fs::create_directory(dir);
}

// The device side of one particle container: a dem_ctx plus the
// parameters it was created with.  Structural parameters (count,
// capacities, cutoff, binning box) recreate the context when they change;
// the others go through dem_set_params.
struct DeviceBinding {
  dem_ctx* ctx = nullptr;
  dem_config cfg{};
  int generation = 0;            // bumped on every (re)create
  const void* walls_owner = nullptr;
  int walls_generation = -1;     // context generation the walls were uploaded to
  long walls_version = -1;
  bool list_requested = false; // a neighbour list has been asked for
  std::function<void(dem_particles_view&)> fill_view; // host arrays of the owner
  std::function<int()> max_cnts;
  std::function<int()> max_walls;
  std::function<int()> count;
  Mirror* mirror = nullptr;
  // Optional state beyond the particle view (bonded containers): run on a
  // new context before its first upload, and beside every upload/download.
  std::function<void()> after_create;
  std::function<void()> upload_extra;
  std::function<void()> download_extra;

  DeviceBinding() {
    cfg.max_contacts = 16;
    cfg.dt = 1e-6;
    cfg.cor_pp = cfg.cor_pw = 1.0;
    cfg.cutoff = 0.;
    cfg.rebuild_every = 1;
  }
  ~DeviceBinding() { reset(); }
  DeviceBinding(const DeviceBinding&) = delete;
  DeviceBinding& operator=(const DeviceBinding&) = delete;

  void reset() {
    if (ctx) dem_destroy(ctx);
    ctx = nullptr;
  }

  void download() {
    if (!ctx) return;
    dem_particles_view v{};
    fill_view(v);
    check(dem_download_particles(ctx, &v), ctx);
    if (download_extra) download_extra();
  }
  void upload() {
    dem_particles_view v{};
    fill_view(v);
    check(dem_upload_particles(ctx, &v), ctx);
    if (upload_extra) upload_extra();
  }

  static bool structural_equal(const dem_config& a, const dem_config& b) {
    bool same = a.n == b.n && a.max_contacts == b.max_contacts && a.cutoff == b.cutoff &&
                a.max_neighbors == b.max_neighbors;
    for (int d = 0; d < 3; ++d) same = same && a.lo[d] == b.lo[d] && a.hi[d] == b.hi[d];
    return same;
  }

  // Makes the context match `want`, carrying the state across a recreate,
  // and uploads the host mirror if the host may have changed it.
  void ensure(dem_config want) {
    want.n = count();
    want.max_contacts = max_cnts();
    if (!(want.cutoff > 0.)) want.cutoff = cfg.cutoff > 0. ? cfg.cutoff : default_cutoff();
    bool box_ok = true;
    for (int d = 0; d < 3; ++d) box_ok = box_ok && want.hi[d] > want.lo[d];
    if (!box_ok) default_box(want);
    if (ctx && !structural_equal(want, cfg)) {
      if (mirror && mirror->device_newer) {
        mirror->device_newer = false;
        download();
      }
      reset();
      if (mirror) mirror->host_touched = true;
    }
    if (!ctx) {
      want.n_walls = 0;
      check(dem_create(&want, &ctx), nullptr);
      cfg = want;
      ++generation;
      if (after_create) after_create();
      if (mirror) mirror->host_touched = true;
    } else {
      check(dem_set_params(ctx, &want), ctx);
      int walls = cfg.n_walls;
      cfg = want;
      cfg.n_walls = walls;
    }
    if (mirror && mirror->host_touched) {
      if (mirror->device_newer) { // host touched after a run: take the device state first
        mirror->device_newer = false;
        download();
      }
      upload();
      mirror->host_touched = false;
    }
  }

  // Uploads the walls when they changed or the context is new.
  void bind_walls(const void* owner, long version, const dem_walls_view& w) {
    if (owner == walls_owner && version == walls_version && walls_generation == generation) return;
    check(dem_upload_walls(ctx, &w), ctx);
    walls_owner = owner;
    walls_version = version;
    walls_generation = generation;
    cfg.n_walls = w.n;
  }

  void after_compute() {
    if (mirror) mirror->device_newer = true;
  }

  double default_cutoff() {
    // three radii, as the reference drivers pass delta = 3 r
    dem_particles_view v{};
    fill_view(v);
    double rmax = 0.;
    for (int i = 0; i < count(); ++i) rmax = v.rad[i] > rmax ? v.rad[i] : rmax;
    return rmax > 0. ? 3. * rmax : 1.;
  }
  void default_box(dem_config& c) {
    dem_particles_view v{};
    fill_view(v);
    for (int d = 0; d < 3; ++d) {
      double lo = 1e300, hi = -1e300;
      for (int i = 0; i < count(); ++i) {
        lo = v.x[3 * i + d] < lo ? v.x[3 * i + d] : lo;
        hi = v.x[3 * i + d] > hi ? v.x[3 * i + d] : hi;
      }
      if (!(hi >= lo)) { lo = 0.; hi = 1.; }
      c.lo[d] = lo - c.cutoff;
      c.hi[d] = hi + c.cutoff;
    }
  }
};

} // namespace detail
} // namespace CabanaDEM
