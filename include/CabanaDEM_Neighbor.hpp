// CabanaDEM_Neighbor.hpp -- the full Verlet list the solver builds.
//
// Stands in for Cabana::VerletList<MemorySpace, FullNeighborTag,
// VerletLayout2D, TeamOpTag> as constructed and rebuilt at
// src/CabanaDEM_Solver.hpp:60-63,87-88 (x slice, index range, cutoff delta,
// cell ratio, mesh bounds).  The list lives on the GPU next to the particles
// of the slice it was built from: every j != i with |x_i - x_j|^2 <= delta^2.
#pragma once

#include <vector>

#include "CabanaDEM_Core.hpp"

namespace CabanaDEM {

template <class MemorySpace, class AlgorithmTag = FullNeighborTag, class LayoutTag = VerletLayout2D,
          class BuildTag = TeamOpTag>
class VerletList {
 public:
  using memory_space = MemorySpace;

  VerletList() = default;

  template <class SliceType>
  VerletList(const SliceType& x, std::size_t begin, std::size_t end, double neighborhood_radius,
             double cell_size_ratio, const double grid_min[3], const double grid_max[3]) {
    build(x, begin, end, neighborhood_radius, cell_size_ratio, grid_min, grid_max);
  }

  // The cell ratio only sizes Cabana's bins; the set is independent of it.
  template <class SliceType>
  void build(const SliceType& x, std::size_t begin, std::size_t end, double neighborhood_radius,
             double cell_size_ratio, const double grid_min[3], const double grid_max[3]) {
    (void)cell_size_ratio;
    detail::Mirror* m = x.mirror();
    if (!m || !m->binding) throw Error(DEM_ERR_ARG, "VerletList needs the position slice of a Particles container");
    if (begin != 0 || end != x.size())
      throw Error(DEM_ERR_ARG, "VerletList over a sub-range of the particles is not supported");
    binding_ = m->binding;
    dem_config want = binding_->cfg;
    want.cutoff = neighborhood_radius;
    for (int d = 0; d < 3; ++d) {
      want.lo[d] = grid_min[d];
      want.hi[d] = grid_max[d];
    }
    binding_->ensure(want);
    check(dem_build_neighbors(binding_->ctx), binding_->ctx);
    binding_->list_requested = true;
    generation_ = binding_->generation;
    radius_ = neighborhood_radius;
  }

  // Rebuilds if the particles' device context was recreated since.
  void refresh() const {
    if (binding_ && binding_->generation != generation_) {
      check(dem_build_neighbors(binding_->ctx), binding_->ctx);
      generation_ = binding_->generation;
    }
  }

  double radius() const { return radius_; }

  // Rows in original particle ids, ascending (a host copy, for inspection).
  void neighbors(std::vector<int>& counts, std::vector<int>& ids, int max_n) const {
    counts.assign(binding_->cfg.n, 0);
    ids.assign((std::size_t)binding_->cfg.n * max_n, -1);
    check(dem_get_neighbors(binding_->ctx, counts.data(), ids.data(), max_n), binding_->ctx);
  }

 private:
  detail::DeviceBinding* binding_ = nullptr;
  mutable int generation_ = -1;
  double radius_ = 0.;
};

} // namespace CabanaDEM
