// dem_api.hip -- context, memory and the C ABI of include/cabanadem.h.
//
// The context owns every device array of one system (layout in
// dem_internal.h) and a stream of default priority.  Host arrays cross the
// ABI in the reference layout and original particle order; the context
// keeps the particles in cell order on the device and translates through
// the id array (sorted slot -> original id) on upload and download.
#include "dem_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

using namespace demhip;

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string& s) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = s;
}

template <class T>
T* dalloc(size_t count, hipError_t* err) {
  void* p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess && *err == hipSuccess) *err = e;
  return static_cast<T*>(p);
}

// src/CabanaDEM_Force.hpp:206-210 and :215 -- the damping prefactor
// (-2 sqrt(5/6)) * beta, evaluated on the host with the same operations.
double eta_prefactor(double cor) {
  double t1 = std::log(cor);
  double t2 = t1 * t1 + M_PI * M_PI;
  double beta = t1 / std::sqrt(t2);
  return -2. * std::sqrt(5. / 6.) * beta;
}

// defined below
Grid make_grid(const dem_config& c, long* ncell, int rows);

} // namespace

struct dem_ctx {
  dem_config cfg{};
  std::string err;
  hipStream_t stream = nullptr;
  int n = 0, ld = 0, max_n = 0, max_c = 0;
  long ncell = 0;
  double rc2 = 0.;

  // systems: one, or an ensemble of independent systems in one context.
  // System s owns slots [sys_off[s], sys_off[s+1]) and, for the force
  // kernels' tile size kForceBlocks[b], tiles [tile0_b[b][s], tile0_b[b][s+1]).
  int nsys = 1;
  std::vector<int> sys_off;
  std::vector<dem_config> sys_cfg; // each system's configuration (binning box)
  std::vector<Grid> grids;
  std::vector<long> kbase;
  std::vector<SysParams> h_sys;
  long ncell_cap = 0;        // cells the cell arrays hold
  int tile_rows = kTileRows; // y-rows per cell tile of the binning grids
  int bidx = 1;              // force-kernel tile size kForceBlocks[bidx] (no bonds)
  std::vector<int> tile0_b[3];
  int ntiles_b[3] = {0, 0, 0};
  int4* d_tiles_b[3] = {nullptr, nullptr, nullptr}; // ensembles only
  int* d_tile0_b[3] = {nullptr, nullptr, nullptr};
  // the current tile set: kForceBlocks[bonds_on ? 1 : bidx] (select_block)
  int ntiles = 0;
  int4* d_tiles = nullptr;
  int* d_tile0 = nullptr;
  SysParams* d_sys = nullptr;
  int* d_sysid = nullptr;    // ensembles only
  Grid* d_grids = nullptr;   // ensembles only
  long* d_kbase = nullptr;   // ensembles only
  double* d_sysdiag = nullptr; // [nsys][8]
  int grid_blocks = 0;       // force-kernel workgroups; 0 = one per tile

  // captured step cycles (dem_set_graph): valid while the launch arguments
  // (generation gen) and the buffer parity are those of the capture
  bool use_graph = false;
  int force_pass = 0;        // DEM_FORCE_PASS_* (systems without bonds)
  // translating walls (corrected mode): positions hold at step wall_step0;
  // the force pass of step s sits at (s + 1 - wall_step0) dt
  int64_t wall_step0 = 0;
  bool walls_move = false;

  // automatic rebuild cadence (rebuild_every = DEM_REBUILD_AUTO)
  int* d_work = nullptr;       // [tiles][block] particles with contacts (split force passes)
  int* d_work_tile = nullptr;  // [tiles] their number per tile
  int cus = 256;               // compute units of the device (autotuner launch grids)
  unsigned long long* d_disp = nullptr; // [2] max |x - xb|^2, max |u|^2 (bits)
  unsigned long long* h_disp = nullptr; // pinned host copy of d_disp (a pageable
                                        // destination costs a staged copy per check)
  hipEvent_t ev_disp = nullptr;         // recorded behind the copy into h_disp
  int auto_left = 0;         // force passes until the next displacement check
  int auto_first = 0;        // interval planned right after the last build
  int64_t auto_since = 0;    // step at which the positions were last set
  int64_t auto_checks = 0;   // displacement checks so far
  bool auto_due = false;     // rebuild before the next force pass
  double auto_v = -1.;       // largest speed at the last check (< 0: none yet)
  bool capturing = false;
  hipGraphExec_t gexec = nullptr;
  int gen = 0, g_gen = -1, g_parity = -1, g_K = -1;

  // double-buffered groups: P (xr, uw), S (mat, mi, I, id, ft, wall), C (contacts)
  double4* xr[2] = {nullptr, nullptr};
  double* uw[2] = {nullptr, nullptr};
  double4* mat[2] = {nullptr, nullptr};
  double2* mI[2] = {nullptr, nullptr};
  int* id[2] = {nullptr, nullptr};
  double* ft[2] = {nullptr, nullptr};
  double* sw[2] = {nullptr, nullptr};
  int* cmeta[2] = {nullptr, nullptr};
  int* cj[2] = {nullptr, nullptr};
  double* ch[2] = {nullptr, nullptr};
  int cP = 0, cS = 0, cC = 0;

  // bonded DEM, groups and external force: group S (permuted at rebuilds)
  bool bonds_on = false;
  dem_bond_config bcfg{};
  int max_b = 0;
  int* nbond[2] = {nullptr, nullptr};
  int* bj[2] = {nullptr, nullptr};
  double* bl[2] = {nullptr, nullptr};
  double* bh[2] = {nullptr, nullptr};
  double2* bEG[2] = {nullptr, nullptr};
  int* vgroup[2] = {nullptr, nullptr};
  int n_groups = 0;
  int gkind[DEM_MAX_GROUPS] = {};
  double gvel[DEM_MAX_GROUPS][3] = {};
  double* fext[2] = {nullptr, nullptr};

  int* nn = nullptr;
  int* nbr = nullptr;
  double* xb = nullptr;
  int* key = nullptr;
  int* rank = nullptr;
  int* key_sorted = nullptr;
  int* cell_count = nullptr;
  int* cell_start = nullptr;
  int* perm = nullptr;
  int* inv = nullptr;
  int* scan_tmp = nullptr;
  DevFlags* flags = nullptr;
  double* diag_part = nullptr;
  double* diag_rows = nullptr;
  double* h_diag_rows = nullptr; // pinned host copy of diag_rows
  int diag_rows_cap = 0;
  DevFlags* h_flags = nullptr;   // pinned host copy of flags

  bool list_valid = false;
  bool built_once = false;
  int64_t step = 0;
  int64_t rebuilds = 0;
  int diag_every = 0;
  std::vector<double> diag_hist; // rows of 9
  double last_sample[3] = {0., 0., 0.};
  int max_nn_seen = 0;

  // host copies of the static properties, original order
  std::vector<double> h_m, h_rad, h_E, h_nu, h_G, h_I;
  double max_rad = 0.; // largest uploaded radius (0: none uploaded yet)

  // timing
  bool timing = false;
  int timing_every = 1;    // time every timing_every-th force pass
  int64_t timing_seen = 0; // force passes eligible for timing so far
  bool mark_open = false;  // mark_begin recorded a start event
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_marks; // (slot, index of start event)
  double t_ms[2] = {0., 0.};
  int64_t t_launches[2] = {0, 0};

  ~dem_ctx() {
    for (int b = 0; b < 2; ++b) {
      hipFree(xr[b]); hipFree(uw[b]); hipFree(mat[b]); hipFree(mI[b]);
      hipFree(id[b]); hipFree(ft[b]); hipFree(sw[b]);
      hipFree(cmeta[b]); hipFree(cj[b]); hipFree(ch[b]);
      hipFree(nbond[b]); hipFree(bj[b]); hipFree(bl[b]); hipFree(bh[b]); hipFree(bEG[b]);
      hipFree(vgroup[b]); hipFree(fext[b]);
    }
    hipFree(nn); hipFree(nbr); hipFree(xb); hipFree(key); hipFree(rank); hipFree(key_sorted);
    hipFree(cell_count); hipFree(cell_start); hipFree(perm); hipFree(inv); hipFree(scan_tmp);
    hipFree(flags); hipFree(diag_part); hipFree(diag_rows); hipFree(d_disp);
    hipFree(d_work); hipFree(d_work_tile);
    if (h_disp) hipHostFree(h_disp);
    if (ev_disp) hipEventDestroy(ev_disp);
    if (h_diag_rows) hipHostFree(h_diag_rows);
    if (h_flags) hipHostFree(h_flags);
    for (int b = 0; b < 3; ++b) { hipFree(d_tiles_b[b]); hipFree(d_tile0_b[b]); }
    hipFree(d_sys); hipFree(d_sysid); hipFree(d_grids);
    hipFree(d_kbase); hipFree(d_sysdiag);
    if (gexec) hipGraphExecDestroy(gexec);
    for (auto e : ev_pool) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }

  int fail(int code, const std::string& msg) {
    err = msg;
    set_global_error(msg);
    return code;
  }
  int hip_fail(hipError_t e, const char* where) {
    return fail(DEM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
  }

  StepArgs step_args() const {
    StepArgs a{};
    a.n = n; a.ld = ld; a.max_n = max_n; a.max_c = max_c;
    a.physics_mode = cfg.physics_mode;
    a.dt = cfg.dt;
    a.hdt = 0.5 * cfg.dt; // src/CabanaDEM_Integrator.hpp:25
    for (int d = 0; d < 3; ++d) a.g[d] = cfg.gravity[d];
    a.gravity_on = (cfg.gravity[0] != 0. || cfg.gravity[1] != 0. || cfg.gravity[2] != 0.) ? 1 : 0;
    a.sys = d_sys;
    a.sys0 = h_sys[0];
    a.tiles = nsys > 1 ? d_tiles : nullptr;
    a.ntiles = ntiles;
    a.block = kForceBlocks[bonds_on ? 1 : bidx];
    a.xr_in = xr[cP]; a.xr_out = xr[cP ^ 1];
    a.uw_in = uw[cP]; a.uw_out = uw[cP ^ 1];
    a.mat = mat[cS]; a.mI = mI[cS]; a.ft = ft[cS];
    a.sw = sw[cS];
    a.nn = nn; a.nbr = nbr;
    a.cmeta_in = cmeta[cC]; a.cj_in = cj[cC]; a.ch_in = ch[cC];
    a.cmeta_out = cmeta[cC ^ 1]; a.cj_out = cj[cC ^ 1]; a.ch_out = ch[cC ^ 1];
    a.flags = flags;
    a.diag_part = diag_part;
    if (bonds_on) {
      a.max_b = max_b;
      a.nbond = nbond[cS]; a.bj = bj[cS]; a.bl = bl[cS]; a.bh = bh[cS]; a.bEG = bEG[cS];
      a.beta_bond = bcfg.beta_bond;
      a.bond_radius_factor = bcfg.radius_factor;
      a.tensile_strength = bcfg.tensile_strength;
      a.unbonded_contacts = bcfg.unbonded_contacts;
    }
    if (n_groups > 0) {
      a.vgroup = vgroup[cS];
      a.n_vgroups = n_groups;
      for (int g = 0; g < DEM_MAX_GROUPS; ++g) {
        a.vgkind[g] = gkind[g];
        for (int d = 0; d < 3; ++d) a.vgv[g][d] = gvel[g][d];
      }
    }
    a.fext = fext[cS];
    a.split = (force_pass == DEM_FORCE_PASS_SPLIT && !bonds_on) ? 1 : 0;
    a.work = d_work;
    a.work_tile = d_work_tile;
    a.wave = (force_pass != DEM_FORCE_PASS_FUSED && !bonds_on) ? 1 : 0;
    a.tw = (double)(step + 1 - wall_step0) * cfg.dt;
    return a;
  }

  BuildArgs build_args(bool check_skin, bool permute_ft = true) const {
    BuildArgs b{};
    b.keep_xb = cfg.skin > 0. ? 1 : 0;
    b.permute_ft = permute_ft ? 1 : 0;
    b.n = n; b.ld = ld; b.max_n = max_n; b.max_c = max_c;
    b.grid = grids[0]; b.ncell = ncell; b.rc2 = rc2;
    if (nsys > 1) { b.sysid = d_sysid; b.grids = d_grids; b.kbase = d_kbase; }
    double sh = 0.5 * cfg.skin;
    b.skin_half_sq = sh * sh;
    b.check_skin = (check_skin && cfg.skin > 0.) ? 1 : 0;
    b.xr_in = xr[cP]; b.xr_out = xr[cP ^ 1];
    b.uw_in = uw[cP]; b.uw_out = uw[cP ^ 1];
    b.mat_in = mat[cS]; b.mat_out = mat[cS ^ 1];
    b.mI_in = mI[cS]; b.mI_out = mI[cS ^ 1];
    b.id_in = id[cS]; b.id_out = id[cS ^ 1];
    b.ft_in = ft[cS]; b.ft_out = ft[cS ^ 1];
    b.sw_in = sw[cS]; b.sw_out = sw[cS ^ 1];
    b.cmeta_in = cmeta[cC]; b.cmeta_out = cmeta[cC ^ 1];
    b.cj_in = cj[cC]; b.cj_out = cj[cC ^ 1];
    b.ch_in = ch[cC]; b.ch_out = ch[cC ^ 1];
    b.xb = xb; b.key = key; b.rank = rank; b.key_sorted = key_sorted;
    b.cell_count = cell_count; b.cell_start = cell_start;
    b.perm = perm; b.inv = inv; b.scan_tmp = scan_tmp;
    b.nn = nn; b.nbr = nbr; b.flags = flags;
    b.bonds_on = bonds_on ? 1 : 0;
    b.max_b = max_b;
    if (bonds_on) {
      b.nbond_in = nbond[cS]; b.nbond_out = nbond[cS ^ 1];
      b.bj_in = bj[cS]; b.bj_out = bj[cS ^ 1];
      b.bl_in = bl[cS]; b.bl_out = bl[cS ^ 1];
      b.bh_in = bh[cS]; b.bh_out = bh[cS ^ 1];
      b.bEG_in = bEG[cS]; b.bEG_out = bEG[cS ^ 1];
    }
    if (n_groups > 0) { b.vgroup_in = vgroup[cS]; b.vgroup_out = vgroup[cS ^ 1]; }
    if (fext[cS]) { b.fext_in = fext[cS]; b.fext_out = fext[cS ^ 1]; }
    return b;
  }

  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, DEM_TIMING_EVENT_FLAGS) != hipSuccess) return nullptr;
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  // Events around a force pass (slot 0) or a rebuild (slot 1).  Force passes
  // are timed every timing_every-th one; passes that also sample diagnostics
  // (another kernel variant) or check the displacements (extra reads) never,
  // and without advancing the count.  Each pair of events costs the GPU a few
  // microseconds between the kernels.
  void mark_begin(int slot, bool skip = false) {
    mark_open = false;
    if (!timing || capturing) return;
    if (slot == 0 && (skip || timing_seen++ % timing_every != 0)) return;
    const size_t i = ev_used;
    hipEvent_t e0 = next_event();
    hipEvent_t e1 = e0 ? next_event() : nullptr;
    if (!e1) return;
    ev_marks.push_back({slot, i});
    hipEventRecord(e0, stream);
    mark_open = true;
  }
  void mark_end() {
    if (!mark_open) return;
    mark_open = false;
    hipEventRecord(ev_pool[ev_marks.back().second + 1], stream);
  }
  void collect_timing() {
    if (!timing) { ev_used = 0; ev_marks.clear(); return; }
    for (auto& m : ev_marks) {
      if (m.second + 1 >= ev_used) continue;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev_pool[m.second], ev_pool[m.second + 1]) == hipSuccess) {
        t_ms[m.first] += ms;
        t_launches[m.first] += 1;
      }
    }
    ev_used = 0;
    ev_marks.clear();
  }

  // The flags' read-back into pinned memory, enqueued (allocated on first use).
  hipError_t enqueue_flags() {
    hipError_t e = hipSuccess;
    if (!h_flags) e = hipHostMalloc(reinterpret_cast<void**>(&h_flags), sizeof(DevFlags), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemcpyAsync(h_flags, flags, sizeof(DevFlags), hipMemcpyDeviceToHost, stream);
    return e;
  }
  // enqueued: the caller has enqueued the read-back already
  int check_flags(DevFlags* out = nullptr, bool enqueued = false) {
    hipError_t e = enqueued ? hipSuccess : enqueue_flags();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "flag read-back");
    const DevFlags f = *h_flags;
    max_nn_seen = std::max(max_nn_seen, f.max_nn);
    if (out) *out = f;
    if (f.nonfinite) return fail(DEM_ERR_NONFINITE, "non-finite particle position met by the cell binning");
    if (f.cell_overflow)
      return fail(DEM_ERR_ARG, "more than 4096 particles in one cell: the binning box (lo/hi) "
                               "does not cover the particles");
    if (f.nbr_overflow)
      return fail(DEM_ERR_NEIGH_OVERFLOW, "neighbour row of " + std::to_string(f.nbr_overflow) +
                                              " exceeds max_neighbors = " + std::to_string(max_n));
    if (f.contact_overflow)
      return fail(DEM_ERR_CONTACT_OVERFLOW, "particle with " + std::to_string(f.contact_overflow) +
                                                " overlapping neighbours exceeds max_contacts = " +
                                                std::to_string(max_c));
    if (f.skin) return fail(DEM_ERR_SKIN, "a particle moved more than skin/2 between neighbour rebuilds");
    if (f.bond_overflow)
      return fail(DEM_ERR_BOND_OVERFLOW, "particle with " + std::to_string(f.bond_overflow) +
                                             " bond partners exceeds max_bonds = " + std::to_string(max_b));
    return DEM_OK;
  }
  // Reports and clears the sticky device error flags.
  int check_and_clear_flags(bool enqueued = false) {
    DevFlags f{};
    int r = check_flags(&f, enqueued);
    if (r != DEM_OK && r != DEM_ERR_HIP) {
      DevFlags z{};
      z.bonds_broken = f.bonds_broken; // a running count, not an error
      hipMemcpyAsync(flags, &z, sizeof(z), hipMemcpyHostToDevice, stream);
      hipStreamSynchronize(stream);
    }
    return r;
  }

  // Device rows for the sampled diagnostics of one call and their pinned
  // host copy.
  int ensure_diag_rows(int rows) {
    if (rows <= diag_rows_cap) return DEM_OK;
    hipFree(diag_rows);
    diag_rows = nullptr;
    if (h_diag_rows) hipHostFree(h_diag_rows);
    h_diag_rows = nullptr;
    diag_rows_cap = 0;
    hipError_t e = hipSuccess;
    diag_rows = dalloc<double>((size_t)rows * 9, &e);
    if (e == hipSuccess)
      e = hipHostMalloc(reinterpret_cast<void**>(&h_diag_rows), (size_t)rows * 9 * sizeof(double),
                        hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "diagnostic rows");
    diag_rows_cap = rows;
    return DEM_OK;
  }

  // Particle-particle contacts are found through the neighbour rows (every
  // plain system; a bonded one only with contacts between unbonded pairs).
  bool pp_live() const { return !bonds_on || bcfg.unbonded_contacts; }

  // The rows must hold every pair that can touch before the next rebuild:
  // cutoff >= 2 r_max, plus the skin when the rows are reused for several
  // steps (rebuild_every > 1 or automatic, which need a skin to be checked
  // at all).
  int check_list_cover() {
    if (!pp_live() || max_rad <= 0.) return DEM_OK;
    const bool reuse = cfg.rebuild_every != 1;
    if (reuse && !(cfg.skin > 0.))
      return fail(DEM_ERR_ARG, "rebuild_every > 1 needs skin > 0: without it no displacement check "
                               "guards the contacts that form between rebuilds");
    const double need = 2. * max_rad + (reuse ? cfg.skin : 0.);
    if (cfg.cutoff < need * (1. - 1e-12))
      return fail(DEM_ERR_ARG, "cutoff " + std::to_string(cfg.cutoff) + " is below 2 r_max" +
                               (reuse ? " + skin" : "") + " = " + std::to_string(need) +
                               ": contacts would be missed");
    return DEM_OK;
  }

  // The tile set of the force kernels: kForceBlocks[k] for systems without
  // bonds, kBlock (k = 1) with bonds.
  void select_block(int k) {
    bidx = k;
    const int e = bonds_on ? 1 : k;
    ntiles = ntiles_b[e];
    d_tiles = d_tiles_b[e];
    d_tile0 = d_tile0_b[e];
  }
  // Binning grids with `rows` y-rows per cell tile (the cell order of the
  // re-sort; any value gives exact rows).  Grows the cell arrays when the
  // padded cell count does; the rows are rebuilt at the next step.
  int set_tile_rows(int rows) {
    std::vector<Grid> g(nsys);
    std::vector<long> kb(nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
      long nc = 0;
      g[s] = make_grid(sys_cfg[s], &nc, rows);
      kb[s + 1] = kb[s] + nc;
    }
    if (kb[nsys] > (1l << 30)) return fail(DEM_ERR_ARG, "the binning grids need more than 2^30 cells in total");
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess && kb[nsys] > ncell_cap) {
      hipFree(cell_count); hipFree(cell_start); hipFree(scan_tmp);
      cell_count = cell_start = scan_tmp = nullptr;
      cell_count = dalloc<int>((size_t)kb[nsys], &e);
      cell_start = dalloc<int>((size_t)kb[nsys] + 1, &e);
      scan_tmp = dalloc<int>((size_t)(kb[nsys] / 4096 + 2), &e);
      if (e == hipSuccess) ncell_cap = kb[nsys];
    }
    if (e == hipSuccess && nsys > 1) {
      e = hipMemcpy(d_grids, g.data(), nsys * sizeof(Grid), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(d_kbase, kb.data(), nsys * sizeof(long), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return hip_fail(e, "binning grids");
    grids = g;
    kbase = kb;
    ncell = kb[nsys];
    tile_rows = rows;
    return DEM_OK;
  }

  // Automatic rebuild cadence.  Rows built with cutoff >= 2 r_max + skin
  // hold every pair that can touch as long as no particle has moved more
  // than skin/2 since the build.  A check measures the largest
  // displacement since the build, d, and the largest speed, v (inside a
  // step's force pass, read on the host once the next pass is queued), and
  // plans the next check M force passes ahead so
  // that d + 2 v dt (M + 1) stays within 0.9 skin/2: speeds may double in
  // the meantime, and a call boundary adds a drift.  M is at most the
  // number of force passes since the positions were set, so that under a
  // constant acceleration from rest v >= a t bounds the a (M dt)^2 / 2 that
  // the factor 2 leaves out; walls that move (corrected mode) count with
  // twice their surface speed, the most a bounce off them can give.  Right
  // after a build the first check is planned from the speed measured just
  // before it; the rows are rebuilt once not one more pass fits, or once
  // the plan has shrunk below a quarter of that first interval (each check
  // roughly halves what is left, so this caps the checks per build at about
  // three).  The rebuild's own skin check stays on behind this as the guard
  // (a violent impulse can still trip it).
  static constexpr int kAutoMaxInterval = 1000;
#ifndef DEM_AUTO_SHRINK
#define DEM_AUTO_SHRINK 4 // rebuild once the plan is below 1/DEM_AUTO_SHRINK of the first (0: never)
#endif
  double wall_speed() const {
    if (cfg.physics_mode != DEM_MODE_CORRECTED) return 0.; // walls never move in reference mode
    double vmax = 0.;
    for (int s = 0; s < nsys; ++s) {
      const Grid& g = grids[s];
      for (int j = 0; j < h_sys[s].n_walls; ++j) {
        const WallConst& W = h_sys[s].walls[j];
        double reach2 = 0.; // farthest corner of the binning box from x_w
        for (int d = 0; d < 3; ++d) {
          const double a = std::fabs(g.lo[d] - W.x[d]), b = std::fabs(g.lo[d] + g.nc[d] * g.h[d] - W.x[d]);
          reach2 += std::max(a, b) * std::max(a, b);
        }
        const double u = std::sqrt(W.u[0] * W.u[0] + W.u[1] * W.u[1] + W.u[2] * W.u[2]);
        const double w = std::sqrt(W.w[0] * W.w[0] + W.w[1] * W.w[1] + W.w[2] * W.w[2]);
        vmax = std::max(vmax, u + w * std::sqrt(reach2));
      }
    }
    return vmax;
  }
  int auto_interval(double budget) const {
    const double per_step = 2. * auto_v * cfg.dt;
    if (!(budget > 0.)) return 0;
    const int64_t age = step - auto_since;
    const int cap = (int)std::min<int64_t>(kAutoMaxInterval, std::max<int64_t>(1, age));
    if (!(per_step > 0.)) return cap;
    const double m = std::floor(budget / per_step) - 1.;
    return m < 1. ? 0 : (m > cap ? cap : (int)m);
  }
  // A check is read in two parts: the maxima that the step's force pass has
  // computed (StepArgs::disp) are copied to pinned memory behind it, and read
  // once the next force pass is queued too, so the GPU runs that pass while
  // the host waits and plans (no idle gap).  The plan then starts one pass
  // later: that pass's motion, at most 2 v dt, comes off the budget first.
  int auto_check_enqueue() {
    ++auto_checks;
    hipError_t e = hipSuccess;
    if (!h_disp) e = hipHostMalloc(reinterpret_cast<void**>(&h_disp), 2 * sizeof(unsigned long long),
                                   hipHostMallocDefault);
    if (e == hipSuccess && !ev_disp) e = hipEventCreateWithFlags(&ev_disp, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_disp, d_disp, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(ev_disp, stream);
    return e == hipSuccess ? DEM_OK : hip_fail(e, "displacement check");
  }
  // passes_after: force passes queued after the one that computed the maxima
  int auto_check_read(int passes_after) {
    hipError_t e = hipEventSynchronize(ev_disp);
    if (e != hipSuccess) return hip_fail(e, "displacement check");
    double d2, v2;
    std::memcpy(&d2, &h_disp[0], sizeof(double));
    std::memcpy(&v2, &h_disp[1], sizeof(double));
    auto_v = std::max(std::sqrt(v2), 2. * wall_speed()); // NaN stays NaN
    if (v2 != v2) auto_v = v2;
    const double gone = passes_after * 2. * auto_v * cfg.dt;
    const int m = auto_interval(0.45 * cfg.skin - std::sqrt(d2) - gone); // NaN: 0, rebuild
    // the quarter rule counts from the measurement, as the first interval does
    auto_due = m < 1 || (DEM_AUTO_SHRINK > 0 && DEM_AUTO_SHRINK * (m + passes_after) < auto_first);
    auto_left = m;
    return DEM_OK;
  }
  // after a build: the first check planned from the last measured speed
  // (one pass ahead without one)
  void auto_after_rebuild() {
    auto_due = false;
    auto_first = auto_v >= 0. ? std::max(1, auto_interval(0.45 * cfg.skin)) : 1;
    auto_left = auto_first;
  }
  // positions or velocities set from outside: measure afresh
  void auto_reset() {
    auto_since = step;
    auto_v = -1.;
    auto_first = 0;
    auto_left = 0;
    auto_due = false;
  }

  // Materials and walls of every system to the device (synchronous).
  int upload_sys() {
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
      e = hipMemcpy(d_sys, h_sys.data(), h_sys.size() * sizeof(SysParams), hipMemcpyHostToDevice);
    return e == hipSuccess ? DEM_OK : hip_fail(e, "upload system parameters");
  }

  // original id of each sorted slot (host copy of the current order)
  int sorted_ids(std::vector<int>* id) {
    id->resize(n);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(id->data(), this->id[cS], n * sizeof(int), hipMemcpyDeviceToHost);
    return e == hipSuccess ? DEM_OK : hip_fail(e, "id read-back");
  }

  // E, G of each sorted slot for the bond stiffnesses.
  int upload_bEG(const std::vector<int>& id) {
    std::vector<double2> eg(n);
    for (int p = 0; p < n; ++p) { eg[p].x = h_E[id[p]]; eg[p].y = h_G[id[p]]; }
    hipError_t e = hipMemcpy(bEG[cS], eg.data(), n * sizeof(double2), hipMemcpyHostToDevice);
    return e == hipSuccess ? DEM_OK : hip_fail(e, "upload bond moduli");
  }

  int alloc_nbr(int cap) {
    ++gen; // the rows move: captured launches are stale
    hipFree(nbr);
    nbr = nullptr;
    hipError_t e = hipSuccess;
    nbr = dalloc<int>((size_t)cap * ld, &e);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc neighbour rows");
    max_n = cap;
    return DEM_OK;
  }

  // Rebuild: re-sort into cell order, gather, search.  Grows the rows on
  // overflow when max_neighbors was left to auto sizing.
  // permute_ft: F and T are live across this rebuild (false inside dem_step,
  // where the fused kernel has already applied them).
  int rebuild(bool sync_check, bool check_skin, bool permute_ft = true) {
    mark_begin(1);
    BuildArgs b = build_args(check_skin && built_once, permute_ft);
    hipError_t e = launch_rebuild(b, stream);
    if (e != hipSuccess) return hip_fail(e, "rebuild launch");
    cP ^= 1; cS ^= 1; cC ^= 1;
    b = build_args(false);
    b.xr_out = xr[cP];
    if ((e = launch_neighbors(b, stream)) != hipSuccess) return hip_fail(e, "neighbour launch");
    mark_end();
    built_once = true;
    list_valid = true;
    ++rebuilds;
    if (!sync_check) return DEM_OK;
    DevFlags f;
    e = hipMemcpyAsync(&f, flags, sizeof(f), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "flag read-back");
    if (f.nbr_overflow && cfg.max_neighbors <= 0) {
      int cap = f.nbr_overflow + f.nbr_overflow / 2 + 8;
      int rc = alloc_nbr(cap);
      if (rc) return rc;
      DevFlags z{};
      z.max_nn = f.max_nn;
      z.bonds_broken = f.bonds_broken;
      hipMemcpyAsync(flags, &z, sizeof(z), hipMemcpyHostToDevice, stream);
      b.max_n = max_n;
      b.nbr = nbr;
      if ((e = launch_neighbors(b, stream)) != hipSuccess) return hip_fail(e, "neighbour launch");
    }
    return check_and_clear_flags();
  }
};

namespace {

int validate(const dem_config* c, std::string* msg) {
  if (!c) { *msg = "null config"; return DEM_ERR_ARG; }
  if (c->n <= 0) { *msg = "n must be positive"; return DEM_ERR_ARG; }
  if (c->n_walls < 0 || c->n_walls > DEM_MAX_WALLS) { *msg = "n_walls must be in [0, 8]"; return DEM_ERR_ARG; }
  if (c->max_contacts <= 0 || c->max_contacts > kMaxContacts) { *msg = "max_contacts must be in [1, 32]"; return DEM_ERR_ARG; }
  if (!(c->dt > 0.)) { *msg = "dt must be positive"; return DEM_ERR_ARG; }
  if (!(c->cutoff > 0.)) { *msg = "cutoff must be positive"; return DEM_ERR_ARG; }
  if (!(c->cor_pp > 0. && c->cor_pp <= 1.) || !(c->cor_pw > 0. && c->cor_pw <= 1.)) {
    *msg = "restitution coefficients must be in (0, 1] (log(cor) at src/CabanaDEM_Force.hpp:206)";
    return DEM_ERR_ARG;
  }
  for (int d = 0; d < 3; ++d)
    if (!(c->hi[d] > c->lo[d])) { *msg = "binning box must have hi > lo"; return DEM_ERR_ARG; }
  if (c->physics_mode != DEM_MODE_REFERENCE && c->physics_mode != DEM_MODE_CORRECTED) {
    *msg = "physics_mode must be 0 or 1"; return DEM_ERR_ARG;
  }
  if (c->rebuild_every == DEM_REBUILD_AUTO && !(c->skin > 0.)) {
    *msg = "rebuild_every = DEM_REBUILD_AUTO needs skin > 0 (the displacement budget)";
    return DEM_ERR_ARG;
  }
#define DEM_STR2(x) #x
#define DEM_STR(x) DEM_STR2(x)
  if (c->block_size != 0 && c->block_size != 64 && c->block_size != 128 && c->block_size != 256) {
    *msg = "block_size must be 0 (" DEM_STR(DEM_BLOCK) "), 64, 128 or 256";
    return DEM_ERR_ARG;
  }
  return DEM_OK;
}

// Cells of edge >= cutoff (1e-9 margin against rounding in (x-lo)/h),
// capped at 64M cells by coarsening (coarser cells stay exact); cells
// numbered in y-tiles of `rows` rows.
Grid make_grid(const dem_config& c, long* ncell, int rows) {
  Grid g{};
  long cap = 64l << 20;
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    double ext = c.hi[d] - c.lo[d];
    double cs = c.cutoff * (1.0 + 1e-9);
    double f = std::floor(ext / cs);
    if (f < 1.) f = 1.;
    if (f > 1e6) f = 1e6;
    nc[d] = (int)f;
  }
  auto padded = [&]() { return (long)nc[0] * ((nc[1] + rows - 1) / rows) * rows * nc[2]; };
  while (padded() > cap) {
    int d = (nc[0] >= nc[1] && nc[0] >= nc[2]) ? 0 : (nc[1] >= nc[2] ? 1 : 2);
    nc[d] = std::max(1, nc[d] / 2);
  }
  for (int d = 0; d < 3; ++d) {
    double ext = c.hi[d] - c.lo[d];
    g.lo[d] = c.lo[d];
    g.nc[d] = nc[d];
    g.h[d] = ext / nc[d];
    double cs = c.cutoff * (1.0 + 1e-9);
    if (g.h[d] < cs) g.h[d] = cs;
  }
  g.tr = rows;
  g.nty = (nc[1] + rows - 1) / rows;
  *ncell = cell_key_count(g);
  return g;
}

int check_ctx(dem_ctx* c) {
  if (!c) { set_global_error("null context"); return DEM_ERR_ARG; }
  return DEM_OK;
}

} // namespace

extern "C" {

int dem_abi_version(void) { return DEM_ABI_VERSION; }

int dem_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* dem_last_error(const dem_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  static thread_local std::string copy;
  copy = g_err;
  return copy.c_str();
}

}  // extern "C" (reopened below)

namespace {

// Materials of one system into its parameter block.
void set_materials(SysParams& P, double cor_pp, double cor_pw, double fr_pp, double fr_pw) {
  P.c_eta_pp = eta_prefactor(cor_pp);
  P.c_eta_pw = eta_prefactor(cor_pw);
  P.fr_pp = fr_pp;
  P.fr_pw = fr_pw;
}

// Walls of one system: src/CabanaDEM_Force.hpp:562 and :613 per-wall factors.
int set_walls(SysParams& P, const dem_walls_view* w, std::string* msg) {
  if (!w || w->n < 0 || w->n > DEM_MAX_WALLS) { *msg = "wall count must be in [0, 8]"; return DEM_ERR_ARG; }
  if (w->n > 0 && (!w->E || !w->nu || !w->G)) { *msg = "walls need E, nu and G"; return DEM_ERR_ARG; }
  P.n_walls = w->n;
  for (int j = 0; j < w->n; ++j) {
    WallConst& W = P.walls[j];
    for (int d = 0; d < 3; ++d) {
      W.x[d] = w->x ? w->x[3 * j + d] : 0.;
      W.n[d] = w->normal ? w->normal[3 * j + d] : 0.;
    }
    W.a1 = (1. - (w->nu[j] * w->nu[j])) / w->E[j];
    W.a2 = (2. - w->nu[j]) / w->G[j];
    for (int d = 0; d < 3; ++d) {
      W.u[d] = w->u ? w->u[3 * j + d] : 0.;
      W.w[d] = w->omega ? w->omega[3 * j + d] : 0.;
    }
  }
  return DEM_OK;
}

bool walls_translate(const SysParams& P) {
  for (int j = 0; j < P.n_walls; ++j)
    if (P.walls[j].u[0] != 0. || P.walls[j].u[1] != 0. || P.walls[j].u[2] != 0.) return true;
  return false;
}

// One system's configuration: the common fields plus its own count,
// materials and binning box.
dem_config system_config(const dem_config& common, const dem_system_desc& d) {
  dem_config c = common;
  c.n = d.n;
  c.n_walls = d.n_walls;
  c.cor_pp = d.cor_pp; c.cor_pw = d.cor_pw;
  c.friction_pp = d.friction_pp; c.friction_pw = d.friction_pw;
  for (int k = 0; k < 3; ++k) { c.lo[k] = d.lo[k]; c.hi[k] = d.hi[k]; }
  return c;
}

int create_ctx(const dem_config* common, int nsys, const dem_system_desc* sd, dem_ctx** out) {
  if (!out) { set_global_error("null output pointer"); return DEM_ERR_ARG; }
  *out = nullptr;
  if (!common || nsys < 1 || !sd) { set_global_error("need a configuration and at least one system"); return DEM_ERR_ARG; }
  std::string msg;
  long total = 0;
  for (int s = 0; s < nsys; ++s) {
    dem_config cs = system_config(*common, sd[s]);
    int rc = validate(&cs, &msg);
    if (rc) { set_global_error(nsys > 1 ? "system " + std::to_string(s) + ": " + msg : msg); return rc; }
    total += sd[s].n;
  }
  if (total > (1l << 30)) { set_global_error("more than 2^30 particles in one context"); return DEM_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_global_error("no HIP device available");
    return DEM_ERR_HIP;
  }
  dem_ctx* c = new dem_ctx();
  c->cfg = system_config(*common, sd[0]);
  c->cfg.n = (int)total;
  if (c->cfg.rebuild_every <= 0 && c->cfg.rebuild_every != DEM_REBUILD_AUTO) c->cfg.rebuild_every = 1;
  c->n = (int)total;
  c->ld = (c->n + 63) / 64 * 64;
  c->max_c = common->max_contacts;
  c->rc2 = common->cutoff * common->cutoff; // Cabana: rsqr = delta * delta
  // systems: slot ranges, tiles, grids and key ranges, materials
  c->nsys = nsys;
  c->sys_off.assign(nsys + 1, 0);
  c->sys_cfg.resize(nsys);
  c->grids.resize(nsys);
  c->kbase.assign(nsys + 1, 0);
  c->h_sys.assign(nsys, SysParams{});
  std::vector<int4> tiles[3];
  std::vector<int> sysid;
  for (int b = 0; b < 3; ++b) c->tile0_b[b].assign(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    dem_config cs = system_config(*common, sd[s]);
    c->sys_cfg[s] = cs;
    long nc = 0;
    c->grids[s] = make_grid(cs, &nc, c->tile_rows);
    c->kbase[s + 1] = c->kbase[s] + nc;
    c->sys_off[s + 1] = c->sys_off[s] + sd[s].n;
    for (int b = 0; b < 3; ++b) {
      const int B = kForceBlocks[b];
      const int nt = (sd[s].n + B - 1) / B;
      c->tile0_b[b][s + 1] = c->tile0_b[b][s] + nt;
      for (int t = 0; t < nt; ++t)
        tiles[b].push_back(make_int4(s, c->sys_off[s] + t * B, c->sys_off[s + 1], 0));
    }
    if (nsys > 1) sysid.insert(sysid.end(), (size_t)sd[s].n, s);
    set_materials(c->h_sys[s], sd[s].cor_pp, sd[s].cor_pw, sd[s].friction_pp, sd[s].friction_pw);
    c->h_sys[s].n_walls = 0;
  }
  c->ncell = c->kbase[nsys];
  c->ncell_cap = c->ncell;
  for (int b = 0; b < 3; ++b) c->ntiles_b[b] = c->tile0_b[b][nsys];
  if (c->ncell > (1l << 30)) {
    delete c;
    set_global_error("the binning grids need more than 2^30 cells in total");
    return DEM_ERR_ARG;
  }
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) { int r = c->hip_fail(e, "hipStreamCreate"); delete c; return r; }
  {
    int dev = 0, cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
      c->cus = cu;
  }
  const size_t n = (size_t)c->n, ld = (size_t)c->ld;
  e = hipSuccess;
  for (int b = 0; b < 2; ++b) {
    c->xr[b] = dalloc<double4>(n, &e);
    c->uw[b] = dalloc<double>(6 * n, &e);
    c->mat[b] = dalloc<double4>(n, &e);
    c->mI[b] = dalloc<double2>(n, &e);
    c->id[b] = dalloc<int>(n, &e);
    c->ft[b] = dalloc<double>(6 * n, &e);
    c->sw[b] = dalloc<double>((size_t)DEM_MAX_WALLS * 3 * ld, &e);
    c->cmeta[b] = dalloc<int>(n, &e);
    c->cj[b] = dalloc<int>((size_t)c->max_c * ld, &e);
    c->ch[b] = dalloc<double>((size_t)c->max_c * 3 * ld, &e);
  }
  c->nn = dalloc<int>(n, &e);
  c->xb = dalloc<double>(3 * n, &e);
  c->key = dalloc<int>(n, &e);
  c->rank = dalloc<int>(n, &e);
  c->key_sorted = dalloc<int>(n, &e);
  c->cell_count = dalloc<int>((size_t)c->ncell, &e);
  c->cell_start = dalloc<int>((size_t)c->ncell + 1, &e);
  c->perm = dalloc<int>(n, &e);
  c->inv = dalloc<int>(n, &e);
  c->scan_tmp = dalloc<int>((size_t)(c->ncell / 4096 + 2), &e);
  c->flags = dalloc<DevFlags>(1, &e);
  c->d_disp = dalloc<unsigned long long>(2, &e);
  {
    // work lists: every tile's slots at every tile size
    size_t slots = 0;
    for (int b = 0; b < 3; ++b) slots = std::max(slots, (size_t)c->ntiles_b[b] * kForceBlocks[b]);
    c->d_work = dalloc<int>(slots, &e);
    c->d_work_tile = dalloc<int>((size_t)c->ntiles_b[0], &e); // the most tiles (64 slots)
  }
  // the most tiles (64 slots), then the reduction's group partials
  c->diag_part = dalloc<double>(((size_t)c->ntiles_b[0] + kDiagThreads) * 8, &e);
  c->d_sys = dalloc<SysParams>((size_t)nsys, &e);
  for (int b = 0; b < 3; ++b) c->d_tile0_b[b] = dalloc<int>((size_t)nsys + 1, &e);
  c->d_sysdiag = dalloc<double>((size_t)nsys * 8, &e);
  if (nsys > 1) {
    for (int b = 0; b < 3; ++b) c->d_tiles_b[b] = dalloc<int4>(tiles[b].size(), &e);
    c->d_sysid = dalloc<int>(n, &e);
    c->d_grids = dalloc<Grid>((size_t)nsys, &e);
    c->d_kbase = dalloc<long>((size_t)nsys, &e);
  }
  if (e != hipSuccess) { int r = c->hip_fail(e, "hipMalloc"); delete c; return r; }
  int rc;
  int cap = common->max_neighbors > 0 ? common->max_neighbors : 48;
  if ((rc = c->alloc_nbr(cap))) { delete c; return rc; }
  // zero state: identity order, no contacts, no wall history
  std::vector<int> ids(n);
  for (size_t i = 0; i < n; ++i) ids[i] = (int)i;
  e = hipMemcpy(c->id[0], ids.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->cmeta[0], 0, n * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->ft[0], 0, 6 * n * sizeof(double));
  if (e == hipSuccess) e = hipMemset(c->uw[0], 0, 6 * n * sizeof(double));
  if (e == hipSuccess) e = hipMemset(c->xr[0], 0, n * sizeof(double4));
  if (e == hipSuccess) e = hipMemset(c->nn, 0, n * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->flags, 0, sizeof(DevFlags));
  for (int b = 0; b < 3; ++b) {
    if (e == hipSuccess)
      e = hipMemcpy(c->d_tile0_b[b], c->tile0_b[b].data(), (nsys + 1) * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && nsys > 1)
      e = hipMemcpy(c->d_tiles_b[b], tiles[b].data(), tiles[b].size() * sizeof(int4), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && nsys > 1) {
    e = hipMemcpy(c->d_sysid, sysid.data(), n * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_grids, c->grids.data(), nsys * sizeof(Grid), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_kbase, c->kbase.data(), nsys * sizeof(long), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) { int r = c->hip_fail(e, "initial fill"); delete c; return r; }
  c->select_block(common->block_size == 64 ? 0 : (common->block_size == 256 ? 2 : 1));
  if ((rc = c->upload_sys())) { delete c; return rc; }
  c->h_m.assign(n, 1.); c->h_rad.assign(n, 1.); c->h_E.assign(n, 1.);
  c->h_nu.assign(n, 0.); c->h_G.assign(n, 1.); c->h_I.assign(n, 1.);
  *out = c;
  return DEM_OK;
}

} // namespace

extern "C" {

int dem_create(const dem_config* cfg, dem_ctx** out) {
  if (!cfg) { set_global_error("null config"); if (out) *out = nullptr; return DEM_ERR_ARG; }
  dem_system_desc d{};
  d.n = cfg->n;
  d.n_walls = cfg->n_walls;
  d.cor_pp = cfg->cor_pp; d.cor_pw = cfg->cor_pw;
  d.friction_pp = cfg->friction_pp; d.friction_pw = cfg->friction_pw;
  for (int k = 0; k < 3; ++k) { d.lo[k] = cfg->lo[k]; d.hi[k] = cfg->hi[k]; }
  return create_ctx(cfg, 1, &d, out);
}

int dem_create_ensemble(const dem_config* common, int n_systems, const dem_system_desc* systems,
                        dem_ctx** out) {
  return create_ctx(common, n_systems, systems, out);
}

void dem_destroy(dem_ctx* ctx) {
  if (!ctx) return;
  hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int dem_set_params(dem_ctx* c, const dem_config* cfg) {
  if (int r = check_ctx(c)) return r;
  dem_config want = *cfg;
  if (c->nsys > 1) {
    // ensembles: only the common fields change; count, box and materials
    // stay per system (dem_set_system_params)
    want.n = c->cfg.n;
    for (int d = 0; d < 3; ++d) { want.lo[d] = c->cfg.lo[d]; want.hi[d] = c->cfg.hi[d]; }
    want.cor_pp = c->cfg.cor_pp; want.cor_pw = c->cfg.cor_pw;
    want.friction_pp = c->cfg.friction_pp; want.friction_pw = c->cfg.friction_pw;
  }
  std::string msg;
  if (int r = validate(&want, &msg)) return c->fail(r, msg);
  const dem_config& o = c->cfg;
  bool same = want.n == o.n && want.max_contacts == o.max_contacts && want.cutoff == o.cutoff;
  for (int d = 0; d < 3; ++d) same = same && want.lo[d] == o.lo[d] && want.hi[d] == o.hi[d];
  if (!same) return c->fail(DEM_ERR_ARG, "dem_set_params cannot change n, max_contacts, cutoff or the box");
  hipStreamSynchronize(c->stream);
  dem_config keep = c->cfg;
  // positions at the last build (xb) are kept only while skin > 0: after a
  // change the next rebuild runs unchecked and records them afresh
  if (want.skin != keep.skin) c->list_valid = false;
  c->cfg = want;
  c->cfg.max_neighbors = keep.max_neighbors;
  c->cfg.n_walls = keep.n_walls;
  if (c->cfg.rebuild_every <= 0 && c->cfg.rebuild_every != DEM_REBUILD_AUTO) c->cfg.rebuild_every = 1;
  if (c->cfg.rebuild_every == DEM_REBUILD_AUTO && !(c->cfg.skin > 0.)) {
    c->cfg = keep;
    return c->fail(DEM_ERR_ARG, "rebuild_every = DEM_REBUILD_AUTO needs skin > 0");
  }
  ++c->gen;
  if (c->nsys == 1) {
    set_materials(c->h_sys[0], want.cor_pp, want.cor_pw, want.friction_pp, want.friction_pw);
    return c->upload_sys();
  }
  return DEM_OK;
}

int dem_set_system_params(dem_ctx* c, int system, double cor_pp, double cor_pw,
                          double friction_pp, double friction_pw) {
  if (int r = check_ctx(c)) return r;
  if (system < 0 || system >= c->nsys) return c->fail(DEM_ERR_ARG, "system index out of range");
  if (!(cor_pp > 0. && cor_pp <= 1.) || !(cor_pw > 0. && cor_pw <= 1.))
    return c->fail(DEM_ERR_ARG, "restitution coefficients must be in (0, 1]");
  set_materials(c->h_sys[system], cor_pp, cor_pw, friction_pp, friction_pw);
  if (c->nsys == 1) {
    c->cfg.cor_pp = cor_pp; c->cfg.cor_pw = cor_pw;
    c->cfg.friction_pp = friction_pp; c->cfg.friction_pw = friction_pw;
  }
  return c->upload_sys();
}

int dem_upload_walls(dem_ctx* c, const dem_walls_view* w) {
  if (int r = check_ctx(c)) return r;
  if (c->nsys > 1) return c->fail(DEM_ERR_STATE, "an ensemble takes walls per system (dem_upload_system_walls)");
  return dem_upload_system_walls(c, 0, w);
}

int dem_upload_system_walls(dem_ctx* c, int system, const dem_walls_view* w) {
  if (int r = check_ctx(c)) return r;
  if (system < 0 || system >= c->nsys) return c->fail(DEM_ERR_ARG, "system index out of range");
  std::string msg;
  if (int r = set_walls(c->h_sys[system], w, &msg)) return c->fail(r, msg);
  if (c->nsys == 1) c->cfg.n_walls = w->n;
  // the uploaded positions hold now: translating walls move from here on
  c->wall_step0 = c->step;
  c->walls_move = false;
  for (int s = 0; s < c->nsys; ++s) c->walls_move = c->walls_move || walls_translate(c->h_sys[s]);
  ++c->gen;
  return c->upload_sys();
}

int dem_get_systems(dem_ctx* c, int* n_systems, int* offsets) {
  if (int r = check_ctx(c)) return r;
  if (n_systems) *n_systems = c->nsys;
  if (offsets) std::copy(c->sys_off.begin(), c->sys_off.end(), offsets);
  return DEM_OK;
}

int dem_system_diagnostics(dem_ctx* c, double* out) {
  if (int r = check_ctx(c)) return r;
  if (!out) return c->fail(DEM_ERR_ARG, "null output");
  StepArgs a = c->step_args();
  hipError_t e = launch_diag_current(a, c->uw[c->cP], c->stream);
  if (e == hipSuccess) e = launch_diag_systems(c->diag_part, c->d_tile0, c->nsys, c->d_sysdiag, c->stream);
  std::vector<double> h((size_t)c->nsys * 8);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), c->d_sysdiag, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "system diagnostics");
  for (int s = 0; s < c->nsys; ++s)
    for (int k = 0; k < 5; ++k) out[5 * s + k] = h[8 * (size_t)s + k];
  return DEM_OK;
}

int dem_set_launch(dem_ctx* c, int grid_blocks) {
  if (int r = check_ctx(c)) return r;
  if (grid_blocks < 0) return c->fail(DEM_ERR_ARG, "grid_blocks must be >= 0");
  c->grid_blocks = grid_blocks;
  ++c->gen;
  return DEM_OK;
}

int dem_set_force_pass(dem_ctx* c, int mode) {
  if (int r = check_ctx(c)) return r;
  if (mode != DEM_FORCE_PASS_FUSED && mode != DEM_FORCE_PASS_SPLIT && mode != DEM_FORCE_PASS_WAVE)
    return c->fail(DEM_ERR_ARG, "force pass must be DEM_FORCE_PASS_FUSED, _SPLIT or _WAVE");
  c->force_pass = mode;
  ++c->gen;
  return DEM_OK;
}

int dem_set_shape(dem_ctx* c, int block, int tile_rows) {
  if (int r = check_ctx(c)) return r;
  if (block != 0 && block != 64 && block != 128 && block != 256)
    return c->fail(DEM_ERR_ARG, "block must be 0 (128), 64, 128 or 256");
  if (tile_rows < 0 || tile_rows > kMaxTileRows)
    return c->fail(DEM_ERR_ARG, "tile_rows must be 0 (8) or 1 to 32");
  if (block == 0) block = kBlock;
  if (tile_rows == 0) tile_rows = kTileRows;
  if (c->bonds_on && block != kBlock)
    return c->fail(DEM_ERR_ARG, "bonded systems run force tiles of 128");
  const int k = block == 64 ? 0 : (block == 256 ? 2 : 1);
  if (k != c->bidx) {
    c->select_block(k);
    ++c->gen;
  }
  if (tile_rows != c->tile_rows) {
    if (int r = c->set_tile_rows(tile_rows)) return r;
    c->list_valid = false; // the next step re-sorts into the new cell order
    ++c->gen;
  }
  return DEM_OK;
}

int dem_get_shape(dem_ctx* c, int* block, int* tile_rows) {
  if (int r = check_ctx(c)) return r;
  if (block) *block = kForceBlocks[c->bonds_on ? 1 : c->bidx];
  if (tile_rows) *tile_rows = c->tile_rows;
  return DEM_OK;
}

int dem_upload_particles(dem_ctx* c, const dem_particles_view* v) {
  if (int r = check_ctx(c)) return r;
  if (!v || v->n != c->n) return c->fail(DEM_ERR_ARG, "particle count differs from dem_config.n");
  const int n = c->n;
  const size_t ld = c->ld;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "sync");
  // current device state, sorted order
  std::vector<int> id(n), meta(n);
  std::vector<double4> xr(n);
  std::vector<double> uw(6 * (size_t)n), ft(6 * (size_t)n);
  e = hipMemcpy(id.data(), c->id[c->cS], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(meta.data(), c->cmeta[c->cC], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(xr.data(), c->xr[c->cP], n * sizeof(double4), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(uw.data(), c->uw[c->cP], 6 * n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ft.data(), c->ft[c->cS], 6 * n * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return c->hip_fail(e, "upload read-back");
  std::vector<int> pos(n);
  for (int p = 0; p < n; ++p) pos[id[p]] = p;
  // static properties (original order) -> derived factors
  auto setv = [&](std::vector<double>& dst, const double* src) {
    if (src) std::copy(src, src + n, dst.begin());
  };
  setv(c->h_m, v->m); setv(c->h_rad, v->rad); setv(c->h_E, v->E);
  setv(c->h_nu, v->nu); setv(c->h_G, v->G); setv(c->h_I, v->I);
  if (v->rad) c->max_rad = *std::max_element(c->h_rad.begin(), c->h_rad.end());
  std::vector<double4> mat(n);
  std::vector<double2> mI(n);
  for (int p = 0; p < n; ++p) {
    const int I = id[p];
    const double nu = c->h_nu[I];
    // the per-particle factors of src/CabanaDEM_Force.hpp:196-220
    mat[p].x = (1. - (nu * nu)) / c->h_E[I];
    mat[p].y = (2. - nu) / c->h_G[I];
    mat[p].z = 1. / c->h_rad[I];
    mat[p].w = 1. / c->h_m[I];
    mI[p].x = c->h_m[I];
    mI[p].y = c->h_I[I];
    xr[p].w = c->h_rad[I];
    for (int d = 0; d < 3; ++d) {
      if (v->x) (&xr[p].x)[d] = v->x[3 * I + d];
      if (v->u) uw[6 * p + d] = v->u[3 * I + d];
      if (v->omega) uw[6 * p + 3 + d] = v->omega[3 * I + d];
      if (v->force) ft[6 * p + d] = v->force[3 * I + d];
      if (v->torque) ft[6 * p + 3 + d] = v->torque[3 * I + d];
    }
  }
  e = hipMemcpy(c->xr[c->cP], xr.data(), n * sizeof(double4), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->uw[c->cP], uw.data(), 6 * n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->ft[c->cS], ft.data(), 6 * n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->mat[c->cS], mat.data(), n * sizeof(double4), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->mI[c->cS], mI.data(), n * sizeof(double2), hipMemcpyHostToDevice);
  if (e != hipSuccess) return c->hip_fail(e, "upload");
  // wall history: a set bit marks a (possibly) non-zero history
  if (v->sw_x && v->sw_y && v->sw_z) {
    const int MW = v->max_walls;
    std::vector<double> sw((size_t)DEM_MAX_WALLS * 3 * ld, 0.);
    for (int p = 0; p < n; ++p) {
      const int I = id[p];
      int wm = 0;
      for (int w = 0; w < std::min(MW, DEM_MAX_WALLS); ++w) {
        double h[3] = {v->sw_x[(size_t)I * MW + w], v->sw_y[(size_t)I * MW + w], v->sw_z[(size_t)I * MW + w]};
        if (h[0] != 0. || h[1] != 0. || h[2] != 0.) {
          wm |= 1 << w;
          for (int q = 0; q < 3; ++q) sw[(size_t)(w * 3 + q) * ld + p] = h[q];
        }
      }
      meta[p] = make_meta(meta_count(meta[p]), wm);
    }
    e = hipMemcpy(c->sw[c->cS], sw.data(), sw.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return c->hip_fail(e, "upload wall history");
  }
  // tracked pp contacts: reference rows (ids, history) -> compact rows in
  // sorted index space, ascending j; for a repeated id the last slot wins,
  // as the reference's search does (src/CabanaDEM_Force.hpp:168-174).
  if (v->ss_idx && v->ss_cnt) {
    const int M = v->max_cnts;
    std::vector<int> cjh((size_t)c->max_c * ld, 0);
    std::vector<double> chh((size_t)c->max_c * 3 * ld, 0.);
    for (int p = 0; p < n; ++p) {
      const int I = id[p];
      std::vector<std::pair<int, int>> ent; // (sorted j, slot k)
      for (int k = 0; k < v->ss_cnt[I] && k < M; ++k) {
        int J = v->ss_idx[(size_t)I * M + k];
        if (J < 0 || J >= n) continue;
        int jp = pos[J];
        bool dup = false;
        for (auto& q : ent) if (q.first == jp) { q.second = k; dup = true; }
        if (!dup) ent.push_back({jp, k});
      }
      std::sort(ent.begin(), ent.end());
      if ((int)ent.size() > c->max_c)
        return c->fail(DEM_ERR_CONTACT_OVERFLOW, "uploaded tracked contacts exceed max_contacts");
      meta[p] = make_meta((int)ent.size(), meta_walls(meta[p]));
      for (size_t m = 0; m < ent.size(); ++m) {
        cjh[m * ld + p] = ent[m].first;
        const size_t src = (size_t)I * M + ent[m].second;
        const double h[3] = {v->ss_x ? v->ss_x[src] : 0., v->ss_y ? v->ss_y[src] : 0.,
                             v->ss_z ? v->ss_z[src] : 0.};
        for (int q = 0; q < 3; ++q) chh[(m * 3 + q) * ld + p] = h[q];
      }
    }
    e = hipMemcpy(c->cj[c->cC], cjh.data(), cjh.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->ch[c->cC], chh.data(), chh.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return c->hip_fail(e, "upload contacts");
  }
  e = hipMemcpy(c->cmeta[c->cC], meta.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return c->hip_fail(e, "upload contact counts");
  if (c->bonds_on && (v->E || v->G))
    if (int r = c->upload_bEG(id)) return r;
  if (v->x) c->list_valid = false;
  if (v->x || v->u) c->auto_reset();
  return DEM_OK;
}

int dem_download_particles(dem_ctx* c, dem_particles_view* v) {
  if (int r = check_ctx(c)) return r;
  if (!v || v->n != c->n) return c->fail(DEM_ERR_ARG, "particle count differs from dem_config.n");
  const int n = c->n;
  const size_t ld = c->ld;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "sync");
  std::vector<int> id(n), meta(n);
  std::vector<double4> xr(n);
  std::vector<double> uw(6 * (size_t)n), ft(6 * (size_t)n);
  e = hipMemcpy(id.data(), c->id[c->cS], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(meta.data(), c->cmeta[c->cC], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(xr.data(), c->xr[c->cP], n * sizeof(double4), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(uw.data(), c->uw[c->cP], 6 * n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ft.data(), c->ft[c->cS], 6 * n * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return c->hip_fail(e, "download");
  for (int p = 0; p < n; ++p) {
    const int I = id[p];
    for (int d = 0; d < 3; ++d) {
      if (v->x) v->x[3 * I + d] = (&xr[p].x)[d];
      if (v->u) v->u[3 * I + d] = uw[6 * p + d];
      if (v->omega) v->omega[3 * I + d] = uw[6 * p + 3 + d];
      if (v->force) v->force[3 * I + d] = ft[6 * p + d];
      if (v->torque) v->torque[3 * I + d] = ft[6 * p + 3 + d];
    }
  }
  auto getv = [&](double* dst, const std::vector<double>& src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  getv(v->m, c->h_m); getv(v->rad, c->h_rad); getv(v->E, c->h_E);
  getv(v->nu, c->h_nu); getv(v->G, c->h_G); getv(v->I, c->h_I);
  if (v->sw_x || v->sw_y || v->sw_z) {
    const int MW = v->max_walls;
    std::vector<double> sw((size_t)DEM_MAX_WALLS * 3 * ld);
    e = hipMemcpy(sw.data(), c->sw[c->cS], sw.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "download wall history");
    for (int p = 0; p < n; ++p) {
      const int I = id[p];
      const int wm = meta_walls(meta[p]);
      for (int w = 0; w < MW; ++w) {
        double h[3] = {0., 0., 0.};
        if (w < DEM_MAX_WALLS && (wm & (1 << w)))
          for (int q = 0; q < 3; ++q) h[q] = sw[(size_t)(w * 3 + q) * ld + p];
        if (v->sw_x) v->sw_x[(size_t)I * MW + w] = h[0];
        if (v->sw_y) v->sw_y[(size_t)I * MW + w] = h[1];
        if (v->sw_z) v->sw_z[(size_t)I * MW + w] = h[2];
      }
    }
  }
  if (v->ss_idx || v->ss_cnt || v->ss_x) {
    const int M = v->max_cnts;
    std::vector<int> cjh((size_t)c->max_c * ld);
    std::vector<double> chh((size_t)c->max_c * 3 * ld);
    e = hipMemcpy(cjh.data(), c->cj[c->cC], cjh.size() * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(chh.data(), c->ch[c->cC], chh.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "download contacts");
    for (int p = 0; p < n; ++p) {
      const int I = id[p];
      const int cnt = meta_count(meta[p]);
      int k = 0;
      for (; k < cnt && k < M; ++k) {
        const size_t dst = (size_t)I * M + k;
        if (v->ss_idx) v->ss_idx[dst] = id[cjh[(size_t)k * ld + p]];
        if (v->ss_x) v->ss_x[dst] = chh[(size_t)(k * 3 + 0) * ld + p];
        if (v->ss_y) v->ss_y[dst] = chh[(size_t)(k * 3 + 1) * ld + p];
        if (v->ss_z) v->ss_z[dst] = chh[(size_t)(k * 3 + 2) * ld + p];
      }
      for (; k < M; ++k) {
        const size_t dst = (size_t)I * M + k;
        if (v->ss_idx) v->ss_idx[dst] = -1;
        if (v->ss_x) v->ss_x[dst] = 0.;
        if (v->ss_y) v->ss_y[dst] = 0.;
        if (v->ss_z) v->ss_z[dst] = 0.;
      }
      if (v->ss_cnt) v->ss_cnt[I] = std::min(cnt, M);
    }
  }
  return DEM_OK;
}

int dem_build_neighbors(dem_ctx* c) {
  if (int r = check_ctx(c)) return r;
  return c->rebuild(true, c->list_valid);
}

int dem_get_neighbors(dem_ctx* c, int* counts, int* ids, int max_n) {
  if (int r = check_ctx(c)) return r;
  if (!c->list_valid) return c->fail(DEM_ERR_STATE, "no neighbour list has been built");
  if (!counts || !ids || max_n <= 0) return c->fail(DEM_ERR_ARG, "bad output arrays");
  const int n = c->n;
  const size_t ld = c->ld;
  hipError_t e = hipStreamSynchronize(c->stream);
  std::vector<int> id(n), nn(n), nbr((size_t)c->max_n * ld);
  if (e == hipSuccess) e = hipMemcpy(id.data(), c->id[c->cS], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nn.data(), c->nn, n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nbr.data(), c->nbr, nbr.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return c->hip_fail(e, "download neighbours");
  int over = 0;
  std::vector<int> row;
  for (int p = 0; p < n; ++p) {
    row.clear();
    for (int k = 0; k < nn[p]; ++k) row.push_back(id[nbr[(size_t)k * ld + p]]);
    std::sort(row.begin(), row.end());
    const int I = id[p];
    counts[I] = (int)row.size();
    if ((int)row.size() > max_n) over = 1;
    for (int k = 0; k < (int)row.size() && k < max_n; ++k) ids[(size_t)I * max_n + k] = row[k];
  }
  if (over) return c->fail(DEM_ERR_ARG, "max_n too small for the neighbour rows");
  return DEM_OK;
}

int dem_set_kinematics_device(dem_ctx* c, const double* x, const double* u, int clear_history) {
  if (int r = check_ctx(c)) return r;
  if (c->bonds_on && clear_history)
    return c->fail(DEM_ERR_STATE, "clearing histories of a bonded context is not supported");
  hipError_t e = launch_set_kinematics(c->n, c->id[c->cS], x, u, clear_history, c->xr[c->cP], c->uw[c->cP],
                                       c->ft[c->cS], c->cmeta[c->cC], c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "set kinematics");
  if (x) c->list_valid = false;
  if (x || u) c->auto_reset();
  return DEM_OK;
}

int dem_get_kinematics_device(dem_ctx* c, double* x, double* u) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = launch_get_kinematics(c->n, c->id[c->cS], c->xr[c->cP], c->uw[c->cP], x, u, c->stream);
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "get kinematics");
}

int dem_zero_forces(dem_ctx* c) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = launch_zero_forces(c->step_args(), c->stream);
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "zero forces");
}

int dem_update_tangential(dem_ctx* c) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = launch_update_tangential(c->step_args(), c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "update tangential");
  c->cC ^= 1;
  return DEM_OK;
}

static int force_call(dem_ctx* c, ForceKind kind, bool needs_list) {
  if (int r = check_ctx(c)) return r;
  if (needs_list && !c->list_valid) return c->fail(DEM_ERR_STATE, "neighbour list not built for the current positions");
  if (needs_list && c->max_rad > 0. && c->pp_live() && c->cfg.cutoff < 2. * c->max_rad * (1. - 1e-12))
    return c->fail(DEM_ERR_ARG, "cutoff is below 2 r_max: contacts would be missed");
  hipError_t e = launch_force(kind, c->step_args(), false, c->grid_blocks, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "force launch");
  if (kind != FK_PW_ACCUM && kind != FK_BOND_ACCUM) c->cC ^= 1;
  return c->check_and_clear_flags();
}

int dem_force_pp(dem_ctx* c) { return force_call(c, FK_PP_ACCUM, true); }
int dem_force_pw(dem_ctx* c) { return force_call(c, FK_PW_ACCUM, false); }
int dem_compute_forces(dem_ctx* c) { return force_call(c, FK_PP_PW_FORCES, true); }

static int kick_call(dem_ctx* c, int kick, int drift) {
  if (int r = check_ctx(c)) return r;
  StepArgs a = c->step_args();
  hipError_t e = launch_kick(a, c->xr[c->cP], c->uw[c->cP], c->ft[c->cS], kick, drift, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "stage launch");
  return DEM_OK;
}
int dem_stage1(dem_ctx* c) { return kick_call(c, 1, 0); }
int dem_stage2(dem_ctx* c) { return kick_call(c, 0, 1); }
int dem_stage3(dem_ctx* c) { return kick_call(c, 1, 0); }

int dem_set_diag_every(dem_ctx* c, int every) {
  if (int r = check_ctx(c)) return r;
  c->diag_every = every > 0 ? every : 0;
  // rows for the calls of up to 64 samples, allocated here rather than in a step
  return c->diag_every > 0 ? c->ensure_diag_rows(64) : DEM_OK;
}

// Enqueues nsteps full steps (the first kick, then fused steps, the last one
// ending with stage3 and F, T written) without synchronising.  Sampled
// diagnostic rows go to diag_rows[*row...].
static int enqueue_steps(dem_ctx* c, int nsteps, int* row) {
  const int K = c->cfg.rebuild_every;
  const bool autoK = K == DEM_REBUILD_AUTO;
  // stage1 + stage2 of the first step (later ones are fused into the force kernel)
  StepArgs a0 = c->step_args();
  hipError_t e = launch_kick(a0, c->xr[c->cP], c->uw[c->cP], c->ft[c->cS], 1, 1, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "stage1/2 launch");
  bool pending = false; // the previous pass computed a check, to be read after this pass
  for (int s = 0; s < nsteps; ++s) {
    const int64_t gstep = c->step;
    if (!c->list_valid || (autoK ? c->auto_due : gstep % K == 0)) {
      int r = c->rebuild(!c->built_once, c->list_valid, false);
      if (r) return r;
      if (autoK) c->auto_after_rebuild();
      pending = false; // measured against the old build: dropped
    }
    const bool sample = c->diag_every > 0 && gstep % c->diag_every == 0;
    const bool last = (s == nsteps - 1);
    StepArgs a = c->step_args();
    // the displacement check after this step, computed by its own force pass
    const bool check = autoK && !pending && c->auto_left <= 1;
    if (check) {
      e = launch_zero_u64(c->d_disp, 2, c->stream);
      if (e != hipSuccess) return c->hip_fail(e, "displacement check clear");
      a.xb = c->xb;
      a.disp = c->d_disp;
    }
    c->mark_begin(0, sample || check);
    e = launch_force(last ? FK_STEP_LAST : FK_STEP_FUSED, a, sample, c->grid_blocks, c->stream);
    c->mark_end();
    if (e != hipSuccess) return c->hip_fail(e, "step launch");
    c->cP ^= 1;
    c->cC ^= 1;
    if (sample) {
      e = launch_diag_final(c->diag_part, c->ntiles, c->diag_rows + 9 * (*row), (double)gstep, c->stream);
      if (e != hipSuccess) return c->hip_fail(e, "diag launch");
      ++*row;
    }
    c->step++;
    if (pending) {
      if (int r = c->auto_check_read(1)) return r;
      pending = false;
    } else if (autoK && --c->auto_left <= 0) {
      if (int r = c->auto_check_enqueue()) return r;
      pending = true;
    }
  }
  if (pending)
    if (int r = c->auto_check_read(0)) return r;
  return DEM_OK;
}

// One cycle of 2K steps starting at a multiple of K, from a captured graph:
// captured (and run) on first use or when the launch arguments changed,
// replayed afterwards.  A cycle flips every double buffer an even number of
// times, so the captured pointers stay valid from one replay to the next.
static int graph_cycle(dem_ctx* c) {
  const int K = c->cfg.rebuild_every;
  const int parity = c->cP | (c->cS << 1) | (c->cC << 2);
  const bool valid = c->gexec && c->g_gen == c->gen && c->g_parity == parity && c->g_K == K;
  if (valid) {
    hipError_t e = hipGraphLaunch(c->gexec, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "graph launch");
    c->step += 2 * K;
    c->rebuilds += 2;
    return DEM_OK;
  }
  if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return c->hip_fail(e, "begin capture");
  c->capturing = true;
  int row = 0;
  const int64_t step0 = c->step, reb0 = c->rebuilds;
  int r = enqueue_steps(c, 2 * K, &row);
  c->capturing = false;
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(c->stream, &g);
  if (r) { if (g) hipGraphDestroy(g); return r; }
  if (e != hipSuccess) return c->hip_fail(e, "end capture");
  e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) { c->gexec = nullptr; return c->hip_fail(e, "graph instantiate"); }
  // the capture advanced the host-side counters; the device runs it now
  e = hipGraphLaunch(c->gexec, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "graph launch");
  c->g_gen = c->gen;
  c->g_parity = c->cP | (c->cS << 1) | (c->cC << 2);
  c->g_K = K;
  if (c->g_parity != parity || c->step != step0 + 2 * K || c->rebuilds != reb0 + 2)
    return c->fail(DEM_ERR_STATE, "captured cycle did not return to its starting buffers");
  return DEM_OK;
}

int dem_step(dem_ctx* c, int nsteps) {
  if (int r = check_ctx(c)) return r;
  if (nsteps <= 0) return DEM_OK;
  if (int r = c->check_list_cover()) return r;
  const int K = c->cfg.rebuild_every;
  // diagnostic rows of this call
  int rows = 0;
  if (c->diag_every > 0)
    for (int s = 0; s < nsteps; ++s)
      if ((c->step + s) % c->diag_every == 0) ++rows;
  if (int r = c->ensure_diag_rows(rows)) return r;
  int row = 0;
  hipError_t e;
  // captured launches carry fixed arguments, so steps whose walls translate
  // (a new wall time per step) run eagerly
  const bool moving = c->walls_move && c->cfg.physics_mode == DEM_MODE_CORRECTED;
  const bool graphs = c->use_graph && c->diag_every == 0 && !c->timing && K > 0 && K < (1 << 20) &&
                      !moving;
  if (!graphs) {
    if (int r = enqueue_steps(c, nsteps, &row)) return r;
  } else {
    // eager up to a multiple of K (and through the first build), whole
    // cycles of 2K steps from the graph, the rest eager
    int left = nsteps;
    int head = (int)((K - c->step % K) % K);
    if (head == 0 && (!c->built_once || !c->list_valid)) head = K; // build eagerly first
    head = std::min(head, left);
    if (head > 0) {
      if (int r = enqueue_steps(c, head, &row)) return r;
      if (!c->built_once) {
        // the first build sizes the neighbour rows; settle it before capturing
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "step sync");
      }
      left -= head;
    }
    if (c->step % K != 0 && left > 0) {
      // the head could not reach a multiple of K (first build): finish eagerly
      if (int r = enqueue_steps(c, left, &row)) return r;
      left = 0;
    }
    while (left >= 2 * K) {
      if (int r = graph_cycle(c)) return r;
      left -= 2 * K;
    }
    if (left > 0)
      if (int r = enqueue_steps(c, left, &row)) return r;
  }
  // the rows and the flags into pinned memory, then one synchronisation
  if (row > 0) {
    e = hipMemcpyAsync(c->h_diag_rows, c->diag_rows, (size_t)row * 9 * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "diag read-back");
  }
  e = c->enqueue_flags();
  if (e != hipSuccess) return c->hip_fail(e, "flag read-back");
  const int rf = c->check_and_clear_flags(true);
  if (rf == DEM_ERR_HIP) return rf;
  const double nD = (double)c->n;
  for (int r = 0; r < row; ++r) {
    double q[9];
    std::copy(c->h_diag_rows + (size_t)r * 9, c->h_diag_rows + (size_t)r * 9 + 9, q);
    q[6] /= nD; q[7] /= nD; q[8] /= nD; // sums -> per-particle means
    c->diag_hist.insert(c->diag_hist.end(), q, q + 9);
    c->last_sample[0] = q[6]; c->last_sample[1] = q[7]; c->last_sample[2] = q[8];
  }
  c->collect_timing();
  return rf;
}

int dem_set_graph(dem_ctx* c, int enable) {
  if (int r = check_ctx(c)) return r;
  c->use_graph = enable != 0;
  return DEM_OK;
}

} // extern "C"

namespace {

// Exact copy of a context's live state: the current half of every double
// buffer, the neighbour rows and the host-side counters.
struct Snapshot {
  std::vector<std::pair<void*, void*>> bufs; // (live buffer, copy)
  std::vector<size_t> bytes;
  int cP, cS, cC, max_n, max_nn_seen;
  int64_t step, rebuilds;
  bool list_valid, built_once;
  size_t diag_rows;
  double last_sample[3];
  dem_config cfg;
  int grid_blocks, diag_every;
  bool use_graph, timing;
  int force_pass, bidx, tile_rows;
  int auto_left, auto_first;
  int64_t auto_since;
  bool auto_due;
  double auto_v;

  void add(void* p, size_t nbytes, hipError_t* e) {
    if (!p || nbytes == 0) return;
    void* q = nullptr;
    if (*e == hipSuccess) *e = hipMalloc(&q, nbytes);
    if (*e == hipSuccess) *e = hipMemcpy(q, p, nbytes, hipMemcpyDeviceToDevice);
    bufs.push_back({p, q});
    bytes.push_back(nbytes);
  }
  ~Snapshot() {
    for (auto& b : bufs) hipFree(b.second);
  }
};

// Buffers in the order restore() expects them; nbr is the last one.
void snapshot_buffers(dem_ctx* c, Snapshot* s, hipError_t* e) {
  const size_t n = c->n, ld = c->ld;
  s->add(c->xr[c->cP], n * sizeof(double4), e);
  s->add(c->uw[c->cP], 6 * n * sizeof(double), e);
  s->add(c->ft[c->cS], 6 * n * sizeof(double), e);
  s->add(c->mat[c->cS], n * sizeof(double4), e);
  s->add(c->mI[c->cS], n * sizeof(double2), e);
  s->add(c->id[c->cS], n * sizeof(int), e);
  s->add(c->sw[c->cS], (size_t)DEM_MAX_WALLS * 3 * ld * sizeof(double), e);
  s->add(c->cmeta[c->cC], n * sizeof(int), e);
  s->add(c->cj[c->cC], (size_t)c->max_c * ld * sizeof(int), e);
  s->add(c->ch[c->cC], (size_t)c->max_c * 3 * ld * sizeof(double), e);
  s->add(c->nn, n * sizeof(int), e);
  s->add(c->xb, 3 * n * sizeof(double), e);
  s->add(c->flags, sizeof(DevFlags), e);
  if (c->bonds_on) {
    s->add(c->nbond[c->cS], n * sizeof(int), e);
    s->add(c->bj[c->cS], (size_t)c->max_b * ld * sizeof(int), e);
    s->add(c->bl[c->cS], (size_t)c->max_b * ld * sizeof(double), e);
    s->add(c->bh[c->cS], (size_t)9 * c->max_b * ld * sizeof(double), e);
    s->add(c->bEG[c->cS], n * sizeof(double2), e);
  }
  if (c->n_groups > 0) s->add(c->vgroup[c->cS], n * sizeof(int), e);
  if (c->fext[c->cS]) s->add(c->fext[c->cS], 3 * n * sizeof(double), e);
  s->add(c->nbr, (size_t)c->max_n * ld * sizeof(int), e);
}

int take_snapshot(dem_ctx* c, Snapshot* s) {
  hipError_t e = hipStreamSynchronize(c->stream);
  s->cP = c->cP; s->cS = c->cS; s->cC = c->cC;
  s->max_n = c->max_n; s->max_nn_seen = c->max_nn_seen;
  s->step = c->step; s->rebuilds = c->rebuilds;
  s->list_valid = c->list_valid; s->built_once = c->built_once;
  s->diag_rows = c->diag_hist.size();
  std::copy(c->last_sample, c->last_sample + 3, s->last_sample);
  s->cfg = c->cfg; s->grid_blocks = c->grid_blocks; s->diag_every = c->diag_every;
  s->use_graph = c->use_graph; s->timing = c->timing; s->force_pass = c->force_pass;
  s->bidx = c->bidx; s->tile_rows = c->tile_rows;
  s->auto_left = c->auto_left; s->auto_first = c->auto_first; s->auto_due = c->auto_due;
  s->auto_since = c->auto_since;
  s->auto_v = c->auto_v;
  if (e == hipSuccess) snapshot_buffers(c, s, &e);
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "autotune snapshot");
}

int restore_snapshot(dem_ctx* c, const Snapshot& s) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "autotune restore");
  if (c->max_n != s.max_n)
    if (int r = c->alloc_nbr(s.max_n)) return r;
  c->cP = s.cP; c->cS = s.cS; c->cC = s.cC;
  // the live pointers of the snapshot are the current halves again; the
  // neighbour rows may have been re-allocated, so take the current one
  for (size_t k = 0; k < s.bufs.size(); ++k) {
    void* dst = (k + 1 == s.bufs.size()) ? (void*)c->nbr : s.bufs[k].first;
    if (e == hipSuccess) e = hipMemcpy(dst, s.bufs[k].second, s.bytes[k], hipMemcpyDeviceToDevice);
  }
  if (e != hipSuccess) return c->hip_fail(e, "autotune restore");
  c->max_nn_seen = s.max_nn_seen;
  c->step = s.step; c->rebuilds = s.rebuilds;
  c->list_valid = s.list_valid; c->built_once = s.built_once;
  c->diag_hist.resize(s.diag_rows);
  std::copy(s.last_sample, s.last_sample + 3, c->last_sample);
  c->cfg = s.cfg; c->grid_blocks = s.grid_blocks; c->diag_every = s.diag_every;
  c->use_graph = s.use_graph; c->timing = s.timing; c->force_pass = s.force_pass;
  c->select_block(s.bidx);
  if (s.tile_rows != c->tile_rows)
    if (int r = c->set_tile_rows(s.tile_rows)) return r;
  c->auto_left = s.auto_left; c->auto_first = s.auto_first; c->auto_due = s.auto_due;
  c->auto_since = s.auto_since;
  c->auto_v = s.auto_v;
  ++c->gen;
  c->err.clear();
  return DEM_OK;
}

} // namespace

extern "C" {

int dem_autotune(dem_ctx* c, int trial_steps, dem_tune_result* out) {
  if (int r = check_ctx(c)) return r;
  if (!out) return c->fail(DEM_ERR_ARG, "null result");
  if (trial_steps <= 0) return c->fail(DEM_ERR_ARG, "trial_steps must be positive");
  *out = dem_tune_result{};
  Snapshot snap;
  if (int r = take_snapshot(c, &snap)) return r;
  const int cus = c->cus;
  std::vector<int> Ks;
  const int K0 = snap.cfg.rebuild_every;
  if (snap.cfg.skin > 0.) Ks = {5, 10, 20, 40, 80, DEM_REBUILD_AUTO};
  if (std::find(Ks.begin(), Ks.end(), K0) == Ks.end()) Ks.push_back(K0);
  const std::vector<int> grids = {0, 2 * cus, 8 * cus};
  // the force-pass shapes and tile sizes exist for systems without bonds
  const int n_pass = c->bonds_on ? 1 : 3;
  double best = -1.;
  int nt = 0;
  auto trial = [&](int K, int g, int graph, int pass, int bk, int rows) -> int {
    if (nt >= DEM_TUNE_MAX_TRIALS) return DEM_OK;
    if (int r = restore_snapshot(c, snap)) return r;
    c->cfg.rebuild_every = K;
    c->grid_blocks = g;
    c->use_graph = graph != 0;
    c->force_pass = pass;
    c->diag_every = 0;
    c->timing = false;
    ++c->gen;
    int r = dem_set_shape(c, kForceBlocks[bk], rows);
    // whole captured cycles of 2K steps (automatic cadence: periods of 40)
    const int period = K > 0 ? 2 * K : 40;
    const int steps = ((std::max(trial_steps, period) + period - 1) / period) * period;
    double ms = -1.;
    if (r == DEM_OK) r = dem_step(c, period); // warm-up: builds, captures
    if (r == DEM_OK) {
      auto t0 = std::chrono::steady_clock::now();
      r = dem_step(c, steps);
      auto t1 = std::chrono::steady_clock::now();
      if (r == DEM_OK) ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
    }
    if (r == DEM_ERR_HIP) return r;
    out->trial_rebuild_every[nt] = K;
    out->trial_grid_blocks[nt] = g;
    out->trial_graph[nt] = graph;
    out->trial_force_pass[nt] = pass;
    out->trial_block[nt] = kForceBlocks[bk];
    out->trial_tile_rows[nt] = rows;
    out->trial_ms_per_step[nt] = ms;
    if (ms > 0. && (best < 0. || ms < best)) {
      best = ms;
      out->rebuild_every = K;
      out->grid_blocks = g;
      out->use_graph = graph;
      out->force_pass = pass;
      out->block = kForceBlocks[bk];
      out->tile_rows = rows;
    }
    ++nt;
    return DEM_OK;
  };
  // stage 1: cadence x launch grid x graphs x force-pass shape, at the
  // context's tile size and cell tiles
  for (int K : Ks)
    for (int g : grids)
      for (int graph = 0; graph < 2; ++graph)
        for (int pass = 0; pass < n_pass; ++pass)
          if (int r = trial(K, g, graph, pass, snap.bidx, snap.tile_rows)) return r;
  // stage 2: tile sizes (systems without bonds) x cell-tile rows at the best
  // of stage 1
  if (best > 0.) {
    const int K = out->rebuild_every, g = out->grid_blocks, graph = out->use_graph, pass = out->force_pass;
    const std::vector<int> bks = c->bonds_on ? std::vector<int>{1} : std::vector<int>{0, 1, 2};
    for (int bk : bks)
      for (int rows : {4, 8, 16})
        if (bk != snap.bidx || rows != snap.tile_rows)
          if (int r = trial(K, g, graph, pass, bk, rows)) return r;
  }
  out->n_trials = nt;
  out->ms_per_step = best;
  if (int r = restore_snapshot(c, snap)) return r;
  if (best < 0.) return c->fail(DEM_ERR_STATE, "every autotune candidate failed");
  c->cfg.rebuild_every = out->rebuild_every;
  c->grid_blocks = out->grid_blocks;
  c->use_graph = out->use_graph != 0;
  c->force_pass = out->force_pass;
  ++c->gen;
  return dem_set_shape(c, out->block, out->tile_rows);
}

int dem_get_diag_history(dem_ctx* c, double* rows, int cap, int* nrows) {
  if (int r = check_ctx(c)) return r;
  const int total = (int)(c->diag_hist.size() / 9);
  if (nrows) *nrows = total;
  if (rows) {
    const int m = std::min(cap, total);
    std::copy(c->diag_hist.begin(), c->diag_hist.begin() + (size_t)m * 9, rows);
  }
  return DEM_OK;
}

int dem_diagnostics(dem_ctx* c, double out[5]) {
  if (int r = check_ctx(c)) return r;
  StepArgs a = c->step_args();
  hipError_t e = launch_diag_current(a, c->uw[c->cP], c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "diag launch");
  if (int r = c->ensure_diag_rows(1)) return r;
  e = launch_diag_final(c->diag_part, c->ntiles, c->diag_rows, (double)c->step, c->stream);
  double h[9];
  if (e == hipSuccess) e = hipMemcpyAsync(h, c->diag_rows, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "diagnostics");
  for (int k = 0; k < 5; ++k) out[k] = h[1 + k];
  return DEM_OK;
}

int dem_get_counters(dem_ctx* c, dem_counters* out) {
  if (int r = check_ctx(c)) return r;
  if (!out) return c->fail(DEM_ERR_ARG, "null output");
  DevFlags f{};
  hipError_t e = hipMemcpyAsync(&f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "counters");
  out->steps = c->step;
  out->rebuilds = c->rebuilds;
  out->max_neighbors_seen = std::max(c->max_nn_seen, f.max_nn);
  out->neighbor_capacity = c->max_n;
  out->mean_neighbors = c->last_sample[0];
  out->mean_contacts = c->last_sample[1];
  out->mean_wall_contacts = c->last_sample[2];
  out->bonds_broken = f.bonds_broken;
  out->displacement_checks = c->auto_checks;
  return DEM_OK;
}

int dem_synchronize(dem_ctx* c) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "sync");
}

int dem_set_timing(dem_ctx* c, int enable) {
  if (int r = check_ctx(c)) return r;
  hipStreamSynchronize(c->stream);
  c->collect_timing();
  c->timing = enable != 0;
  c->timing_every = enable > 1 ? enable : 1;
  c->timing_seen = 0;
  c->t_ms[0] = c->t_ms[1] = 0.;
  c->t_launches[0] = c->t_launches[1] = 0;
  return DEM_OK;
}

int dem_get_timing(dem_ctx* c, double ms[2], int64_t launches[2]) {
  if (int r = check_ctx(c)) return r;
  hipStreamSynchronize(c->stream);
  c->collect_timing();
  for (int k = 0; k < 2; ++k) {
    if (ms) ms[k] = c->t_ms[k];
    if (launches) launches[k] = c->t_launches[k];
  }
  return DEM_OK;
}

// ---- bonded DEM, groups, external force -----------------------------------

int dem_enable_bonds(dem_ctx* c, const dem_bond_config* b) {
  if (int r = check_ctx(c)) return r;
  if (!b) return c->fail(DEM_ERR_ARG, "null bond config");
  if (c->nsys > 1) return c->fail(DEM_ERR_STATE, "bonds are not available in an ensemble context");
  if (b->max_bonds <= 0 || b->max_bonds > DEM_MAX_BONDS)
    return c->fail(DEM_ERR_ARG, "max_bonds must be in [1, 32]");
  if (!(b->beta_bond >= 0.) || !(b->radius_factor > 0.))
    return c->fail(DEM_ERR_ARG, "beta_bond must be >= 0 and radius_factor > 0");
  if (c->bonds_on) {
    if (b->max_bonds != c->max_b) return c->fail(DEM_ERR_ARG, "max_bonds cannot change once bonds are enabled");
    hipStreamSynchronize(c->stream);
    c->bcfg = *b;
    ++c->gen;
    return DEM_OK;
  }
  std::vector<int> id;
  if (int r = c->sorted_ids(&id)) return r;
  const size_t n = (size_t)c->n, ld = (size_t)c->ld, mb = (size_t)b->max_bonds;
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2; ++k) {
    c->nbond[k] = dalloc<int>(n, &e);
    c->bj[k] = dalloc<int>(mb * ld, &e);
    c->bl[k] = dalloc<double>(mb * ld, &e);
    c->bh[k] = dalloc<double>(9 * mb * ld, &e);
    c->bEG[k] = dalloc<double2>(n, &e);
  }
  if (e == hipSuccess) e = hipMemset(c->nbond[c->cS], 0, n * sizeof(int));
  if (e != hipSuccess) {
    for (int k = 0; k < 2; ++k) {
      hipFree(c->nbond[k]); hipFree(c->bj[k]); hipFree(c->bl[k]); hipFree(c->bh[k]); hipFree(c->bEG[k]);
      c->nbond[k] = c->bj[k] = nullptr; c->bl[k] = c->bh[k] = nullptr; c->bEG[k] = nullptr;
    }
    return c->hip_fail(e, "hipMalloc bonds");
  }
  c->max_b = b->max_bonds;
  c->bcfg = *b;
  c->bonds_on = true;
  c->select_block(1); // bonded systems run tiles of kBlock
  ++c->gen;
  return c->upload_bEG(id);
}

int dem_setup_bonds(dem_ctx* c, double factor) {
  if (int r = check_ctx(c)) return r;
  if (!c->bonds_on) return c->fail(DEM_ERR_STATE, "bonds are not enabled (dem_enable_bonds)");
  if (!(factor > 0.)) return c->fail(DEM_ERR_ARG, "bond factor must be positive");
  const double rmax = *std::max_element(c->h_rad.begin(), c->h_rad.end());
  if (factor * rmax > c->cfg.cutoff)
    return c->fail(DEM_ERR_ARG, "factor * max radius exceeds the neighbour cutoff: bonds would be missed");
  if (!c->list_valid)
    if (int r = c->rebuild(true, false)) return r;
  BondSetupArgs a{};
  a.n = c->n; a.ld = c->ld; a.max_b = c->max_b; a.factor = factor;
  a.xr = c->xr[c->cP]; a.nn = c->nn; a.nbr = c->nbr;
  a.nbond = c->nbond[c->cS]; a.bj = c->bj[c->cS]; a.bl = c->bl[c->cS]; a.bh = c->bh[c->cS];
  a.flags = c->flags;
  hipError_t e = launch_setup_bonds(a, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "bond set-up launch");
  return c->check_and_clear_flags();
}

int dem_upload_bonds(dem_ctx* c, const dem_bonds_view* v) {
  if (int r = check_ctx(c)) return r;
  if (!c->bonds_on) return c->fail(DEM_ERR_STATE, "bonds are not enabled (dem_enable_bonds)");
  if (!v || v->n != c->n || v->max_bonds <= 0 || !v->idx || !v->count)
    return c->fail(DEM_ERR_ARG, "bond view needs n, max_bonds, idx and count");
  std::vector<int> id;
  if (int r = c->sorted_ids(&id)) return r;
  const int n = c->n, M = v->max_bonds, MB = c->max_b;
  const size_t ld = c->ld;
  std::vector<int> pos(n);
  for (int p = 0; p < n; ++p) pos[id[p]] = p;
  std::vector<int> nb(n, 0), bjh((size_t)MB * ld, -1);
  std::vector<double> blh((size_t)MB * ld, 0.), bhh((size_t)9 * MB * ld, 0.);
  const double* hist[9] = {v->ft_x, v->ft_y, v->ft_z, v->mn_x, v->mn_y, v->mn_z, v->mt_x, v->mt_y, v->mt_z};
  for (int p = 0; p < n; ++p) {
    const int I = id[p];
    const int cnt = v->count[I];
    if (cnt < 0 || cnt > M) return c->fail(DEM_ERR_ARG, "bond count outside [0, max_bonds] of the view");
    if (cnt > MB) return c->fail(DEM_ERR_BOND_OVERFLOW, "uploaded bonds exceed max_bonds");
    nb[p] = cnt;
    for (int k = 0; k < cnt; ++k) {
      const size_t src = (size_t)I * M + k;
      const int J = v->idx[src];
      if (J >= n) return c->fail(DEM_ERR_ARG, "bond partner id out of range");
      bjh[(size_t)k * ld + p] = J >= 0 ? pos[J] : -1;
      blh[(size_t)k * ld + p] = v->init_len ? v->init_len[src] : 0.;
      for (int q = 0; q < 9; ++q) bhh[(size_t)(9 * k + q) * ld + p] = hist[q] ? hist[q][src] : 0.;
    }
  }
  hipError_t e = hipMemcpy(c->nbond[c->cS], nb.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->bj[c->cS], bjh.data(), bjh.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->bl[c->cS], blh.data(), blh.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->bh[c->cS], bhh.data(), bhh.size() * sizeof(double), hipMemcpyHostToDevice);
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "upload bonds");
}

int dem_download_bonds(dem_ctx* c, dem_bonds_view* v) {
  if (int r = check_ctx(c)) return r;
  if (!c->bonds_on) return c->fail(DEM_ERR_STATE, "bonds are not enabled (dem_enable_bonds)");
  if (!v || v->n != c->n || v->max_bonds <= 0) return c->fail(DEM_ERR_ARG, "bond view needs n and max_bonds");
  std::vector<int> id;
  if (int r = c->sorted_ids(&id)) return r;
  const int n = c->n, M = v->max_bonds, MB = c->max_b;
  const size_t ld = c->ld;
  std::vector<int> nb(n), bjh((size_t)MB * ld);
  std::vector<double> blh((size_t)MB * ld), bhh((size_t)9 * MB * ld);
  hipError_t e = hipMemcpy(nb.data(), c->nbond[c->cS], n * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(bjh.data(), c->bj[c->cS], bjh.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(blh.data(), c->bl[c->cS], blh.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(bhh.data(), c->bh[c->cS], bhh.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return c->hip_fail(e, "download bonds");
  double* hist[9] = {v->ft_x, v->ft_y, v->ft_z, v->mn_x, v->mn_y, v->mn_z, v->mt_x, v->mt_y, v->mt_z};
  for (int p = 0; p < n; ++p) {
    if (nb[p] > M) return c->fail(DEM_ERR_ARG, "view max_bonds is smaller than a particle's bond count");
  }
  for (int p = 0; p < n; ++p) {
    const int I = id[p];
    for (int k = 0; k < M; ++k) {
      const size_t dst = (size_t)I * M + k;
      const bool live = k < nb[p];
      const int j = live ? bjh[(size_t)k * ld + p] : -1;
      if (v->idx) v->idx[dst] = j >= 0 ? id[j] : -1;
      if (v->init_len) v->init_len[dst] = live ? blh[(size_t)k * ld + p] : 0.;
      for (int q = 0; q < 9; ++q)
        if (hist[q]) hist[q][dst] = live ? bhh[(size_t)(9 * k + q) * ld + p] : 0.;
    }
    if (v->count) v->count[I] = nb[p];
  }
  return DEM_OK;
}

int dem_bond_forces(dem_ctx* c) {
  if (int r = check_ctx(c)) return r;
  if (!c->bonds_on) return c->fail(DEM_ERR_STATE, "bonds are not enabled (dem_enable_bonds)");
  return force_call(c, FK_BOND_ACCUM, false);
}

int dem_set_external_force(dem_ctx* c, const double* f) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "sync");
  if (!f) {
    for (int k = 0; k < 2; ++k) { hipFree(c->fext[k]); c->fext[k] = nullptr; }
    ++c->gen;
    return DEM_OK;
  }
  std::vector<int> id;
  if (int r = c->sorted_ids(&id)) return r;
  const size_t n = (size_t)c->n;
  for (int k = 0; k < 2; ++k)
    if (!c->fext[k]) c->fext[k] = dalloc<double>(3 * n, &e);
  if (e != hipSuccess) return c->hip_fail(e, "hipMalloc external force");
  std::vector<double> h(3 * n);
  for (size_t p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) h[3 * p + d] = f[3 * (size_t)id[p] + d];
  e = hipMemcpy(c->fext[c->cS], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  ++c->gen;
  return e == hipSuccess ? DEM_OK : c->hip_fail(e, "upload external force");
}

int dem_set_groups(dem_ctx* c, const int* group, int ngroups, const int* kind, const double* vel) {
  if (int r = check_ctx(c)) return r;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "sync");
  if (!group) {
    for (int k = 0; k < 2; ++k) { hipFree(c->vgroup[k]); c->vgroup[k] = nullptr; }
    c->n_groups = 0;
    ++c->gen;
    return DEM_OK;
  }
  if (ngroups < 1 || ngroups > DEM_MAX_GROUPS || !kind)
    return c->fail(DEM_ERR_ARG, "ngroups must be in [1, 8] with a kind per group");
  for (int g = 0; g < ngroups; ++g) {
    if (kind[g] != DEM_GROUP_VELOCITY && kind[g] != DEM_GROUP_FIXED_FORCE)
      return c->fail(DEM_ERR_ARG, "unknown group kind");
    if (kind[g] == DEM_GROUP_VELOCITY && !vel)
      return c->fail(DEM_ERR_ARG, "a velocity group needs vel");
  }
  const int n = c->n;
  for (int i = 0; i < n; ++i)
    if (group[i] < 0 || group[i] > ngroups) return c->fail(DEM_ERR_ARG, "group index outside [0, ngroups]");
  std::vector<int> id;
  if (int r = c->sorted_ids(&id)) return r;
  for (int k = 0; k < 2; ++k)
    if (!c->vgroup[k]) c->vgroup[k] = dalloc<int>((size_t)n, &e);
  if (e != hipSuccess) return c->hip_fail(e, "hipMalloc groups");
  std::vector<int> h(n);
  for (int p = 0; p < n; ++p) h[p] = group[id[p]];
  e = hipMemcpy(c->vgroup[c->cS], h.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return c->hip_fail(e, "upload groups");
  for (int g = 0; g < DEM_MAX_GROUPS; ++g) {
    c->gkind[g] = g < ngroups ? kind[g] : 0;
    for (int d = 0; d < 3; ++d) c->gvel[g][d] = (g < ngroups && vel) ? vel[3 * g + d] : 0.;
  }
  c->n_groups = ngroups;
  ++c->gen;
  return DEM_OK;
}

} // extern "C"
