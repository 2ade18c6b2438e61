// Particles::output with HDF5 enabled: the file <out>/particles_<step>.h5
// read back through the HDF5 C API holds every field the reference writes
// (src/CabanaDEM_Particles.hpp:305-326) under its slice label, with the
// values written through the slices, and the "Time" attribute that the
// reference's post-processing reads (automate.py:133-135); the .xmf file
// names the same datasets.  Host only.
//
//   test_hdf5_output <output_dir>      prints OK on success
#include <hdf5.h>

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <CabanaDEM.hpp>

static int fail = 0;
static void check(bool ok, const std::string& what) {
  if (!ok) {
    std::cout << "FAIL " << what << "\n";
    ++fail;
  }
}

template <class T>
static std::vector<T> read_ds(hid_t file, const char* name, std::vector<hsize_t>* dims) {
  std::vector<T> out;
  hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
  if (ds < 0) return out;
  hid_t sp = H5Dget_space(ds);
  const int rank = H5Sget_simple_extent_ndims(sp);
  dims->assign(rank, 0);
  H5Sget_simple_extent_dims(sp, dims->data(), nullptr);
  hsize_t total = 1;
  for (hsize_t d : *dims) total *= d;
  out.resize(total);
  H5Dread(ds, std::is_same<T, int>::value ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL,
          H5P_DEFAULT, out.data());
  H5Sclose(sp);
  H5Dclose(ds);
  return out;
}

int main(int argc, char* argv[]) {
  if (argc < 2) return 2;
  const std::string out = argv[1];
  using exec_space = CabanaDEM::DefaultExecutionSpace;
  using memory_space = typename exec_space::memory_space;
  constexpr int MaxCnts = 6;
  const int n = 11;
  auto p = std::make_shared<CabanaDEM::Particles<memory_space, 3, MaxCnts, 1>>(exec_space(), n, out);
  auto x = p->slicePosition();
  auto u = p->sliceVelocity();
  auto w = p->sliceOmega();
  auto au = p->sliceAcceleration();
  auto f = p->sliceForce();
  auto m = p->sliceMass();
  auto rho = p->sliceDensity();
  auto rad = p->sliceRadius();
  auto ssx = p->sliceTangentialDispSSx();
  auto ssy = p->sliceTangentialDispSSy();
  auto ssz = p->sliceTangentialDispSSz();
  auto idx = p->sliceTangentialDispIdx();
  auto cnt = p->sliceTotalNoTangentialContacts();
  auto val = [](int i, int k, int field) { return 0.5 * i - 0.25 * k + 100. * field + 1e-3 * (i * k); };
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      x(i, d) = val(i, d, 1);
      u(i, d) = val(i, d, 2);
      w(i, d) = val(i, d, 3);
      au(i, d) = val(i, d, 4);
      f(i, d) = val(i, d, 5);
    }
    m(i) = val(i, 0, 6);
    rho(i) = val(i, 0, 7);
    rad(i) = val(i, 0, 8);
    for (int k = 0; k < MaxCnts; ++k) {
      ssx(i, k) = val(i, k, 9);
      ssy(i, k) = val(i, k, 10);
      ssz(i, k) = val(i, k, 11);
      idx(i, k) = (k < i % 4) ? (i + k + 1) % n : -1;
    }
    cnt(i) = i % 4;
  }
  const double time = 1.25e-3;
  p->output(7, time);

  const std::string h5 = out + "/particles_7.h5";
  hid_t file = H5Fopen(h5.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  check(file >= 0, "open " + h5);
  if (file < 0) return 1;
  double t = -1.;
  hid_t at = H5Aopen(file, "Time", H5P_DEFAULT);
  check(at >= 0 && H5Aread(at, H5T_NATIVE_DOUBLE, &t) >= 0 && t == time, "Time attribute");
  if (at >= 0) H5Aclose(at);

  struct Vec { const char* name; int field; };
  for (Vec v : {Vec{"positions", 1}, Vec{"velocities", 2}, Vec{"omega", 3}, Vec{"accelerations", 4},
                Vec{"forces", 5}}) {
    std::vector<hsize_t> dims;
    auto a = read_ds<double>(file, v.name, &dims);
    check(dims.size() == 2 && dims[0] == (hsize_t)n && dims[1] == 3, std::string(v.name) + " shape");
    bool same = a.size() == 3u * n;
    for (int i = 0; same && i < n; ++i)
      for (int d = 0; d < 3; ++d) same = same && a[3 * i + d] == val(i, d, v.field);
    check(same, std::string(v.name) + " values");
  }
  for (Vec v : {Vec{"mass", 6}, Vec{"density", 7}, Vec{"radius", 8}}) {
    std::vector<hsize_t> dims;
    auto a = read_ds<double>(file, v.name, &dims);
    check(dims.size() == 1 && dims[0] == (hsize_t)n, std::string(v.name) + " shape");
    bool same = a.size() == (std::size_t)n;
    for (int i = 0; same && i < n; ++i) same = a[i] == val(i, 0, v.field);
    check(same, std::string(v.name) + " values");
  }
  for (Vec v : {Vec{"tangential_disp_ss_x", 9}, Vec{"tangential_disp_ss_y", 10},
                Vec{"tangential_disp_ss_z", 11}}) {
    std::vector<hsize_t> dims;
    auto a = read_ds<double>(file, v.name, &dims);
    check(dims.size() == 2 && dims[0] == (hsize_t)n && dims[1] == MaxCnts, std::string(v.name) + " shape");
    bool same = a.size() == (std::size_t)n * MaxCnts;
    for (int i = 0; same && i < n; ++i)
      for (int k = 0; k < MaxCnts; ++k) same = same && a[MaxCnts * i + k] == val(i, k, v.field);
    check(same, std::string(v.name) + " values");
  }
  {
    std::vector<hsize_t> dims;
    auto a = read_ds<int>(file, "tangential_disp_idx", &dims);
    check(dims.size() == 2 && dims[1] == MaxCnts, "tangential_disp_idx shape");
    bool same = a.size() == (std::size_t)n * MaxCnts;
    for (int i = 0; same && i < n; ++i)
      for (int k = 0; k < MaxCnts; ++k) same = same && a[MaxCnts * i + k] == ((k < i % 4) ? (i + k + 1) % n : -1);
    check(same, "tangential_disp_idx values");
    auto c = read_ds<int>(file, "total_no_tangential_contacts", &dims);
    check(dims.size() == 1 && dims[0] == (hsize_t)n, "total_no_tangential_contacts shape");
    same = c.size() == (std::size_t)n;
    for (int i = 0; same && i < n; ++i) same = c[i] == i % 4;
    check(same, "total_no_tangential_contacts values");
  }
  H5Fclose(file);

  std::ifstream xs(out + "/particles_7.xmf");
  std::stringstream ss;
  ss << xs.rdbuf();
  const std::string xmf = ss.str();
  for (const char* name : {"particles_7.h5:/positions", "particles_7.h5:/forces",
                           "particles_7.h5:/total_no_tangential_contacts", "Polyvertex"})
    check(xmf.find(name) != std::string::npos, std::string("xmf names ") + name);
  // the bonded container: its own field set and labels
  // (src/CabanaBondedDEM_Particles.hpp:244-267)
  {
    constexpr int MB = 4;
    const int nb = 5;
    CabanaBondedDEM::ParticlesBondedDEM<memory_space, 3, MB> b(exec_space(), nb);
    b.setOutputFolder(out + "/bonded");
    auto bx = b.slicePosition();
    auto bidx = b.sliceBondIdx();
    auto bft = b.sliceBondFtX();
    auto bcnt = b.sliceTotalNoBonds();
    for (int i = 0; i < nb; ++i) {
      for (int d = 0; d < 3; ++d) bx(i, d) = val(i, d, 12);
      for (int k = 0; k < MB; ++k) {
        bidx(i, k) = k < 2 ? (i + k + 1) % nb : -1;
        bft(i, k) = val(i, k, 13);
      }
      bcnt(i) = 2;
    }
    b.output(3, 0.5);
    hid_t bf = H5Fopen((out + "/bonded/particles_3.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    check(bf >= 0, "open bonded step file");
    if (bf >= 0) {
      std::vector<hsize_t> dims;
      auto x = read_ds<double>(bf, "x", &dims);
      bool same = dims.size() == 2 && dims[0] == (hsize_t)nb && dims[1] == 3 && x.size() == 3u * nb;
      for (int i = 0; same && i < nb; ++i)
        for (int d = 0; d < 3; ++d) same = same && x[3 * i + d] == val(i, d, 12);
      check(same, "bonded x");
      auto id = read_ds<int>(bf, "bond_idx", &dims);
      same = dims.size() == 2 && dims[1] == MB && id.size() == (std::size_t)nb * MB;
      for (int i = 0; same && i < nb; ++i)
        for (int k = 0; k < MB; ++k) same = same && id[MB * i + k] == (k < 2 ? (i + k + 1) % nb : -1);
      check(same, "bonded bond_idx");
      auto ft = read_ds<double>(bf, "bond_ft_x", &dims);
      same = ft.size() == (std::size_t)nb * MB;
      for (int i = 0; same && i < nb; ++i)
        for (int k = 0; k < MB; ++k) same = same && ft[MB * i + k] == val(i, k, 13);
      check(same, "bonded bond_ft_x");
      auto c = read_ds<int>(bf, "total_no_bonds", &dims);
      check(c.size() == (std::size_t)nb && c[0] == 2 && c[nb - 1] == 2, "bonded total_no_bonds");
      for (const char* name : {"u", "omega", "au", "force", "m", "rho", "rad", "bond_fn_x", "bond_mt_z",
                               "bond_init_len"})
        check(H5Lexists(bf, name, H5P_DEFAULT) > 0, std::string("bonded dataset ") + name);
      H5Fclose(bf);
    }
  }
  if (fail == 0) std::cout << "OK\n";
  return fail == 0 ? 0 : 1;
}
