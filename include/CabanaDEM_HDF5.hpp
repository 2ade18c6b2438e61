// CabanaDEM_HDF5.hpp -- particle output in the reference's HDF5 layout.
//
// The reference writes each output step through Cabana's HDF5 particle
// writer (src/CabanaDEM_Particles.hpp:299-330): a file
// <output_folder>/particles_<step>.h5 holding one dataset per field, named by
// the field's slice label, with a "Time" attribute on the root group, and an
// XDMF file particles_<step>.xmf that describes those datasets to XDMF readers
// (ParaView).  The post-processing reads f["forces"][i][c] and
// f.attrs["Time"] (automate.py:133-135,531).  This header writes the same
// layout with the HDF5 C API on one process: scalars as [n], vectors as
// [n][3], per-contact histories as [n][MaxCnts]; FP64 and int32 as stored.
//
// Compiled in when CABANADEM_ENABLE_HDF5 is defined (include hdf5.h, link
// -lhdf5); Particles::output falls back to CSV otherwise.
#pragma once

#ifdef CABANADEM_ENABLE_HDF5

#include <hdf5.h>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace CabanaDEM {
namespace detail {

struct H5Field {
  std::string label;       // dataset name (the slice label)
  const void* data;        // row-major [n][width]
  int width;               // 1: scalar per particle
  bool is_int;             // int32 (else FP64)
};

// Writes <prefix>_<step>.h5 and <prefix>_<step>.xmf; the first field must be
// the positions (the XDMF geometry).  Returns false on any HDF5 failure.
inline bool write_h5_step(const std::string& prefix, int step, double time, std::size_t n,
                          const std::vector<H5Field>& fields) {
  const std::string stem = prefix + "_" + std::to_string(step);
  const std::string h5 = stem + ".h5";
  hid_t file = H5Fcreate(h5.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return false;
  bool ok = true;
  {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t at = H5Acreate2(file, "Time", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    ok = ok && sp >= 0 && at >= 0 && H5Awrite(at, H5T_NATIVE_DOUBLE, &time) >= 0;
    if (at >= 0) H5Aclose(at);
    if (sp >= 0) H5Sclose(sp);
  }
  for (const H5Field& f : fields) {
    if (!ok) break;
    hsize_t dims[2] = {(hsize_t)n, (hsize_t)f.width};
    hid_t sp = H5Screate_simple(f.width == 1 ? 1 : 2, dims, nullptr);
    hid_t mem = f.is_int ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE;
    hid_t disk = f.is_int ? H5T_STD_I32LE : H5T_IEEE_F64LE;
    hid_t ds = H5Dcreate2(file, f.label.c_str(), disk, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    ok = sp >= 0 && ds >= 0 && (n == 0 || H5Dwrite(ds, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, f.data) >= 0);
    if (ds >= 0) H5Dclose(ds);
    if (sp >= 0) H5Sclose(sp);
  }
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return false;

  // XDMF description: polyvertex topology, XYZ geometry from the positions,
  // one node-centred attribute per other field
  const std::size_t slash = h5.find_last_of('/');
  const std::string h5name = slash == std::string::npos ? h5 : h5.substr(slash + 1);
  std::ofstream x(stem + ".xmf");
  if (!x) return false;
  x << "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
    << "<Xdmf Version=\"2.0\">\n  <Domain>\n    <Grid Name=\"particles\" GridType=\"Uniform\">\n";
  char tbuf[64];
  std::snprintf(tbuf, sizeof(tbuf), "%.17g", time);
  x << "      <Time Value=\"" << tbuf << "\"/>\n"
    << "      <Topology TopologyType=\"Polyvertex\" Dimensions=\"" << n
    << "\" NodesPerElement=\"1\"/>\n";
  auto item = [&](const H5Field& f) {
    x << "        <DataItem Dimensions=\"" << n;
    if (f.width > 1) x << " " << f.width;
    x << "\" NumberType=\"" << (f.is_int ? "Int" : "Float") << "\" Precision=\""
      << (f.is_int ? 4 : 8) << "\" Format=\"HDF\">" << h5name << ":/" << f.label << "</DataItem>\n";
  };
  x << "      <Geometry GeometryType=\"XYZ\">\n";
  item(fields.front());
  x << "      </Geometry>\n";
  for (std::size_t k = 1; k < fields.size(); ++k) {
    const H5Field& f = fields[k];
    const char* kind = f.width == 1 ? "Scalar" : (f.width == 3 ? "Vector" : "Matrix");
    x << "      <Attribute AttributeType=\"" << kind << "\" Center=\"Node\" Name=\"" << f.label << "\">\n";
    item(f);
    x << "      </Attribute>\n";
  }
  x << "    </Grid>\n  </Domain>\n</Xdmf>\n";
  return (bool)x;
}

} // namespace detail
} // namespace CabanaDEM

#endif // CABANADEM_ENABLE_HDF5
