// psx_h5.cpp -- PhysField dump/restore in the reference's HDF5 layout
// (host C++ over the HDF5 C API; declared in include/psx_h5.h).
#include "../../include/psx_h5.h"

#include <hdf5.h>

#include <cstdio>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Closes an HDF5 identifier with the matching call at scope exit.
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() { if (id >= 0) close(id); }
  operator hid_t() const { return id; }
  bool ok() const { return id >= 0; }
};

bool attr_string(hid_t obj, const char* name, const std::string& v) {
  Id type(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!type.ok() || H5Tset_size(type, v.empty() ? 1 : v.size()) < 0) return false;
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Awrite(a, type, v.c_str()) >= 0;
}

template <typename T>
bool attr_vec(hid_t obj, const char* name, hid_t mem_type, hid_t file_type, const std::vector<T>& v) {
  const hsize_t n = v.size();
  Id space(H5Screate_simple(1, &n, nullptr), H5Sclose);
  Id a(H5Acreate2(obj, name, file_type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Awrite(a, mem_type, v.data()) >= 0;
}

template <typename T>
bool attr_scalar(hid_t obj, const char* name, hid_t mem_type, hid_t file_type, T v) {
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, name, file_type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Awrite(a, mem_type, &v) >= 0;
}

}  // namespace

extern "C" void psx_h5_last_error(char* buf, unsigned long len) {
  if (!buf || len == 0) return;
  std::snprintf(buf, len, "%s", g_err.c_str());
}

extern "C" int psx_h5_write_physfield(const char* path, const char* field, int nx, int ny, int nz,
                                      const double* cell, const char* dens_name, const double* dens,
                                      const char* conj_name, const double* conj, long long step) {
  if (!path || !field || !cell || !dens_name || !dens || !conj_name || !conj)
    return fail(-1, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "grid dims must be >= 1");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  {
    // the reference writes a dump only if its file does not exist yet
    std::FILE* f = std::fopen(path, "rb");
    if (f) { std::fclose(f); return 1; }
  }
  Id file(H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot create ") + path);
  // dataset shape (nx, ny, 1) in 2D, (nx, ny, nz, 1) in 3D
  const int idim = nz == 1 ? 2 : 3;
  std::vector<hsize_t> dims = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
  if (idim == 3) dims.push_back(1);
  const std::string mesh = std::string(field) + "Global";
  const std::string limits = std::string(field) + "GlobalLimits";
  const std::string labels = idim == 2 ? "x,y" : "x,y,z";
  const std::vector<double> sizes = {cell[0], cell[1], cell[2]};
  const char* names[2] = {dens_name, conj_name};
  const double* data[2] = {dens, conj};
  for (int k = 0; k < 2; ++k) {
    Id space(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose);
    Id ds(H5Dcreate2(file, names[k], H5T_IEEE_F64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
          H5Dclose);
    if (!ds.ok() || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data[k]) < 0)
      return fail(-3, std::string("cannot write dataset ") + names[k]);
    bool ok = attr_string(ds, "vsType", "variable") && attr_string(ds, "vsMesh", mesh) &&
              attr_string(ds, "vsLimits", limits) &&
              attr_vec(ds, "GridSizes", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, sizes) &&
              attr_string(ds, "vsCentering", "zonal") && attr_string(ds, "vsLabels", labels) &&
              attr_string(ds, "vsAxisLabels", labels) && attr_string(ds, "vsTimeGroup", "time");
    if (!ok) return fail(-3, std::string("cannot write the attributes of ") + names[k]);
  }
  std::vector<int> ncell(idim), start(idim, 0);
  std::vector<double> lower(idim, 0.0), upper(idim);
  const int n[3] = {nx, ny, nz};
  for (int d = 0; d < idim; ++d) {
    ncell[d] = n[d];
    upper[d] = (double)n[d] * cell[d];
  }
  {
    Id g(H5Gcreate2(file, limits.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    bool ok = g.ok() && attr_string(g, "vsType", "limits") && attr_string(g, "vsKind", "Cartesian") &&
              attr_vec(g, "vsLowerBounds", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, lower) &&
              attr_vec(g, "vsUpperBounds", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, upper);
    if (!ok) return fail(-3, "cannot write group " + limits);
  }
  {
    Id g(H5Gcreate2(file, mesh.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    bool ok = g.ok() && attr_string(g, "vsKind", "uniform") && attr_string(g, "vsType", "mesh") &&
              attr_vec(g, "vsLowerBounds", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, lower) &&
              attr_vec(g, "vsUpperBounds", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, upper) &&
              attr_vec(g, "vsNumCells", H5T_NATIVE_INT, H5T_STD_I32LE, ncell) &&
              attr_vec(g, "vsStartCell", H5T_NATIVE_INT, H5T_STD_I32LE, start);
    if (!ok) return fail(-3, "cannot write group " + mesh);
  }
  {
    Id g(H5Gcreate2(file, "time", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    bool ok = g.ok() && attr_string(g, "vsType", "time") &&
              attr_scalar(g, "vsStep", H5T_NATIVE_LLONG, H5T_STD_I64LE, step) &&
              attr_scalar(g, "vsTime", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, (double)step);
    if (!ok) return fail(-3, "cannot write group time");
  }
  return 0;
}

extern "C" int psx_h5_read_physfield(const char* path, int nx, int ny, int nz,
                                     const char* dens_name, double* dens,
                                     const char* conj_name, double* conj) {
  if (!path) return fail(-1, "null path");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  const hsize_t want = (hsize_t)nx * ny * nz;
  const char* names[2] = {dens_name, conj_name};
  double* out[2] = {dens, conj};
  for (int k = 0; k < 2; ++k) {
    if (!out[k]) continue;
    if (!names[k]) return fail(-1, "null dataset name");
    Id ds(H5Dopen2(file, names[k], H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(-4, std::string("no dataset ") + names[k] + " in " + path);
    Id space(H5Dget_space(ds), H5Sclose);
    const hssize_t cnt = H5Sget_simple_extent_npoints(space);
    if (cnt < 0 || (hsize_t)cnt != want)
      return fail(-5, std::string("dataset ") + names[k] + " has " + std::to_string((long long)cnt) +
                      " elements, expected " + std::to_string((unsigned long long)want));
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out[k]) < 0)
      return fail(-3, std::string("cannot read dataset ") + names[k]);
  }
  return 0;
}
