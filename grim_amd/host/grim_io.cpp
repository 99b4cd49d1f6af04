// grim's HDF5 primitive dumps (grid::dump / grid::load, src/grid/grid.cpp:
// 367-381, 482-499) for the C ABI of include/grim_io.h.
//
// PETSc's HDF5 viewer writes a DMDA global vector as one dataset in natural
// order (k, j, i, then the component), without ghost zones.  This file writes
// and reads that layout with the HDF5 C library, loaded at run time so that
// the library needs no HDF5 headers or link-time dependency: the handful of
// entry points used are declared here with their public C signatures
// (HDF5 1.10: hid_t is int64_t, hsize_t unsigned long long).
#include "grim_io.h"

#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;       // H5P_DEFAULT
constexpr hid_t kAll = 0;           // H5S_ALL
constexpr unsigned kAccRdonly = 0;  // H5F_ACC_RDONLY
constexpr unsigned kAccTrunc = 2;   // H5F_ACC_TRUNC
constexpr hid_t kErrDefault = 0;    // H5E_DEFAULT

thread_local std::string g_error;

int fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

struct Hdf5 {
  void* handle = nullptr;
  herr_t (*open)();
  herr_t (*eSetAuto)(hid_t, void*, void*);
  hid_t (*fCreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*fOpen)(const char*, unsigned, hid_t);
  herr_t (*fClose)(hid_t);
  hid_t (*sCreateSimple)(int, const hsize_t*, const hsize_t*);
  int (*sGetNdims)(hid_t);
  int (*sGetDims)(hid_t, hsize_t*, hsize_t*);
  herr_t (*sClose)(hid_t);
  hid_t (*dCreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dOpen)(hid_t, const char*, hid_t);
  hid_t (*dGetSpace)(hid_t);
  herr_t (*dWrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dRead)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  herr_t (*dClose)(hid_t);
  const hid_t* nativeDouble;  // H5T_NATIVE_DOUBLE_g
  const hid_t* ieeeF64LE;     // H5T_IEEE_F64LE_g
};

template <class F>
bool sym(void* h, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(h, name));
  return f != nullptr;
}

// The HDF5 library, loaded once per process; nullptr (with g_error set)
// when it cannot be found.
const Hdf5* hdf5() {
  static Hdf5 lib;
  static std::string err;
  static bool tried = false;
  if (tried) {
    if (!lib.handle) g_error = err;
    return lib.handle ? &lib : nullptr;
  }
  tried = true;
  std::vector<std::string> cands;
  if (const char* e = std::getenv("GRIM_HDF5_LIB")) cands.push_back(e);
  cands.push_back("libhdf5.so");
  cands.push_back("/opt/conda/lib/libhdf5.so");
  void* h = nullptr;
  for (const std::string& c : cands)
    if ((h = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!h) {
    err = "grim_io: the HDF5 C library was not found (set GRIM_HDF5_LIB)";
    g_error = err;
    return nullptr;
  }
  Hdf5 L;
  L.handle = h;
  const bool ok = sym(h, "H5open", L.open) && sym(h, "H5Eset_auto2", L.eSetAuto) &&
                  sym(h, "H5Fcreate", L.fCreate) && sym(h, "H5Fopen", L.fOpen) &&
                  sym(h, "H5Fclose", L.fClose) && sym(h, "H5Screate_simple", L.sCreateSimple) &&
                  sym(h, "H5Sget_simple_extent_ndims", L.sGetNdims) &&
                  sym(h, "H5Sget_simple_extent_dims", L.sGetDims) && sym(h, "H5Sclose", L.sClose) &&
                  sym(h, "H5Dcreate2", L.dCreate) && sym(h, "H5Dopen2", L.dOpen) &&
                  sym(h, "H5Dget_space", L.dGetSpace) && sym(h, "H5Dwrite", L.dWrite) &&
                  sym(h, "H5Dread", L.dRead) && sym(h, "H5Dclose", L.dClose) &&
                  sym(h, "H5T_NATIVE_DOUBLE_g", L.nativeDouble) && sym(h, "H5T_IEEE_F64LE_g", L.ieeeF64LE);
  if (!ok || L.open() < 0) {
    err = "grim_io: the HDF5 library lacks an expected entry point";
    g_error = err;
    dlclose(h);
    return nullptr;
  }
  L.eSetAuto(kErrDefault, nullptr, nullptr);  // errors come back as status codes
  lib = L;
  return &lib;
}

// Bulk extents (X1, X2, X3), ghost widths and ghosted plane shape of a grid
// (grim_grid_zones's rule: ghost zones only along active directions).
struct Shape {
  int N[3], g[3], NT[3];
  int64_t n;
  explicit Shape(const grim_grid_config& c) {
    const int Ns[3] = {c.N1, c.N2, c.N3};
    n = 1;
    for (int d = 0; d < 3; d++) {
      N[d] = Ns[d];
      g[d] = d < c.dim ? c.numGhost : 0;
      NT[d] = N[d] + 2 * g[d];
      n *= NT[d];
    }
  }
  int64_t bulkZones() const { return (int64_t)N[0] * N[1] * N[2]; }
  // dataset dims in PETSc's order: (N3,) (N2,) N1, dof
  std::vector<hsize_t> dims(int dim, int dof) const {
    std::vector<hsize_t> d;
    for (int a = dim - 1; a >= 0; a--) d.push_back((hsize_t)N[a]);
    d.push_back((hsize_t)dof);
    return d;
  }
};

bool validGrid(const grim_grid_config* c, int dof) {
  return c && c->dim >= 1 && c->dim <= 3 && c->N1 > 0 && c->N2 > 0 && c->N3 > 0 && c->numGhost >= 0 &&
         dof > 0 && (c->dim > 1 || c->N2 == 1) && (c->dim > 2 || c->N3 == 1);
}

}  // namespace

extern "C" {

const char* grim_io_last_error(void) { return g_error.c_str(); }

int grim_io_dump(const grim_grid_config* c, int dof, const double* prim, const char* dataset,
                 const char* file) {
  if (!validGrid(c, dof) || !prim || !dataset || !file) return fail("grim_io_dump: bad argument");
  const Hdf5* H = hdf5();
  if (!H) return -2;
  const Shape S(*c);
  // ghosted SoA planes -> bulk, natural order, component fastest
  std::vector<double> buf((size_t)S.bulkZones() * dof);
  size_t o = 0;
  for (int k = S.g[2]; k < S.g[2] + S.N[2]; k++)
    for (int j = S.g[1]; j < S.g[1] + S.N[1]; j++)
      for (int i = S.g[0]; i < S.g[0] + S.N[0]; i++) {
        const int64_t z = i + (int64_t)S.NT[0] * (j + (int64_t)S.NT[1] * k);
        for (int v = 0; v < dof; v++) buf[o++] = prim[(int64_t)v * S.n + z];
      }
  const std::vector<hsize_t> dims = S.dims(c->dim, dof);
  const hid_t f = H->fCreate(file, kAccTrunc, kDefault, kDefault);
  if (f < 0) return fail(std::string("grim_io_dump: cannot create ") + file);
  const hid_t sp = H->sCreateSimple((int)dims.size(), dims.data(), nullptr);
  const hid_t ds = sp < 0 ? -1 : H->dCreate(f, dataset, *H->ieeeF64LE, sp, kDefault, kDefault, kDefault);
  const herr_t w = ds < 0 ? -1 : H->dWrite(ds, *H->nativeDouble, kAll, kAll, kDefault, buf.data());
  if (ds >= 0) H->dClose(ds);
  if (sp >= 0) H->sClose(sp);
  const herr_t fc = H->fClose(f);
  if (ds < 0 || w < 0 || fc < 0) return fail(std::string("grim_io_dump: writing ") + dataset + " to " + file);
  return 0;
}

int grim_io_load(const grim_grid_config* c, int dof, double* prim, const char* dataset, const char* file) {
  if (!validGrid(c, dof) || !prim || !dataset || !file) return fail("grim_io_load: bad argument");
  const Hdf5* H = hdf5();
  if (!H) return -2;
  const Shape S(*c);
  const std::vector<hsize_t> want = S.dims(c->dim, dof);
  const hid_t f = H->fOpen(file, kAccRdonly, kDefault);
  if (f < 0) return fail(std::string("grim_io_load: cannot open ") + file);
  const hid_t ds = H->dOpen(f, dataset, kDefault);
  if (ds < 0) {
    H->fClose(f);
    return fail(std::string("grim_io_load: no dataset ") + dataset + " in " + file);
  }
  const hid_t sp = H->dGetSpace(ds);
  bool shapeOk = sp >= 0 && H->sGetNdims(sp) == (int)want.size();
  if (shapeOk) {
    std::vector<hsize_t> got(want.size());
    shapeOk = H->sGetDims(sp, got.data(), nullptr) == (int)want.size() && got == want;
  }
  std::vector<double> buf((size_t)S.bulkZones() * dof);
  const herr_t r = shapeOk ? H->dRead(ds, *H->nativeDouble, kAll, kAll, kDefault, buf.data()) : -1;
  if (sp >= 0) H->sClose(sp);
  H->dClose(ds);
  H->fClose(f);
  if (!shapeOk) return fail(std::string("grim_io_load: ") + dataset + " in " + file + " does not have the grid's shape");
  if (r < 0) return fail(std::string("grim_io_load: reading ") + dataset + " from " + file);
  size_t o = 0;
  for (int k = S.g[2]; k < S.g[2] + S.N[2]; k++)
    for (int j = S.g[1]; j < S.g[1] + S.N[1]; j++)
      for (int i = S.g[0]; i < S.g[0] + S.N[0]; i++) {
        const int64_t z = i + (int64_t)S.NT[0] * (j + (int64_t)S.NT[1] * k);
        for (int v = 0; v < dof; v++) prim[(int64_t)v * S.n + z] = buf[o++];
      }
  return 0;
}

int grim_io_dump_vts(const grim_grid_config* c, int nvars, const char* const* names, const double* vars,
                     const char* file) {
  if (!validGrid(c, nvars + 3) || !names || !vars || !file) return fail("grim_io_dump_vts: bad argument");
  for (int v = 0; v < nvars; v++)
    if (!names[v] || !names[v][0] || std::string(names[v]).find_first_of("\"<>&") != std::string::npos)
      return fail("grim_io_dump_vts: variable " + std::to_string(v) + " has no usable name");
  const int n1 = c->N1, n2 = c->N2, n3 = c->N3;
  const uint64_t nb = (uint64_t)n1 * n2 * n3;
  const uint64_t bytesVar = nb * sizeof(double), bytesPts = 3 * bytesVar;
  std::string x = "<?xml version=\"1.0\"?>\n"
                  "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                  "header_type=\"UInt64\">\n";
  const std::string extent =
      "0 " + std::to_string(n1 - 1) + " 0 " + std::to_string(n2 - 1) + " 0 " + std::to_string(n3 - 1);
  x += "  <StructuredGrid WholeExtent=\"" + extent + "\">\n    <Piece Extent=\"" + extent + "\">\n";
  x += "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" "
       "offset=\"0\"/>\n      </Points>\n      <PointData>\n";
  uint64_t off = sizeof(uint64_t) + bytesPts;
  for (int v = 0; v < nvars; v++) {
    x += "        <DataArray type=\"Float64\" Name=\"" + std::string(names[v]) +
         "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
    off += sizeof(uint64_t) + bytesVar;
  }
  x += "      </PointData>\n    </Piece>\n  </StructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
  FILE* f = std::fopen(file, "wb");
  if (!f) return fail(std::string("grim_io_dump_vts: cannot create ") + file);
  bool ok = std::fwrite(x.data(), 1, x.size(), f) == x.size();
  // points (x1, x2, x3) interleaved; the x planes follow the nvars planes
  const double* xc = vars + (size_t)nvars * nb;
  ok = ok && std::fwrite(&bytesPts, sizeof bytesPts, 1, f) == 1;
  std::vector<double> pts;
  const uint64_t chunk = 1 << 16;
  for (uint64_t t0 = 0; ok && t0 < nb; t0 += chunk) {
    const uint64_t m = nb - t0 < chunk ? nb - t0 : chunk;
    pts.resize(3 * m);
    for (uint64_t t = 0; t < m; t++)
      for (int d = 0; d < 3; d++) pts[3 * t + d] = d < c->dim ? xc[d * nb + t0 + t] : 0.;
    ok = std::fwrite(pts.data(), sizeof(double), 3 * m, f) == 3 * m;
  }
  for (int v = 0; ok && v < nvars; v++) {
    ok = std::fwrite(&bytesVar, sizeof bytesVar, 1, f) == 1 &&
         std::fwrite(vars + (size_t)v * nb, sizeof(double), nb, f) == nb;
  }
  static const char kEnd[] = "\n  </AppendedData>\n</VTKFile>\n";
  ok = ok && std::fwrite(kEnd, 1, sizeof kEnd - 1, f) == sizeof kEnd - 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(std::string("grim_io_dump_vts: writing ") + file);
  return 0;
}

}  // extern "C"
