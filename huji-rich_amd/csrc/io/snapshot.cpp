// HDF5 snapshots in RICH's 2D layout through the HDF5 C API (see
// include/richio.h for the layout and the reference lines it follows).
#include "../../../include/richio.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

// HDF5 handle closed on scope exit.
struct H {
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  H(const H&) = delete;
  H& operator=(const H&) = delete;
  bool ok() const { return id >= 0; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// write_std_vector_to_hdf5 (hdf5_utils.hpp:29-51): 1D, chunk min(n, 1e5),
// deflate 6, the native type with little-endian order.
int write_vec(hid_t where, const char* name, const void* data, size_t n, bool integer) {
  if (n == 0) return RIO_E_ARG;
  hsize_t dims[1] = {static_cast<hsize_t>(n)};
  H space(H5Screate_simple(1, dims, nullptr), H5Sclose);
  H plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  H type(H5Tcopy(integer ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE), H5Tclose);
  if (!space.ok() || !plist.ok() || !type.ok()) return RIO_E_HDF5;
  if (H5Tset_order(type.id, H5T_ORDER_LE) < 0) return RIO_E_HDF5;
  hsize_t chunk[1] = {std::min<hsize_t>(dims[0], 100000)};
  if (H5Pset_chunk(plist.id, 1, chunk) < 0 || H5Pset_deflate(plist.id, 6) < 0) return RIO_E_HDF5;
  H ds(H5Dcreate2(where, name, type.id, space.id, H5P_DEFAULT, plist.id, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return RIO_E_HDF5;
  if (H5Dwrite(ds.id, type.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return RIO_E_HDF5;
  return RIO_OK;
}

// Length of a 1D dataset, or -1.
int64_t vec_len(hid_t where, const char* name) {
  H ds(H5Dopen2(where, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return -1;
  H space(H5Dget_space(ds.id), H5Sclose);
  if (!space.ok() || H5Sget_simple_extent_ndims(space.id) != 1) return -1;
  hsize_t d[1];
  H5Sget_simple_extent_dims(space.id, d, nullptr);
  return static_cast<int64_t>(d[0]);
}

int read_vec(hid_t where, const char* name, void* out, size_t n, bool integer) {
  if (vec_len(where, name) != static_cast<int64_t>(n)) return RIO_E_LAYOUT;
  if (!out) return RIO_OK;
  H ds(H5Dopen2(where, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return RIO_E_LAYOUT;
  if (H5Dread(ds.id, integer ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
              out) < 0)
    return RIO_E_HDF5;
  return RIO_OK;
}

bool mesh_ok(const rio_mesh* m) {
  return m && m->n_cells > 0 && m->pt_x && m->pt_y && m->cm_x && m->cm_y && m->tess_volume &&
         m->volume && m->csr_off && m->csr_face && m->v1x && m->v1y && m->v2x && m->v2y;
}

// ConvexHull (ConvexHull.cpp:30-58) of cell i: the vertices of its faces in
// face order, a vertex dropped when within squared distance 1e-14 R^2 of
// one already taken (R = GetWidth), then sorted by the angle about the CM.
void cell_hull(const rio_mesh* m, int32_t i, std::vector<double>& xs, std::vector<double>& ys) {
  const int32_t b = m->csr_off[i], e = m->csr_off[i + 1];
  xs.clear();
  ys.clear();
  if (e <= b) return;
  const double R = std::sqrt(m->tess_volume[i] / M_PI);
  const double tol = 1e-14 * std::pow(R, 2);
  auto add = [&](double px, double py, bool check) {
    if (check)
      for (size_t k = 0; k < xs.size(); ++k) {
        const double dx = xs[k] - px, dy = ys[k] - py;
        if (dx * dx + dy * dy < tol) return;
      }
    xs.push_back(px);
    ys.push_back(py);
  };
  const int32_t f0 = m->csr_face[b];
  add(m->v1x[f0], m->v1y[f0], false);
  add(m->v2x[f0], m->v2y[f0], false);
  for (int32_t k = b + 1; k < e; ++k) {
    const int32_t f = m->csr_face[k];
    add(m->v1x[f], m->v1y[f], true);
    add(m->v2x[f], m->v2y[f], true);
  }
  const double cx = m->cm_x[i], cy = m->cm_y[i];
  std::vector<double> ang(xs.size());
  for (size_t k = 0; k < xs.size(); ++k) ang[k] = std::atan2(ys[k] - cy, xs[k] - cx);
  std::vector<size_t> idx(xs.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return ang[a] < ang[c]; });
  std::vector<double> sx(xs.size()), sy(ys.size());
  for (size_t k = 0; k < idx.size(); ++k) {
    sx[k] = xs[idx[k]];
    sy[k] = ys[idx[k]];
  }
  xs.swap(sx);
  ys.swap(sy);
}

int hulls(const rio_mesh* m, std::vector<double>& n, std::vector<double>& xv,
          std::vector<double>& yv) {
  const int32_t N = m->n_cells;
  n.assign(static_cast<size_t>(N), 0.0);
  xv.clear();
  yv.clear();
  xv.reserve(7 * static_cast<size_t>(N));
  yv.reserve(7 * static_cast<size_t>(N));
  std::vector<double> xs, ys;
  for (int32_t i = 0; i < N; ++i) {
    cell_hull(m, i, xs, ys);
    xv.insert(xv.end(), xs.begin(), xs.end());
    yv.insert(yv.end(), ys.begin(), ys.end());
    n[i] = static_cast<double>(static_cast<int>(xs.size()));
  }
  return RIO_OK;
}

struct Listing {
  std::vector<std::string> lines;
};

herr_t visit(hid_t obj, const char* name, const H5O_info_t* info, void* op) {
  if (info->type != H5O_TYPE_DATASET) return 0;
  H ds(H5Dopen2(obj, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return -1;
  H type(H5Dget_type(ds.id), H5Tclose);
  H space(H5Dget_space(ds.id), H5Sclose);
  H plist(H5Dget_create_plist(ds.id), H5Pclose);
  const H5T_class_t cls = H5Tget_class(type.id);
  const size_t size = H5Tget_size(type.id);
  hsize_t d[1] = {0};
  if (H5Sget_simple_extent_ndims(space.id) == 1) H5Sget_simple_extent_dims(space.id, d, nullptr);
  hsize_t chunk[1] = {0};
  if (H5Pget_layout(plist.id) == H5D_CHUNKED) H5Pget_chunk(plist.id, 1, chunk);
  int deflate = -1;
  const int nf = H5Pget_nfilters(plist.id);
  for (int k = 0; k < nf; ++k) {
    unsigned flags = 0, cd[4] = {0, 0, 0, 0};
    size_t ncd = 4;
    const H5Z_filter_t f = H5Pget_filter2(plist.id, static_cast<unsigned>(k), &flags, &ncd, cd, 0,
                                          nullptr, nullptr);
    if (f == H5Z_FILTER_DEFLATE) deflate = ncd > 0 ? static_cast<int>(cd[0]) : 0;
  }
  char line[512];
  const bool le = H5Tget_order(type.id) == H5T_ORDER_LE;
  std::snprintf(line, sizeof(line), "/%s %c%s %zu %llu %llu %d\n", name,
                cls == H5T_FLOAT ? 'f' : (cls == H5T_INTEGER ? 'i' : '?'), le ? "" : "BE", size,
                static_cast<unsigned long long>(d[0]), static_cast<unsigned long long>(chunk[0]),
                deflate);
  static_cast<Listing*>(op)->lines.push_back(line);
  return 0;
}

}  // namespace

extern "C" int rio_cell_vertices(const rio_mesh* m, int64_t* total, double* n, double* xv,
                                 double* yv, int64_t cap) {
  if (!mesh_ok(m) || !total) return RIO_E_ARG;
  std::vector<double> cn, x, y;
  hulls(m, cn, x, y);
  *total = static_cast<int64_t>(x.size());
  if (n) std::copy(cn.begin(), cn.end(), n);
  if (xv || yv) {
    if (cap < static_cast<int64_t>(x.size())) return RIO_E_SIZE;
    if (xv) std::copy(x.begin(), x.end(), xv);
    if (yv) std::copy(y.begin(), y.end(), yv);
  }
  return RIO_OK;
}

extern "C" int rio_write_snapshot(const char* path, const rio_mesh* m, const double* density,
                                  const double* pressure, const double* x_velocity,
                                  const double* y_velocity, double time, int32_t cycle) {
  if (!path || !mesh_ok(m) || !density || !pressure || !x_velocity || !y_velocity)
    return RIO_E_ARG;
  quiet();
  const size_t N = static_cast<size_t>(m->n_cells);
  std::vector<double> nv, xv, yv;
  hulls(m, nv, xv, yv);
  H file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return RIO_E_HDF5;
  H geometry(H5Gcreate2(file.id, "/geometry", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  H appendices(H5Gcreate2(file.id, "/appendices", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               H5Gclose);
  H hydro(H5Gcreate2(file.id, "/hydrodynamic", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  H tracers(H5Gcreate2(file.id, "/tracers", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  H stickers(H5Gcreate2(file.id, "/stickers", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!geometry.ok() || !appendices.ok() || !hydro.ok() || !tracers.ok() || !stickers.ok())
    return RIO_E_HDF5;
  const int cyc = static_cast<int>(cycle);
  int rc = RIO_OK;
  auto put = [&](hid_t where, const char* name, const void* data, size_t n, bool integer) {
    if (rc == RIO_OK) rc = write_vec(where, name, data, n, integer);
  };
  put(file.id, "time", &time, 1, false);
  put(file.id, "cycle", &cyc, 1, true);
  put(file.id, "Volumes", m->volume, N, false);
  put(geometry.id, "x_coordinate", m->pt_x, N, false);
  put(geometry.id, "y_coordinate", m->pt_y, N, false);
  put(geometry.id, "x_vertices", xv.data(), xv.size(), false);
  put(geometry.id, "y_vertices", yv.data(), yv.size(), false);
  put(geometry.id, "n_vertices", nv.data(), N, false);
  put(hydro.id, "density", density, N, false);
  put(hydro.id, "pressure", pressure, N, false);
  put(hydro.id, "x_velocity", x_velocity, N, false);
  put(hydro.id, "y_velocity", y_velocity, N, false);
  return rc;
}

extern "C" int rio_snapshot_cells(const char* path, int32_t* n_cells) {
  if (!path || !n_cells) return RIO_E_ARG;
  quiet();
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return RIO_E_HDF5;
  H hydro(H5Gopen2(file.id, "hydrodynamic", H5P_DEFAULT), H5Gclose);
  if (!hydro.ok()) return RIO_E_LAYOUT;
  const int64_t n = vec_len(hydro.id, "density");
  if (n < 0 || n > INT32_MAX) return RIO_E_LAYOUT;
  *n_cells = static_cast<int32_t>(n);
  return RIO_OK;
}

extern "C" int rio_read_snapshot(const char* path, int32_t n_cells, double* x, double* y,
                                 double* density, double* pressure, double* x_velocity,
                                 double* y_velocity, double* time, int32_t* cycle) {
  if (!path || n_cells <= 0) return RIO_E_ARG;
  int32_t n = 0;
  int rc = rio_snapshot_cells(path, &n);
  if (rc != RIO_OK) return rc;
  if (n != n_cells) return RIO_E_SIZE;
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return RIO_E_HDF5;
  H geometry(H5Gopen2(file.id, "geometry", H5P_DEFAULT), H5Gclose);
  H hydro(H5Gopen2(file.id, "hydrodynamic", H5P_DEFAULT), H5Gclose);
  if (!geometry.ok() || !hydro.ok()) return RIO_E_LAYOUT;
  const size_t N = static_cast<size_t>(n_cells);
  auto get = [&](hid_t where, const char* name, void* out, size_t len, bool integer) {
    if (rc == RIO_OK) rc = read_vec(where, name, out, len, integer);
  };
  get(geometry.id, "x_coordinate", x, N, false);
  get(geometry.id, "y_coordinate", y, N, false);
  get(hydro.id, "density", density, N, false);
  get(hydro.id, "pressure", pressure, N, false);
  get(hydro.id, "x_velocity", x_velocity, N, false);
  get(hydro.id, "y_velocity", y_velocity, N, false);
  double t = 0;
  int c = 0;
  get(file.id, "time", &t, 1, false);
  get(file.id, "cycle", &c, 1, true);
  if (rc != RIO_OK) return rc;
  if (time) *time = t;
  if (cycle) *cycle = static_cast<int32_t>(c);
  return RIO_OK;
}

extern "C" int rio_describe(const char* path, char* buf, int32_t cap) {
  if (!path || !buf || cap <= 0) return RIO_E_ARG;
  quiet();
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return RIO_E_HDF5;
  Listing l;
  if (H5Ovisit(file.id, H5_INDEX_NAME, H5_ITER_INC, visit, &l) < 0) return RIO_E_HDF5;
  std::string all;
  for (const std::string& s : l.lines) all += s;
  if (all.size() + 1 > static_cast<size_t>(cap)) return RIO_E_SIZE;
  std::memcpy(buf, all.c_str(), all.size() + 1);
  return RIO_OK;
}
