// cfdl_cgns.cpp -- CGNS mesh files (HDF5 storage) for the element-list
// ingest of libcfdl_hip.so (host code only; libcfdl_cgns.so, linked against
// HDF5, so that the solver library itself needs no HDF5).
//
// The reference reads its mesh with the CGNS library (src/setup/cell_input.f90:
// 36-101, src/modules/mod_cgns.f90:17-148): base 1, zone 1, the vertex count
// from the zone's sizes, every Elements_t section (name, element type,
// element range, connectivity) in file order, and the coordinate arrays
// CoordinateX/Y/Z.  The CGNS library is not part of this image, so this file
// reads and writes that structure through HDF5 directly, following the CGNS
// HDF5 storage: every node is an HDF5 group carrying the string attributes
// "name" and "label" (33 bytes) and "type" (3 bytes: MT, I4, I8, R4, R8, C1),
// an integer attribute "flags", and its data, if any, in a dataset named
// " data" whose dimensions are the node's in reverse (Fortran) order.
// Children are taken in creation order when the file tracks it, else by name.
//
// HDF5 is opened with dlopen at the first call (CFDL_HDF5_LIB, else
// libhdf5.so.103 / libhdf5.so on the loader's path, else /opt/conda/lib), so
// that loading this library pulls no other C++ runtime into the process.
#include <dlfcn.h>
#include <hdf5.h>

#include <cstdlib>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cfdl.h"
#include "cfdl_cgns.h"
#include "cfdl_internal.h"

namespace {

std::string g_cgns_err;

// H5F_ACC_RDONLY and H5F_ACC_TRUNC without the library-version check their
// header macros call (this file links no HDF5 symbol)
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;

// the HDF5 entry points and type / property-class ids this file uses
struct Hdf5 {
  void *lib = nullptr;
  decltype(&H5open) open;
  decltype(&H5Fopen) Fopen;
  decltype(&H5Fcreate) Fcreate;
  decltype(&H5Fclose) Fclose;
  decltype(&H5Gopen2) Gopen2;
  decltype(&H5Gcreate2) Gcreate2;
  decltype(&H5Gclose) Gclose;
  decltype(&H5Gget_info) Gget_info;
  decltype(&H5Lget_name_by_idx) Lget_name_by_idx;
  decltype(&H5Lexists) Lexists;
  decltype(&H5Aexists) Aexists;
  decltype(&H5Aopen) Aopen;
  decltype(&H5Acreate2) Acreate2;
  decltype(&H5Aread) Aread;
  decltype(&H5Awrite) Awrite;
  decltype(&H5Aclose) Aclose;
  decltype(&H5Aget_type) Aget_type;
  decltype(&H5Tcopy) Tcopy;
  decltype(&H5Tset_size) Tset_size;
  decltype(&H5Tset_strpad) Tset_strpad;
  decltype(&H5Tget_size) Tget_size;
  decltype(&H5Tclose) Tclose;
  decltype(&H5Screate) Screate;
  decltype(&H5Screate_simple) Screate_simple;
  decltype(&H5Sget_simple_extent_npoints) Sget_simple_extent_npoints;
  decltype(&H5Sclose) Sclose;
  decltype(&H5Dopen2) Dopen2;
  decltype(&H5Dcreate2) Dcreate2;
  decltype(&H5Dread) Dread;
  decltype(&H5Dwrite) Dwrite;
  decltype(&H5Dget_space) Dget_space;
  decltype(&H5Dclose) Dclose;
  decltype(&H5Pcreate) Pcreate;
  decltype(&H5Pset_link_creation_order) Pset_link_creation_order;
  decltype(&H5Pclose) Pclose;
  decltype(&H5Eset_auto2) Eset_auto2;
  hid_t int64 = -1, int32 = -1, dbl = -1, flt = -1, chr = -1, c_s1 = -1, i32le = -1,
        f32le = -1, f64le = -1, gcpl = -1, fcpl = -1;
};
Hdf5 H;

template <typename F>
bool sym(F &f, const char *name) {
  f = reinterpret_cast<F>(dlsym(H.lib, name));
  return f != nullptr;
}
bool id_of(hid_t &v, const char *name) {
  const hid_t *p = static_cast<const hid_t *>(dlsym(H.lib, name));
  if (!p) return false;
  v = *p;
  return true;
}

int hdf5_load() {
  if (H.lib) return CFDL_OK;
  const char *names[4] = {std::getenv("CFDL_HDF5_LIB"), "libhdf5.so.103", "libhdf5.so",
                          "/opt/conda/lib/libhdf5.so.103"};
  void *lib = nullptr;
  for (const char *n : names)
    if (n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    g_cgns_err = "CGNS: no HDF5 library (set CFDL_HDF5_LIB to libhdf5.so)";
    return CFDL_ENODEV;
  }
  H.lib = lib;
  bool ok = sym(H.open, "H5open") && sym(H.Fopen, "H5Fopen") && sym(H.Fcreate, "H5Fcreate") &&
            sym(H.Fclose, "H5Fclose") && sym(H.Gopen2, "H5Gopen2") && sym(H.Gcreate2, "H5Gcreate2") &&
            sym(H.Gclose, "H5Gclose") && sym(H.Gget_info, "H5Gget_info") &&
            sym(H.Lget_name_by_idx, "H5Lget_name_by_idx") && sym(H.Lexists, "H5Lexists") &&
            sym(H.Aexists, "H5Aexists") && sym(H.Aopen, "H5Aopen") && sym(H.Acreate2, "H5Acreate2") &&
            sym(H.Aread, "H5Aread") && sym(H.Awrite, "H5Awrite") && sym(H.Aclose, "H5Aclose") &&
            sym(H.Aget_type, "H5Aget_type") && sym(H.Tcopy, "H5Tcopy") &&
            sym(H.Tset_size, "H5Tset_size") && sym(H.Tset_strpad, "H5Tset_strpad") &&
            sym(H.Tget_size, "H5Tget_size") && sym(H.Tclose, "H5Tclose") &&
            sym(H.Screate, "H5Screate") && sym(H.Screate_simple, "H5Screate_simple") &&
            sym(H.Sget_simple_extent_npoints, "H5Sget_simple_extent_npoints") &&
            sym(H.Sclose, "H5Sclose") && sym(H.Dopen2, "H5Dopen2") && sym(H.Dcreate2, "H5Dcreate2") &&
            sym(H.Dread, "H5Dread") && sym(H.Dwrite, "H5Dwrite") && sym(H.Dget_space, "H5Dget_space") &&
            sym(H.Dclose, "H5Dclose") && sym(H.Pcreate, "H5Pcreate") &&
            sym(H.Pset_link_creation_order, "H5Pset_link_creation_order") &&
            sym(H.Pclose, "H5Pclose") && sym(H.Eset_auto2, "H5Eset_auto2");
  // the predefined ids are set by H5open
  ok = ok && H.open() >= 0 && id_of(H.int64, "H5T_NATIVE_INT64_g") &&
       id_of(H.int32, "H5T_NATIVE_INT32_g") && id_of(H.dbl, "H5T_NATIVE_DOUBLE_g") &&
       id_of(H.flt, "H5T_NATIVE_FLOAT_g") && id_of(H.chr, "H5T_NATIVE_SCHAR_g") &&
       id_of(H.c_s1, "H5T_C_S1_g") && id_of(H.i32le, "H5T_STD_I32LE_g") &&
       id_of(H.f32le, "H5T_IEEE_F32LE_g") && id_of(H.f64le, "H5T_IEEE_F64LE_g") &&
       id_of(H.gcpl, "H5P_CLS_GROUP_CREATE_ID_g") && id_of(H.fcpl, "H5P_CLS_FILE_CREATE_ID_g");
  if (!ok) {
    dlclose(lib);
    H = Hdf5{};
    g_cgns_err = "CGNS: the HDF5 library lacks an entry point this reader needs";
    return CFDL_ENODEV;
  }
  H.Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as status codes
  return CFDL_OK;
}

int fail(const std::string &msg) {
  g_cgns_err = msg;
  return CFDL_EINVAL;
}

// HDF5 identifiers closed on scope exit
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  operator hid_t() const { return id; }
  bool ok() const { return id >= 0; }
};

// a string attribute of a node ("" if absent)
std::string str_attr(hid_t g, const char *name) {
  if (H.Aexists(g, name) <= 0) return "";
  Hid a(H.Aopen(g, name, H5P_DEFAULT), H.Aclose);
  if (!a.ok()) return "";
  Hid t(H.Aget_type(a), H.Tclose);
  const size_t n = H.Tget_size(t);
  std::vector<char> buf(n + 1, 0);
  Hid mt(H.Tcopy(H.c_s1), H.Tclose);
  H.Tset_size(mt, n);
  if (H.Aread(a, mt, buf.data()) < 0) return "";
  return std::string(buf.data(), strnlen(buf.data(), n));
}

// children of a group, in creation order if tracked, else by name
std::vector<std::string> children(hid_t g) {
  std::vector<std::string> out;
  H5G_info_t info;
  if (H.Gget_info(g, &info) < 0) return out;
  for (int pass = 0; pass < 2 && out.empty(); pass++) {
    const H5_index_t idx = pass == 0 ? H5_INDEX_CRT_ORDER : H5_INDEX_NAME;
    for (hsize_t i = 0; i < info.nlinks; i++) {
      const ssize_t n = H.Lget_name_by_idx(g, ".", idx, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
      if (n < 0) {
        out.clear();
        break;
      }
      std::vector<char> buf(static_cast<size_t>(n) + 1, 0);
      H.Lget_name_by_idx(g, ".", idx, H5_ITER_INC, i, buf.data(), buf.size(), H5P_DEFAULT);
      out.emplace_back(buf.data());
    }
  }
  // node data is a dataset, not a child node
  out.erase(std::remove_if(out.begin(), out.end(), [](const std::string &s) {
              return !s.empty() && s[0] == ' ';
            }),
            out.end());
  return out;
}

// the first child (in the order above) carrying the label, or ""
std::string child_with_label(hid_t g, const char *label) {
  for (const std::string &c : children(g)) {
    Hid cg(H.Gopen2(g, c.c_str(), H5P_DEFAULT), H.Gclose);
    if (cg.ok() && str_attr(cg, "label") == label) return c;
  }
  return "";
}

// a node's data as int64 or double values
bool read_ints(hid_t node, std::vector<int64_t> &v) {
  if (H.Lexists(node, " data", H5P_DEFAULT) <= 0) return false;
  Hid d(H.Dopen2(node, " data", H5P_DEFAULT), H.Dclose);
  Hid s(H.Dget_space(d), H.Sclose);
  const hssize_t n = H.Sget_simple_extent_npoints(s);
  if (n < 0) return false;
  v.assign(static_cast<size_t>(n), 0);
  return n == 0 || H.Dread(d, H.int64, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) >= 0;
}
bool read_doubles(hid_t node, std::vector<double> &v) {
  if (H.Lexists(node, " data", H5P_DEFAULT) <= 0) return false;
  Hid d(H.Dopen2(node, " data", H5P_DEFAULT), H.Dclose);
  Hid s(H.Dget_space(d), H.Sclose);
  const hssize_t n = H.Sget_simple_extent_npoints(s);
  if (n < 0) return false;
  v.assign(static_cast<size_t>(n), 0.0);
  return n == 0 || H.Dread(d, H.dbl, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) >= 0;
}

// nodes per element of the CGNS element types the reference's meshes use
// (ElementType_t: QUAD_4 = 7, HEXA_8 = 17; TRI_3 = 5, TETRA_4 = 10,
// PYRA_5 = 12, PENTA_6 = 14 also fit the 8 vertex slots per element)
int nodes_per_element(int64_t t) {
  switch (t) {
    case 5: return 3;
    case 7: return 4;
    case 10: return 4;
    case 12: return 5;
    case 14: return 6;
    case 17: return 8;
    default: return 0;
  }
}

// ---- writing ----
herr_t put_str_attr(hid_t g, const char *name, const std::string &v, size_t len) {
  Hid t(H.Tcopy(H.c_s1), H.Tclose);
  H.Tset_size(t, len);
  H.Tset_strpad(t, H5T_STR_NULLTERM);
  Hid s(H.Screate(H5S_SCALAR), H.Sclose);
  Hid a(H.Acreate2(g, name, t, s, H5P_DEFAULT, H5P_DEFAULT), H.Aclose);
  std::vector<char> buf(len, 0);
  std::memcpy(buf.data(), v.data(), std::min(v.size(), len - 1));
  return H.Awrite(a, t, buf.data());
}

// a node: group with name/label/type/flags; creation order tracked
hid_t new_node(hid_t parent, const std::string &name, const std::string &label,
               const std::string &type) {
  Hid pl(H.Pcreate(H.gcpl), H.Pclose);
  H.Pset_link_creation_order(pl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  const hid_t g = H.Gcreate2(parent, name.c_str(), H5P_DEFAULT, pl, H5P_DEFAULT);
  if (g < 0) return g;
  put_str_attr(g, "name", name, 33);
  put_str_attr(g, "label", label, 33);
  put_str_attr(g, "type", type, 3);
  Hid s(H.Screate(H5S_SCALAR), H.Sclose);
  Hid a(H.Acreate2(g, "flags", H.int32, s, H5P_DEFAULT, H5P_DEFAULT), H.Aclose);
  const int32_t one = 1;
  H.Awrite(a, H.int32, &one);
  return g;
}

// node data with CGNS dimensions dims (stored reversed)
herr_t put_data(hid_t g, hid_t file_type, hid_t mem_type, const std::vector<hsize_t> &dims,
                const void *v) {
  std::vector<hsize_t> rd(dims.rbegin(), dims.rend());
  Hid s(H.Screate_simple(static_cast<int>(rd.size()), rd.data(), nullptr), H.Sclose);
  Hid d(H.Dcreate2(g, " data", file_type, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H.Dclose);
  if (!d.ok()) return -1;
  return H.Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
}

}  // namespace

extern "C" {

const char *cfdl_cgns_last_error(void) { return g_cgns_err.c_str(); }

int cfdl_cgns_section_name(const cfdl_mesh *m, int k, char *buf, int cap) {
  if (!m || !buf || cap < 1 || k < 0 || k >= static_cast<int>(m->secname.size()))
    return fail("cfdl_cgns_section_name: no such section");
  const std::string &s = m->secname[k];
  const size_t n = std::min(s.size(), static_cast<size_t>(cap - 1));
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return CFDL_OK;
}

int cfdl_mesh_read_cgns(const char *path, cfdl_mesh **out) {
  if (!path || !out) return fail("cfdl_mesh_read_cgns: null argument");
  *out = nullptr;
  if (const int rc = hdf5_load()) return rc;
  H.Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Hid f(H.Fopen(path, kAccRdonly, H5P_DEFAULT), H.Fclose);
  if (!f.ok()) return fail(std::string("cfdl_mesh_read_cgns: cannot open ") + path + " as HDF5");
  Hid root(H.Gopen2(f, "/", H5P_DEFAULT), H.Gclose);
  const std::string bname = child_with_label(root, "CGNSBase_t");
  if (bname.empty()) return fail("cfdl_mesh_read_cgns: no CGNSBase_t node");
  Hid base(H.Gopen2(root, bname.c_str(), H5P_DEFAULT), H.Gclose);
  const std::string zname = child_with_label(base, "Zone_t");
  if (zname.empty()) return fail("cfdl_mesh_read_cgns: no Zone_t node in base " + bname);
  Hid zone(H.Gopen2(base, zname.c_str(), H5P_DEFAULT), H.Gclose);
  // unstructured zone sizes: vertices, cells, boundary vertices
  std::vector<int64_t> zs;
  if (!read_ints(zone, zs) || zs.empty() || zs[0] <= 0)
    return fail("cfdl_mesh_read_cgns: zone " + zname + " has no vertex count");
  const std::string tname = child_with_label(zone, "ZoneType_t");
  if (!tname.empty()) {
    Hid zt(H.Gopen2(zone, tname.c_str(), H5P_DEFAULT), H.Gclose);
    Hid d(H.Dopen2(zt, " data", H5P_DEFAULT), H.Dclose);
    if (d.ok()) {
      Hid s(H.Dget_space(d), H.Sclose);
      std::vector<char> buf(static_cast<size_t>(H.Sget_simple_extent_npoints(s)) + 1, 0);
      H.Dread(d, H.chr, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
      if (std::string(buf.data()).rfind("Unstructured", 0) != 0)
        return fail("cfdl_mesh_read_cgns: zone " + zname + " is not unstructured");
    }
  }
  const int64_t nvx = zs[0];
  // coordinates
  const std::string gname = child_with_label(zone, "GridCoordinates_t");
  if (gname.empty()) return fail("cfdl_mesh_read_cgns: no GridCoordinates_t node");
  Hid grid(H.Gopen2(zone, gname.c_str(), H5P_DEFAULT), H.Gclose);
  std::vector<double> xyz[3];
  const char *cn[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
  for (int i = 0; i < 3; i++) {
    if (H.Lexists(grid, cn[i], H5P_DEFAULT) <= 0)
      return fail(std::string("cfdl_mesh_read_cgns: no ") + cn[i]);
    Hid c(H.Gopen2(grid, cn[i], H5P_DEFAULT), H.Gclose);
    if (!read_doubles(c, xyz[i]) || static_cast<int64_t>(xyz[i].size()) != nvx)
      return fail(std::string("cfdl_mesh_read_cgns: ") + cn[i] + " does not hold one value per vertex");
  }
  // sections in file order
  struct Sec {
    std::string name;
    int type;
    int64_t start, end;
    std::vector<int64_t> conn;
  };
  std::vector<Sec> secs;
  for (const std::string &c : children(zone)) {
    Hid e(H.Gopen2(zone, c.c_str(), H5P_DEFAULT), H.Gclose);
    if (!e.ok() || str_attr(e, "label") != "Elements_t") continue;
    Sec s;
    s.name = str_attr(e, "name");
    if (s.name.empty()) s.name = c;
    std::vector<int64_t> head, range;
    if (!read_ints(e, head) || head.empty())
      return fail("cfdl_mesh_read_cgns: section " + c + " has no element type");
    s.type = static_cast<int>(head[0]);
    if (nodes_per_element(s.type) == 0)
      return fail("cfdl_mesh_read_cgns: section " + c + " has element type " +
                  std::to_string(s.type) + ", which has no fixed vertex count of at most 8");
    if (H.Lexists(e, "ElementRange", H5P_DEFAULT) <= 0 ||
        H.Lexists(e, "ElementConnectivity", H5P_DEFAULT) <= 0)
      return fail("cfdl_mesh_read_cgns: section " + c + " lacks ElementRange or ElementConnectivity");
    {
      Hid r(H.Gopen2(e, "ElementRange", H5P_DEFAULT), H.Gclose);
      if (!read_ints(r, range) || range.size() != 2)
        return fail("cfdl_mesh_read_cgns: section " + c + " has a bad ElementRange");
    }
    s.start = range[0];
    s.end = range[1];
    {
      Hid k(H.Gopen2(e, "ElementConnectivity", H5P_DEFAULT), H.Gclose);
      if (!read_ints(k, s.conn)) return fail("cfdl_mesh_read_cgns: section " + c + " unreadable");
    }
    const int64_t n = s.end - s.start + 1;
    if (n <= 0 || static_cast<int64_t>(s.conn.size()) != n * nodes_per_element(s.type))
      return fail("cfdl_mesh_read_cgns: section " + c + " connectivity does not match its range");
    secs.push_back(std::move(s));
  }
  if (secs.empty()) return fail("cfdl_mesh_read_cgns: no Elements_t section");
  // the element-list input: sections by element range, 8 vertex slots each
  std::stable_sort(secs.begin(), secs.end(), [](const Sec &a, const Sec &b) { return a.start < b.start; });
  int64_t next = 1;
  for (const Sec &s : secs) {
    if (s.start != next) return fail("cfdl_mesh_read_cgns: element ranges are not contiguous from 1");
    next = s.end + 1;
  }
  const int64_t nelem = next - 1;
  if (nelem > (int64_t{1} << 30) || nvx > (int64_t{1} << 30))
    return fail("cfdl_mesh_read_cgns: mesh too large");
  const int nsec = static_cast<int>(secs.size());
  std::vector<int> etype(nsec), esec(2 * nsec), e2vx(8 * static_cast<size_t>(nelem), 0);
  std::vector<const char *> names(nsec);
  for (int k = 0; k < nsec; k++) {
    const Sec &s = secs[k];
    etype[k] = s.type;
    esec[2 * k] = static_cast<int>(s.start);
    esec[2 * k + 1] = static_cast<int>(s.end);
    const int npe = nodes_per_element(s.type);
    for (int64_t e = s.start; e <= s.end; e++)
      for (int j = 0; j < npe; j++) {
        const int64_t v = s.conn[static_cast<size_t>((e - s.start) * npe + j)];
        if (v < 1 || v > nvx) return fail("cfdl_mesh_read_cgns: vertex id out of range in " + s.name);
        e2vx[static_cast<size_t>(8 * (e - 1) + j)] = static_cast<int>(v);
      }
    names[k] = s.name.c_str();
  }
  cfdl_mesh *m = nullptr;
  int rc = cfdl_mesh_from_elements(static_cast<int>(nvx), xyz[0].data(), xyz[1].data(),
                                   xyz[2].data(), nsec, etype.data(), esec.data(), e2vx.data(), &m);
  if (rc != CFDL_OK) return fail(std::string("cfdl_mesh_read_cgns: ") + cfdl_last_error());
  rc = cfdl_mesh_set_section_names(m, nsec, names.data());
  if (rc != CFDL_OK) {
    cfdl_mesh_free(m);
    return fail(std::string("cfdl_mesh_read_cgns: ") + cfdl_last_error());
  }
  *out = m;
  return CFDL_OK;
}

int cfdl_mesh_write_cgns(const cfdl_mesh *m, const char *path) {
  if (!m || !path) return fail("cfdl_mesh_write_cgns: null argument");
  if (const int rc = hdf5_load()) return rc;
  if (m->nsec < 1 || static_cast<int>(m->secname.size()) < m->nsec)
    return fail("cfdl_mesh_write_cgns: the mesh's sections have no names");
  H.Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Hid fcpl(H.Pcreate(H.fcpl), H.Pclose);
  H.Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  Hid f(H.Fcreate(path, kAccTrunc, fcpl, H5P_DEFAULT), H.Fclose);
  if (!f.ok()) return fail(std::string("cfdl_mesh_write_cgns: cannot create ") + path);
  Hid root(H.Gopen2(f, "/", H5P_DEFAULT), H.Gclose);
  put_str_attr(root, "name", "HDF5 MotherNode", 33);
  put_str_attr(root, "label", "Root Node of HDF5 File", 33);
  put_str_attr(root, "type", "MT", 3);
  {
    const char fmt[] = "IEEE_LITTLE_32";
    std::vector<hsize_t> d{sizeof(fmt) - 1};
    Hid s(H.Screate_simple(1, d.data(), nullptr), H.Sclose);
    Hid ds(H.Dcreate2(root, " format", H.chr, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
           H.Dclose);
    H.Dwrite(ds, H.chr, H5S_ALL, H5S_ALL, H5P_DEFAULT, fmt);
  }
  {
    Hid v(new_node(root, "CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4"), H.Gclose);
    const float ver = 4.2f;
    put_data(v, H.f32le, H.flt, {1}, &ver);
  }
  Hid base(new_node(root, "Base", "CGNSBase_t", "I4"), H.Gclose);
  const int32_t dims[2] = {3, 3};
  put_data(base, H.i32le, H.int32, {2}, dims);
  Hid zone(new_node(base, "Zone", "Zone_t", "I4"), H.Gclose);
  const int32_t zsz[3] = {m->nvx, m->ne, 0};
  put_data(zone, H.i32le, H.int32, {1, 3}, zsz);
  {
    Hid zt(new_node(zone, "ZoneType", "ZoneType_t", "C1"), H.Gclose);
    const char un[] = "Unstructured";
    put_data(zt, H.chr, H.chr, {sizeof(un) - 1}, un);
  }
  {
    Hid g(new_node(zone, "GridCoordinates", "GridCoordinates_t", "MT"), H.Gclose);
    const std::vector<double> *c[3] = {&m->x, &m->y, &m->z};
    const char *cn[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
    for (int i = 0; i < 3; i++) {
      Hid a(new_node(g, cn[i], "DataArray_t", "R8"), H.Gclose);
      if (put_data(a, H.f64le, H.dbl, {static_cast<hsize_t>(m->nvx)},
                   c[i]->data()) < 0)
        return fail("cfdl_mesh_write_cgns: writing coordinates failed");
    }
  }
  for (int k = 0; k < m->nsec; k++) {
    const int t = m->etype[k], npe = nodes_per_element(t);
    if (npe == 0) return fail("cfdl_mesh_write_cgns: unsupported element type " + std::to_string(t));
    const int s0 = m->esec[2 * k], s1 = m->esec[2 * k + 1];
    Hid e(new_node(zone, m->secname[k], "Elements_t", "I4"), H.Gclose);
    if (!e.ok()) return fail("cfdl_mesh_write_cgns: cannot create section " + m->secname[k]);
    const int32_t head[2] = {t, 0};
    put_data(e, H.i32le, H.int32, {2}, head);
    {
      Hid r(new_node(e, "ElementRange", "IndexRange_t", "I4"), H.Gclose);
      const int32_t rg[2] = {s0, s1};
      put_data(r, H.i32le, H.int32, {2}, rg);
    }
    std::vector<int32_t> conn;
    conn.reserve(static_cast<size_t>(s1 - s0 + 1) * npe);
    for (int el = s0; el <= s1; el++)
      for (int j = 0; j < npe; j++) conn.push_back(m->e2vx[static_cast<size_t>(8 * (el - 1) + j)]);
    Hid cc(new_node(e, "ElementConnectivity", "DataArray_t", "I4"), H.Gclose);
    if (put_data(cc, H.i32le, H.int32, {conn.size()}, conn.data()) < 0)
      return fail("cfdl_mesh_write_cgns: writing connectivity failed");
  }
  return CFDL_OK;
}

}  // extern "C"
