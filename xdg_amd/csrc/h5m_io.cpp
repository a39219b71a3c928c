// h5m_io.cpp -- DAGMC/MOAB .h5m files <-> flat geometry arrays
// (include/xdg_h5m.h), host code over the HDF5 C API.
//
// Reads the entity sets a DAGMC model is made of, as MOABMeshManager::init
// and parse_metadata do through MOAB (src/moab/mesh_manager.cpp:37-107,
// 489-590): volume and surface sets by GEOM_DIMENSION, their GLOBAL_IDs,
// the triangles in each surface set, the tetrahedra in each volume set, the
// GEOM_SENSE_2 parent volumes of each surface, and group sets (CATEGORY
// "Group") with their NAME and member sets.
#include <hdf5.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "xdg_h5m.h"

namespace {

thread_local std::string g_err;

struct FormatError {
  std::string msg;
};
struct IoError {
  std::string msg;
};

// MOAB entity types stored with each element group (moab/EntityType.hpp)
constexpr int32_t MB_TRI = 2;
constexpr int32_t MB_TET = 5;
// set flags (mhdf): contents stored as (start, count) ranges
constexpr int64_t SET_RANGE_BIT = 0x8;

// ---- small RAII handles ----------------------------------------------------
struct H {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H() = default;
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  H(const H&) = delete;
  H& operator=(const H&) = delete;
  H(H&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
  ~H()
  {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

H check(hid_t id, herr_t (*close)(hid_t), const std::string& what)
{
  if (id < 0) throw IoError{"HDF5: " + what};
  return H(id, close);
}
void check(herr_t rc, const std::string& what)
{
  if (rc < 0) throw IoError{"HDF5: " + what};
}

H open_group(hid_t loc, const char* name) { return check(H5Gopen2(loc, name, H5P_DEFAULT), H5Gclose, std::string("open group ") + name); }
H make_group(hid_t loc, const char* name)
{
  return check(H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, std::string("create group ") + name);
}
bool exists(hid_t loc, const char* path) { return H5Lexists(loc, path, H5P_DEFAULT) > 0; }

// ---- writing ----------------------------------------------------------------
void write_attr_u64(hid_t obj, const char* name, uint64_t v)
{
  H sp = check(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
  H a = check(H5Acreate2(obj, name, H5T_STD_U64LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, std::string("attr ") + name);
  check(H5Awrite(a, H5T_NATIVE_UINT64, &v), std::string("write attr ") + name);
}
void write_attr_i32(hid_t obj, const char* name, int32_t v)
{
  H sp = check(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
  H a = check(H5Acreate2(obj, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, std::string("attr ") + name);
  check(H5Awrite(a, H5T_NATIVE_INT32, &v), std::string("write attr ") + name);
}

// rows x cols dataset (cols == 0: one dimension)
H write_data(hid_t loc, const char* name, hid_t file_type, hid_t mem_type, int64_t rows, int cols, const void* data)
{
  hsize_t dims[2] = {(hsize_t)rows, (hsize_t)cols};
  H sp = check(H5Screate_simple(cols ? 2 : 1, dims, nullptr), H5Sclose, "data space");
  H d = check(H5Dcreate2(loc, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
              std::string("create ") + name);
  if (rows > 0) check(H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), std::string("write ") + name);
  return d;
}

// a sparse tag: /tstt/tags/<name>/{type, id_list, values}
void write_tag(hid_t tags, const char* name, hid_t value_type, const std::vector<uint64_t>& ids, const void* values)
{
  H g = make_group(tags, name);
  write_attr_i32(g, "class", 1);  // sparse
  H t = check(H5Tcopy(value_type), H5Tclose, "copy tag type");
  check(H5Tcommit2(g, "type", t, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "commit tag type");
  write_data(g, "id_list", H5T_STD_U64LE, H5T_NATIVE_UINT64, (int64_t)ids.size(), 0, ids.data());
  write_data(g, "values", t, t, (int64_t)ids.size(), 0, values);
}

// ---- reading ----------------------------------------------------------------
struct Data {
  std::vector<hsize_t> dims;
  std::vector<char> bytes;
  size_t elem = 0;  // bytes per element of the memory type
};

Data read_data(hid_t loc, const char* name, hid_t mem_type)
{
  H d = check(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose, std::string("open ") + name);
  H sp = check(H5Dget_space(d), H5Sclose, "data space");
  const int nd = H5Sget_simple_extent_ndims(sp);
  if (nd < 1 || nd > 2) throw FormatError{std::string(name) + ": expected 1 or 2 dimensions"};
  Data r;
  r.dims.resize((size_t)nd);
  H5Sget_simple_extent_dims(sp, r.dims.data(), nullptr);
  size_t n = 1;
  for (hsize_t x : r.dims) n *= (size_t)x;
  r.elem = H5Tget_size(mem_type);
  r.bytes.resize(n * r.elem);
  if (n > 0) check(H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, r.bytes.data()), std::string("read ") + name);
  return r;
}

template <class T>
std::vector<T> as_vec(const Data& d)
{
  std::vector<T> v(d.bytes.size() / sizeof(T));
  if (!v.empty()) std::memcpy(v.data(), d.bytes.data(), d.bytes.size());
  return v;
}

uint64_t read_attr_u64(hid_t obj, const char* name)
{
  H a = check(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose, std::string("open attr ") + name);
  uint64_t v = 0;
  check(H5Aread(a, H5T_NATIVE_UINT64, &v), std::string("read attr ") + name);
  return v;
}

// sparse tag values as raw bytes in the file's value type
struct Tag {
  std::vector<uint64_t> ids;
  std::vector<char> values;
  size_t size = 0;  // bytes per value
  bool present = false;
};

Tag read_tag(hid_t tags, const char* name, hid_t mem_type)
{
  Tag t;
  if (!exists(tags, name)) return t;
  H g = open_group(tags, name);
  t.ids = as_vec<uint64_t>(read_data(g, "id_list", H5T_NATIVE_UINT64));
  if (mem_type >= 0) {
    Data v = read_data(g, "values", mem_type);
    t.values = std::move(v.bytes);
    t.size = v.elem;
  } else {  // raw: the stored type itself
    H d = check(H5Dopen2(g, "values", H5P_DEFAULT), H5Dclose, "open values");
    H ft = check(H5Dget_type(d), H5Tclose, "values type");
    H nt = check(H5Tget_native_type(ft, H5T_DIR_ASCEND), H5Tclose, "native type");
    t.size = H5Tget_size(nt);
    t.values.resize(t.ids.size() * t.size);
    if (!t.ids.empty()) check(H5Dread(d, nt, H5S_ALL, H5S_ALL, H5P_DEFAULT, t.values.data()), "read values");
  }
  if (t.values.size() != t.ids.size() * t.size) throw FormatError{std::string("tag ") + name + ": sizes disagree"};
  t.present = true;
  return t;
}

std::string str32(const char* p, size_t n)
{
  size_t k = 0;
  while (k < n && p[k]) ++k;
  return std::string(p, k);
}

}  // namespace

struct xdg_h5m_model {
  std::vector<double> verts;
  std::vector<int64_t> vert_ids;
  std::vector<int32_t> surf_ids;
  std::vector<int64_t> surf_tri_off{0};
  std::vector<int32_t> tris;
  std::vector<int64_t> tri_ids;
  std::vector<int32_t> surf_sense;
  std::vector<int32_t> vol_ids;
  std::vector<int64_t> vol_tet_off{0};
  std::vector<int32_t> tets;
  std::vector<int64_t> tet_ids;
  std::vector<char> group_names;
  std::vector<int64_t> member_off{0};
  std::vector<int32_t> member_dim, member_id;
};

namespace {

template <class F>
int guarded(F&& f)
{
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are reported here, not printed
  try {
    return f();
  } catch (const FormatError& e) {
    g_err = e.msg;
    return XDG_H5M_ERR_FORMAT;
  } catch (const IoError& e) {
    g_err = e.msg;
    return XDG_H5M_ERR_IO;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return XDG_H5M_ERR_INVALID;
  } catch (const std::exception& e) {
    g_err = e.what();
    return XDG_H5M_ERR_IO;
  }
}

void read_file(const char* path, xdg_h5m_model& m)
{
  H f = check(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("open file ") + path);
  if (!exists(f, "/tstt")) throw FormatError{"no /tstt group"};
  H tstt = open_group(f, "/tstt");

  // vertices
  H nodes = open_group(tstt, "nodes");
  Data c = read_data(nodes, "coordinates", H5T_NATIVE_DOUBLE);
  if (c.dims.size() != 2 || c.dims[1] != 3) throw FormatError{"nodes/coordinates must be n x 3"};
  m.verts = as_vec<double>(c);
  const int64_t nv = (int64_t)c.dims[0];
  uint64_t node0 = 0;
  {
    H d = check(H5Dopen2(nodes, "coordinates", H5P_DEFAULT), H5Dclose, "open coordinates");
    node0 = read_attr_u64(d, "start_id");
  }
  m.vert_ids.resize((size_t)nv);
  for (int64_t i = 0; i < nv; ++i) m.vert_ids[(size_t)i] = (int64_t)(node0 + (uint64_t)i);

  // elements: triangles and tetrahedra blocks (by connectivity width)
  struct Block {
    uint64_t start = 0;
    int64_t n = 0;
    int width = 0;
    std::vector<uint64_t> conn;
  };
  std::vector<Block> blocks;
  if (exists(tstt, "elements")) {
    H el = open_group(tstt, "elements");
    H5G_info_t info;
    check(H5Gget_info(el, &info), "elements info");
    for (hsize_t i = 0; i < info.nlinks; ++i) {
      char name[256];
      if (H5Lget_name_by_idx(el, ".", H5_INDEX_NAME, H5_ITER_INC, i, name, sizeof name, H5P_DEFAULT) < 0) continue;
      H g = open_group(el, name);
      if (!exists(g, "connectivity")) continue;
      Data d = read_data(g, "connectivity", H5T_NATIVE_UINT64);
      if (d.dims.size() != 2 || (d.dims[1] != 3 && d.dims[1] != 4)) continue;  // not a triangle/tet block
      Block b;
      b.n = (int64_t)d.dims[0];
      b.width = (int)d.dims[1];
      b.conn = as_vec<uint64_t>(d);
      H cd = check(H5Dopen2(g, "connectivity", H5P_DEFAULT), H5Dclose, "open connectivity");
      b.start = read_attr_u64(cd, "start_id");
      blocks.push_back(std::move(b));
    }
  }
  auto find_elem = [&](uint64_t id, int width, const Block** blk, int64_t* idx) {
    for (const auto& b : blocks)
      if (b.width == width && id >= b.start && id < b.start + (uint64_t)b.n) {
        *blk = &b;
        *idx = (int64_t)(id - b.start);
        return true;
      }
    return false;
  };
  auto vidx = [&](uint64_t id) {
    if (id < node0 || id >= node0 + (uint64_t)nv) throw FormatError{"connectivity names a non-vertex id"};
    return (int32_t)(id - node0);
  };

  // sets
  if (!exists(tstt, "sets")) throw FormatError{"no /tstt/sets"};
  H sets = open_group(tstt, "sets");
  Data ld = read_data(sets, "list", H5T_NATIVE_INT64);
  if (ld.dims.size() != 2 || ld.dims[1] != 4) throw FormatError{"sets/list must be n x 4"};
  const auto list = as_vec<int64_t>(ld);
  const int64_t ns = (int64_t)ld.dims[0];
  uint64_t set0 = 0;
  {
    H d = check(H5Dopen2(sets, "list", H5P_DEFAULT), H5Dclose, "open list");
    set0 = read_attr_u64(d, "start_id");
  }
  const auto contents = exists(sets, "contents") ? as_vec<uint64_t>(read_data(sets, "contents", H5T_NATIVE_UINT64))
                                                 : std::vector<uint64_t>();
  // the ids in a set's contents (explicit list, or (start, count) ranges)
  auto set_contents = [&](int64_t s) {
    const int64_t lo = s == 0 ? 0 : list[(size_t)(4 * (s - 1))] + 1;
    const int64_t hi = list[(size_t)(4 * s)] + 1;
    if (lo < 0 || hi < lo || hi > (int64_t)contents.size()) throw FormatError{"sets/list content index out of range"};
    std::vector<uint64_t> ids;
    if (list[(size_t)(4 * s + 3)] & SET_RANGE_BIT) {
      if ((hi - lo) % 2) throw FormatError{"range contents must be (start, count) pairs"};
      for (int64_t k = lo; k < hi; k += 2)
        for (uint64_t j = 0; j < contents[(size_t)k + 1]; ++j) ids.push_back(contents[(size_t)k] + j);
    } else {
      ids.assign(contents.begin() + lo, contents.begin() + hi);
    }
    return ids;
  };
  auto set_index = [&](uint64_t id) -> int64_t {
    return (id >= set0 && id < set0 + (uint64_t)ns) ? (int64_t)(id - set0) : -1;
  };

  // tags on sets
  std::vector<int32_t> dim((size_t)ns, -1), gid((size_t)ns, 0);
  std::vector<std::string> cat((size_t)ns), name((size_t)ns);
  std::vector<std::pair<int64_t, int64_t>> sense((size_t)ns, {-1, -1});
  if (exists(tstt, "tags")) {
    H tags = open_group(tstt, "tags");
    const Tag td = read_tag(tags, "GEOM_DIMENSION", H5T_NATIVE_INT32);
    for (size_t i = 0; i < td.ids.size(); ++i) {
      const int64_t s = set_index(td.ids[i]);
      if (s >= 0) std::memcpy(&dim[(size_t)s], td.values.data() + 4 * i, 4);
    }
    const Tag tg = read_tag(tags, "GLOBAL_ID", H5T_NATIVE_INT32);
    for (size_t i = 0; i < tg.ids.size(); ++i) {
      const int64_t s = set_index(tg.ids[i]);
      if (s >= 0) std::memcpy(&gid[(size_t)s], tg.values.data() + 4 * i, 4);
    }
    for (const char* tn : {"CATEGORY", "NAME"}) {
      const Tag t = read_tag(tags, tn, -1);
      for (size_t i = 0; i < t.ids.size(); ++i) {
        const int64_t s = set_index(t.ids[i]);
        if (s >= 0) (tn[0] == 'C' ? cat : name)[(size_t)s] = str32(t.values.data() + t.size * i, t.size);
      }
    }
    const Tag tsn = read_tag(tags, "GEOM_SENSE_2", -1);
    if (tsn.present) {
      if (tsn.size != 16 && tsn.size != 8) throw FormatError{"GEOM_SENSE_2 values must hold two handles"};
      for (size_t i = 0; i < tsn.ids.size(); ++i) {
        const int64_t s = set_index(tsn.ids[i]);
        if (s < 0) continue;
        uint64_t h[2] = {0, 0};
        if (tsn.size == 16) {
          std::memcpy(h, tsn.values.data() + 16 * i, 16);
        } else {
          uint32_t h32[2];
          std::memcpy(h32, tsn.values.data() + 8 * i, 8);
          h[0] = h32[0];
          h[1] = h32[1];
        }
        sense[(size_t)s] = {h[0] ? set_index(h[0]) : -1, h[1] ? set_index(h[1]) : -1};
      }
    }
  }

  // surfaces and volumes in file order
  for (int64_t s = 0; s < ns; ++s) {
    if (dim[(size_t)s] == 2) {
      m.surf_ids.push_back(gid[(size_t)s]);
      for (uint64_t id : set_contents(s)) {
        const Block* b = nullptr;
        int64_t k = 0;
        if (!find_elem(id, 3, &b, &k)) continue;  // vertices, curves
        for (int j = 0; j < 3; ++j) m.tris.push_back(vidx(b->conn[(size_t)(3 * k + j)]));
        m.tri_ids.push_back((int64_t)id);
      }
      m.surf_tri_off.push_back((int64_t)m.tri_ids.size());
      for (int j = 0; j < 2; ++j) {
        const int64_t v = j == 0 ? sense[(size_t)s].first : sense[(size_t)s].second;
        if (v >= 0 && dim[(size_t)v] != 3) throw FormatError{"GEOM_SENSE_2 names a set that is not a volume"};
        m.surf_sense.push_back(v >= 0 ? gid[(size_t)v] : -1);
      }
    } else if (dim[(size_t)s] == 3) {
      m.vol_ids.push_back(gid[(size_t)s]);
      for (uint64_t id : set_contents(s)) {
        const Block* b = nullptr;
        int64_t k = 0;
        if (!find_elem(id, 4, &b, &k)) continue;
        for (int j = 0; j < 4; ++j) m.tets.push_back(vidx(b->conn[(size_t)(4 * k + j)]));
        m.tet_ids.push_back((int64_t)id);
      }
      m.vol_tet_off.push_back((int64_t)m.tet_ids.size());
    }
  }
  // groups (parse_metadata reads them by CATEGORY "Group")
  for (int64_t s = 0; s < ns; ++s) {
    if (cat[(size_t)s] != "Group") continue;
    std::string nm = name[(size_t)s];
    nm.resize(XDG_H5M_NAME_SIZE, '\0');
    m.group_names.insert(m.group_names.end(), nm.begin(), nm.end());
    for (uint64_t id : set_contents(s)) {
      const int64_t t = set_index(id);
      if (t < 0 || (dim[(size_t)t] != 2 && dim[(size_t)t] != 3)) continue;
      m.member_dim.push_back(dim[(size_t)t]);
      m.member_id.push_back(gid[(size_t)t]);
    }
    m.member_off.push_back((int64_t)m.member_id.size());
  }
}

}  // namespace

extern "C" {

const char* xdg_h5m_last_error(void) { return g_err.c_str(); }

int xdg_h5m_read(const char* path, xdg_h5m_model** out)
{
  return guarded([&]() {
    if (!path || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto* m = new xdg_h5m_model();
    try {
      read_file(path, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return XDG_H5M_OK;
  });
}

void xdg_h5m_free(xdg_h5m_model* m) { delete m; }

int xdg_h5m_get_counts(const xdg_h5m_model* m, xdg_h5m_counts* c)
{
  return guarded([&]() {
    if (!m || !c) throw std::invalid_argument("null argument");
    std::memset(c, 0, sizeof *c);
    c->n_verts = (int64_t)m->verts.size() / 3;
    c->n_tris = (int64_t)m->tri_ids.size();
    c->n_tets = (int64_t)m->tet_ids.size();
    c->n_surfaces = (int32_t)m->surf_ids.size();
    c->n_volumes = (int32_t)m->vol_ids.size();
    c->n_groups = (int32_t)(m->member_off.size() - 1);
    c->n_group_members = (int64_t)m->member_id.size();
    return XDG_H5M_OK;
  });
}

int xdg_h5m_get_mesh(const xdg_h5m_model* m, double* verts, int64_t* vert_ids, int32_t* surf_ids,
                     int64_t* surf_tri_off, int32_t* tris, int64_t* tri_ids, int32_t* surf_sense, int32_t* vol_ids,
                     int64_t* vol_tet_off, int32_t* tets, int64_t* tet_ids)
{
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null model");
    auto put = [](auto* dst, const auto& src) {
      if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(verts, m->verts);
    put(vert_ids, m->vert_ids);
    put(surf_ids, m->surf_ids);
    put(surf_tri_off, m->surf_tri_off);
    put(tris, m->tris);
    put(tri_ids, m->tri_ids);
    put(surf_sense, m->surf_sense);
    put(vol_ids, m->vol_ids);
    put(vol_tet_off, m->vol_tet_off);
    put(tets, m->tets);
    put(tet_ids, m->tet_ids);
    return XDG_H5M_OK;
  });
}

int xdg_h5m_get_groups(const xdg_h5m_model* m, char* names, int64_t* member_off, int32_t* member_dim,
                       int32_t* member_id)
{
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null model");
    if (names && !m->group_names.empty()) std::memcpy(names, m->group_names.data(), m->group_names.size());
    if (member_off) std::memcpy(member_off, m->member_off.data(), m->member_off.size() * 8);
    if (member_dim && !m->member_dim.empty()) std::memcpy(member_dim, m->member_dim.data(), m->member_dim.size() * 4);
    if (member_id && !m->member_id.empty()) std::memcpy(member_id, m->member_id.data(), m->member_id.size() * 4);
    return XDG_H5M_OK;
  });
}

int xdg_h5m_write(const char* path, int64_t n_verts, const double* verts, int32_t n_surfaces,
                  const int32_t* surf_ids, const int64_t* surf_tri_off, const int32_t* tris,
                  const int32_t* surf_sense, int32_t n_volumes, const int32_t* vol_ids, const int64_t* vol_tet_off,
                  const int32_t* tets, int32_t n_groups, const char* names, const int64_t* member_off,
                  const int32_t* member_dim, const int32_t* member_id)
{
  return guarded([&]() {
    if (!path || n_verts < 0 || n_surfaces < 0 || n_volumes < 0 || n_groups < 0)
      throw std::invalid_argument("bad sizes");
    if ((n_verts && !verts) || (n_surfaces && (!surf_ids || !surf_tri_off || !surf_sense)) ||
        (n_volumes && (!vol_ids || !vol_tet_off)) || (n_groups && (!names || !member_off || !member_dim || !member_id)))
      throw std::invalid_argument("null array");
    const int64_t nt = n_surfaces ? surf_tri_off[n_surfaces] : 0;
    const int64_t nk = n_volumes ? vol_tet_off[n_volumes] : 0;
    if ((nt && !tris) || (nk && !tets)) throw std::invalid_argument("null element array");
    for (int64_t i = 0; i < 3 * nt; ++i)
      if (tris[i] < 0 || tris[i] >= n_verts) throw std::invalid_argument("triangle vertex index out of range");
    for (int64_t i = 0; i < 4 * nk; ++i)
      if (tets[i] < 0 || tets[i] >= n_verts) throw std::invalid_argument("tet vertex index out of range");
    std::map<int32_t, int32_t> vol_slot, surf_slot;
    for (int32_t v = 0; v < n_volumes; ++v)
      if (!vol_slot.emplace(vol_ids[v], v).second) throw std::invalid_argument("duplicate volume id");
    for (int32_t s = 0; s < n_surfaces; ++s)
      if (!surf_slot.emplace(surf_ids[s], s).second) throw std::invalid_argument("duplicate surface id");

    // one id space: nodes, triangles, tets, then sets (surfaces, volumes, groups)
    const uint64_t node0 = 1, tri0 = node0 + (uint64_t)n_verts, tet0 = tri0 + (uint64_t)nt;
    const uint64_t set0 = tet0 + (uint64_t)nk;
    const uint64_t surf_set0 = set0, vol_set0 = set0 + (uint64_t)n_surfaces;
    const uint64_t grp_set0 = vol_set0 + (uint64_t)n_volumes;
    const int64_t nsets = (int64_t)n_surfaces + n_volumes + n_groups;

    H f = check(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, std::string("create ") + path);
    H tstt = make_group(f, "tstt");
    write_attr_u64(tstt, "max_id", set0 + (uint64_t)nsets - 1);
    {
      H nodes = make_group(tstt, "nodes");
      H d = write_data(nodes, "coordinates", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, n_verts, 3, verts);
      write_attr_u64(d, "start_id", node0);
    }
    {
      H el = make_group(tstt, "elements");
      auto block = [&](const char* name, int32_t type, int64_t n, int width, const int32_t* idx, uint64_t start) {
        H g = make_group(el, name);
        write_attr_i32(g, "element_type", type);
        std::vector<uint64_t> conn((size_t)(n * width));
        for (int64_t i = 0; i < n * width; ++i) conn[(size_t)i] = node0 + (uint64_t)idx[i];
        H d = write_data(g, "connectivity", H5T_STD_U64LE, H5T_NATIVE_UINT64, n, width, conn.data());
        write_attr_u64(d, "start_id", start);
      };
      if (nt) block("Tri3", MB_TRI, nt, 3, tris, tri0);
      if (nk) block("Tet4", MB_TET, nk, 4, tets, tet0);
    }
    // set lists: contents / children / parents with inclusive end indices
    std::vector<int64_t> list;
    std::vector<uint64_t> contents, children, parents;
    auto end_set = [&]() {
      list.push_back((int64_t)contents.size() - 1);
      list.push_back((int64_t)children.size() - 1);
      list.push_back((int64_t)parents.size() - 1);
      list.push_back(0x2);  // a set (not ordered), explicit contents
    };
    for (int32_t s = 0; s < n_surfaces; ++s) {
      for (int64_t i = surf_tri_off[s]; i < surf_tri_off[s + 1]; ++i) contents.push_back(tri0 + (uint64_t)i);
      for (int j = 0; j < 2; ++j) {
        const int32_t v = surf_sense[2 * s + j];
        if (v < 0) continue;
        auto it = vol_slot.find(v);
        if (it == vol_slot.end()) throw std::invalid_argument("surface sense names an unknown volume");
        parents.push_back(vol_set0 + (uint64_t)it->second);
      }
      end_set();
    }
    for (int32_t v = 0; v < n_volumes; ++v) {
      for (int64_t i = vol_tet_off[v]; i < vol_tet_off[v + 1]; ++i) contents.push_back(tet0 + (uint64_t)i);
      for (int32_t s = 0; s < n_surfaces; ++s)
        if (surf_sense[2 * s] == vol_ids[v] || surf_sense[2 * s + 1] == vol_ids[v])
          children.push_back(surf_set0 + (uint64_t)s);
      end_set();
    }
    for (int32_t g = 0; g < n_groups; ++g) {
      for (int64_t i = member_off[g]; i < member_off[g + 1]; ++i) {
        const auto& slots = member_dim[i] == 3 ? vol_slot : surf_slot;
        auto it = slots.find(member_id[i]);
        if ((member_dim[i] != 2 && member_dim[i] != 3) || it == slots.end())
          throw std::invalid_argument("group member is not a known surface or volume");
        contents.push_back((member_dim[i] == 3 ? vol_set0 : surf_set0) + (uint64_t)it->second);
      }
      end_set();
    }
    {
      H sets = make_group(tstt, "sets");
      H d = write_data(sets, "list", H5T_STD_I64LE, H5T_NATIVE_INT64, nsets, 4, list.data());
      write_attr_u64(d, "start_id", set0);
      write_data(sets, "contents", H5T_STD_U64LE, H5T_NATIVE_UINT64, (int64_t)contents.size(), 0, contents.data());
      write_data(sets, "children", H5T_STD_U64LE, H5T_NATIVE_UINT64, (int64_t)children.size(), 0, children.data());
      write_data(sets, "parents", H5T_STD_U64LE, H5T_NATIVE_UINT64, (int64_t)parents.size(), 0, parents.data());
    }
    // tags
    H tags = make_group(tstt, "tags");
    std::vector<uint64_t> geo_ids;
    std::vector<int32_t> dims, gids;
    for (int32_t s = 0; s < n_surfaces; ++s) {
      geo_ids.push_back(surf_set0 + (uint64_t)s);
      dims.push_back(2);
      gids.push_back(surf_ids[s]);
    }
    for (int32_t v = 0; v < n_volumes; ++v) {
      geo_ids.push_back(vol_set0 + (uint64_t)v);
      dims.push_back(3);
      gids.push_back(vol_ids[v]);
    }
    write_tag(tags, "GEOM_DIMENSION", H5T_NATIVE_INT32, geo_ids, dims.data());
    std::vector<uint64_t> gid_ids = geo_ids;
    std::vector<int32_t> gid_vals = gids;
    for (int32_t g = 0; g < n_groups; ++g) {  // groups are numbered from 1
      gid_ids.push_back(grp_set0 + (uint64_t)g);
      gid_vals.push_back(g + 1);
    }
    write_tag(tags, "GLOBAL_ID", H5T_NATIVE_INT32, gid_ids, gid_vals.data());
    {
      H opq = check(H5Tcreate(H5T_OPAQUE, XDG_H5M_NAME_SIZE), H5Tclose, "opaque type");
      std::vector<char> cat;
      std::vector<uint64_t> cat_ids;
      auto add = [&](uint64_t id, const char* v) {
        char b[XDG_H5M_NAME_SIZE] = {0};
        std::strncpy(b, v, XDG_H5M_NAME_SIZE - 1);
        cat.insert(cat.end(), b, b + XDG_H5M_NAME_SIZE);
        cat_ids.push_back(id);
      };
      for (int32_t s = 0; s < n_surfaces; ++s) add(surf_set0 + (uint64_t)s, "Surface");
      for (int32_t v = 0; v < n_volumes; ++v) add(vol_set0 + (uint64_t)v, "Volume");
      for (int32_t g = 0; g < n_groups; ++g) add(grp_set0 + (uint64_t)g, "Group");
      write_tag(tags, "CATEGORY", opq, cat_ids, cat.data());
      if (n_groups) {
        std::vector<uint64_t> name_ids;
        for (int32_t g = 0; g < n_groups; ++g) name_ids.push_back(grp_set0 + (uint64_t)g);
        write_tag(tags, "NAME", opq, name_ids, names);
      }
    }
    {
      const hsize_t two = 2;
      H pair = check(H5Tarray_create2(H5T_NATIVE_UINT64, 1, &two), H5Tclose, "handle pair type");
      std::vector<uint64_t> ids, vals;
      for (int32_t s = 0; s < n_surfaces; ++s) {
        ids.push_back(surf_set0 + (uint64_t)s);
        for (int j = 0; j < 2; ++j) {
          const int32_t v = surf_sense[2 * s + j];
          vals.push_back(v < 0 ? 0 : vol_set0 + (uint64_t)vol_slot.at(v));
        }
      }
      write_tag(tags, "GEOM_SENSE_2", pair, ids, vals.data());
    }
    return XDG_H5M_OK;
  });
}

}  // extern "C"
