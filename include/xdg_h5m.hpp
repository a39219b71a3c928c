/*
 * xdg_h5m.hpp -- a mesh manager over a DAGMC .h5m file (libxdg_h5m.so), with
 * the calls include/xdg_hip_mesh.hpp flattens, so that a file goes
 * straight to device trees:
 *
 *   xdg_hip::H5mMeshManager mm("model.h5m");
 *   auto faces = xdg_hip::flatten_volume_surfaces(mm, volume);
 *   auto elems = xdg_hip::flatten_elements(mm, {volume});
 *
 * As MOABMeshManager::init (src/moab/mesh_manager.cpp:37-107): volumes and
 * surfaces are the file's GEOM_DIMENSION 3 / 2 sets by GLOBAL_ID, and the
 * implicit complement is created afterwards (MeshManager::
 * create_implicit_complement, src/mesh_manager_interface.cpp:20-45).  Face
 * and element ids are the file ids.  Element adjacency and the boundary
 * face -> element map are found by matching faces (MOAB's canonical tet
 * faces, include/xdg/moab/direct_access.h:181-183), which MOAB derives from
 * its adjacency tables.
 */
#ifndef XDG_H5M_HPP
#define XDG_H5M_HPP

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "xdg_h5m.h"
#include "xdg_hip_mesh.hpp"

namespace xdg_hip {

class H5mMeshManager {
 public:
  struct Vertex {
    double x, y, z;
  };

  explicit H5mMeshManager(const std::string& path)
  {
    xdg_h5m_model* h = nullptr;
    if (xdg_h5m_read(path.c_str(), &h) != XDG_H5M_OK) throw std::runtime_error(xdg_h5m_last_error());
    xdg_h5m_counts c{};
    xdg_h5m_get_counts(h, &c);
    std::vector<int64_t> vert_ids((size_t)c.n_verts), surf_off((size_t)c.n_surfaces + 1), tri_ids((size_t)c.n_tris);
    std::vector<int64_t> vol_off((size_t)c.n_volumes + 1), tet_ids((size_t)c.n_tets);
    verts_.resize((size_t)(3 * c.n_verts));
    surfaces_.resize((size_t)c.n_surfaces);
    tris_.resize((size_t)(3 * c.n_tris));
    std::vector<int32_t> sense((size_t)(2 * c.n_surfaces));
    volumes_.resize((size_t)c.n_volumes);
    tets_.resize((size_t)(4 * c.n_tets));
    const int rc = xdg_h5m_get_mesh(h, verts_.data(), vert_ids.data(), surfaces_.data(), surf_off.data(), tris_.data(),
                                    tri_ids.data(), sense.data(), volumes_.data(), vol_off.data(), tets_.data(),
                                    tet_ids.data());
    xdg_h5m_free(h);
    if (rc != XDG_H5M_OK) throw std::runtime_error(xdg_h5m_last_error());

    auto as_id = [](int64_t x) {
      if (x < 0 || x > INT32_MAX) throw std::runtime_error("h5m: file id beyond the int32 MeshID range");
      return (MeshID)x;
    };
    // implicit complement: one past the largest volume id, the missing
    // parent of every surface that has only one
    ipc_ = volumes_.empty() ? 1 : *std::max_element(volumes_.begin(), volumes_.end()) + 1;
    for (size_t s = 0; s < surfaces_.size(); ++s) {
      const MeshID sid = surfaces_[s];
      senses_[sid] = {sense[2 * s] >= 0 ? sense[2 * s] : ipc_, sense[2 * s + 1] >= 0 ? sense[2 * s + 1] : ipc_};
      for (int64_t i = surf_off[s]; i < surf_off[s + 1]; ++i) {
        const MeshID f = as_id(tri_ids[(size_t)i]);
        face_index_[f] = i;
        surf_faces_[sid].push_back(f);
      }
    }
    for (size_t v = 0; v < volumes_.size(); ++v)
      for (int64_t i = vol_off[v]; i < vol_off[v + 1]; ++i) {
        const MeshID e = as_id(tet_ids[(size_t)i]);
        elem_index_[e] = i;
        vol_elems_[volumes_[v]].push_back(e);
      }
    volumes_.push_back(ipc_);
    // adjacency by matching faces in the MOAB face order
    std::map<std::array<int32_t, 3>, std::pair<int64_t, int>> open;
    const int64_t ne = (int64_t)tets_.size() / 4;
    adj_.assign((size_t)(4 * ne), -1);
    elem_id_.resize((size_t)ne);
    for (const auto& kv : elem_index_) elem_id_[(size_t)kv.second] = kv.first;
    for (int64_t e = 0; e < ne; ++e)
      for (int f = 0; f < 4; ++f) {
        const auto key = face_key(e, f);
        auto it = open.find(key);
        if (it == open.end()) {
          open[key] = {e, f};
        } else {
          adj_[(size_t)(4 * e + f)] = it->second.first;
          adj_[(size_t)(4 * it->second.first + it->second.second)] = e;
          open.erase(it);
        }
      }
    for (const auto& kv : open) bface_[kv.first] = kv.second.first;  // faces with one owner
  }

  const std::vector<MeshID>& volumes() const { return volumes_; }
  const std::vector<MeshID>& surfaces() const { return surfaces_; }
  MeshID implicit_complement() const { return ipc_; }

  std::vector<MeshID> get_volume_surfaces(MeshID volume) const
  {
    std::vector<MeshID> out;
    for (MeshID s : surfaces_) {
      const auto& p = senses_.at(s);
      if (p.first == volume || p.second == volume) out.push_back(s);
    }
    return out;
  }
  std::pair<MeshID, MeshID> surface_senses(MeshID surface) const { return senses_.at(surface); }
  std::vector<MeshID> get_surface_faces(MeshID surface) const
  {
    auto it = surf_faces_.find(surface);
    return it == surf_faces_.end() ? std::vector<MeshID>() : it->second;
  }
  std::array<Vertex, 3> face_vertices(MeshID face) const
  {
    const int64_t i = face_index_.at(face);
    return {vertex(tris_[(size_t)(3 * i)]), vertex(tris_[(size_t)(3 * i + 1)]), vertex(tris_[(size_t)(3 * i + 2)])};
  }
  std::vector<MeshID> get_volume_elements(MeshID volume) const
  {
    auto it = vol_elems_.find(volume);
    return it == vol_elems_.end() ? std::vector<MeshID>() : it->second;
  }
  // the volume's own vertices, in vertex-index order, and its connectivity
  // over them (as MeshManager::volume_local_mesh_data)
  std::vector<Vertex> get_volume_vertices(MeshID volume) const
  {
    std::vector<Vertex> out;
    for (int32_t v : local_vertices(volume)) out.push_back(vertex(v));
    return out;
  }
  std::vector<int> get_volume_connectivity(MeshID volume) const
  {
    const auto lv = local_vertices(volume);
    std::unordered_map<int32_t, int> loc;
    for (size_t i = 0; i < lv.size(); ++i) loc[lv[i]] = (int)i;
    std::vector<int> c;
    for (MeshID e : get_volume_elements(volume)) {
      const int64_t i = elem_index_.at(e);
      for (int k = 0; k < 4; ++k) c.push_back(loc.at(tets_[(size_t)(4 * i + k)]));
    }
    return c;
  }
  MeshID adjacent_element(MeshID element, int face) const
  {
    const int64_t n = adj_[(size_t)(4 * elem_index_.at(element) + face)];
    return n < 0 ? ID_NONE : elem_id_[(size_t)n];
  }
  MeshID get_boundary_face_element(MeshID face) const
  {
    const int64_t i = face_index_.at(face);
    std::array<int32_t, 3> key = {tris_[(size_t)(3 * i)], tris_[(size_t)(3 * i + 1)], tris_[(size_t)(3 * i + 2)]};
    std::sort(key.begin(), key.end());
    auto it = bface_.find(key);
    return it == bface_.end() ? ID_NONE : elem_id_[(size_t)it->second];
  }

 private:
  Vertex vertex(int32_t v) const { return {verts_[(size_t)(3 * v)], verts_[(size_t)(3 * v + 1)], verts_[(size_t)(3 * v + 2)]}; }
  std::array<int32_t, 3> face_key(int64_t e, int f) const
  {
    std::array<int32_t, 3> k;
    for (int j = 0; j < 3; ++j) k[(size_t)j] = tets_[(size_t)(4 * e + MOAB_TET_FACES[(size_t)(3 * f + j)])];
    std::sort(k.begin(), k.end());
    return k;
  }
  std::vector<int32_t> local_vertices(MeshID volume) const
  {
    std::vector<int32_t> v;
    for (MeshID e : get_volume_elements(volume)) {
      const int64_t i = elem_index_.at(e);
      for (int k = 0; k < 4; ++k) v.push_back(tets_[(size_t)(4 * i + k)]);
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
  }

  std::vector<double> verts_;
  std::vector<int32_t> tris_, tets_;
  std::vector<MeshID> surfaces_, volumes_;
  MeshID ipc_ = ID_NONE;
  std::map<MeshID, std::pair<MeshID, MeshID>> senses_;
  std::map<MeshID, std::vector<MeshID>> surf_faces_, vol_elems_;
  std::unordered_map<MeshID, int64_t> face_index_, elem_index_;
  std::vector<int64_t> adj_;
  std::vector<MeshID> elem_id_;
  std::map<std::array<int32_t, 3>, int64_t> bface_;
};

}  // namespace xdg_hip

#endif  // XDG_H5M_HPP
