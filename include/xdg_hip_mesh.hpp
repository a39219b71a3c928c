/*
 * xdg_hip_mesh.hpp -- a mesh manager's volumes flattened into the arrays the
 * xdg_hip tree builders take: the geometry-gathering half of the reference's
 * EmbreeRayTracer::create_surface_tree / create_element_tree
 * (src/embree/ray_tracer.cpp:50-105, 150-188) and of the global trees
 * (ray_tracer.cpp:190-224), plus the boundary-face -> element map
 * (MeshManager::get_boundary_face_element) that the implicit-complement
 * re-entry of XDG::segments uses (src/xdg.cpp:164-185).
 *
 * Header only, templated on the manager type MM, so that it runs against
 * xdg::MeshManager (MOAB, libMesh, the test mock) and against any class
 * offering the same calls:
 *   std::vector<MeshID> get_volume_surfaces(MeshID volume)
 *   std::pair<MeshID, MeshID> surface_senses(MeshID surface)  (forward, reverse)
 *   std::vector<MeshID> get_surface_faces(MeshID surface)
 *   std::array<P, 3> face_vertices(MeshID face)               P has .x .y .z
 *   std::vector<MeshID> get_volume_elements(MeshID volume)
 *   std::vector<P> get_volume_vertices(MeshID volume)
 *   std::vector<int> get_volume_connectivity(MeshID volume)    4 local vertex
 *                             indices per element, get_volume_elements order
 *   MeshID adjacent_element(MeshID element, int face)          ID_NONE outside
 *   MeshID get_boundary_face_element(MeshID face)
 * The element face order is the mesh library's canonical one (MOAB:
 * include/xdg/moab/direct_access.h:181-183; the reference's test mock:
 * tests/mesh_mock.h:108-115); the caller picks it from mesh_library().
 */
#ifndef XDG_HIP_MESH_HPP
#define XDG_HIP_MESH_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "xdg_hip.hpp"

namespace xdg_hip {

// local vertex triples of faces 0..3
constexpr std::array<int32_t, 12> MOAB_TET_FACES = {0, 1, 3, 1, 2, 3, 2, 0, 3, 0, 2, 1};
constexpr std::array<int32_t, 12> MOCK_TET_FACES = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};

// constants.h: FP_BOX_TOL; ray_tracing_interface.h: numerical_precision_
constexpr double FP_BOX_TOL = 1e-6;
constexpr double BUMP_FLOOR = 1e-3;

// Triangles of one surface tree, one record per face (vertices not shared).
struct SurfaceArrays {
  std::vector<double> verts;
  std::vector<int32_t> tris, face_ids, surf_ids;
  std::vector<uint8_t> flip;
};

// Tetrahedra of one element tree: element index i is elem_ids[i] in the
// manager; adj holds element indices (-1 on the boundary of the tree).
struct ElementArrays {
  std::vector<double> verts;
  std::vector<int32_t> tets, adj;
  std::vector<MeshID> elem_ids;
  std::unordered_map<MeshID, int32_t> index_of;  // manager id -> element index
};

// RayTracer::bounding_box_bump (src/ray_tracing_interface.cpp:20-24):
// max(volume box chord * FP_BOX_TOL, 1e-3).  The volume box is
// MeshManager::volume_bounding_box (mesh_manager_interface.cpp:237-266),
// which starts from the zero box and so always contains the origin.
template <class MM>
double volume_bump(const MM& mm, MeshID volume)
{
  double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
  for (MeshID s : mm.get_volume_surfaces(volume))
    for (MeshID f : mm.get_surface_faces(s))
      for (const auto& p : mm.face_vertices(f)) {
        const double c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], c[k]);
          hi[k] = std::max(hi[k], c[k]);
        }
      }
  const double w[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  return std::max(std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * FP_BOX_TOL, BUMP_FLOOR);
}

// The faces of `surfaces` as one tree.  A face is flipped where `volume` is
// the reverse-sense parent of its surface (the reference records the tree as
// the surface's reverse_vol, ray_tracer.cpp:92-100, and flips the normal for
// rays in that tree, triangle_intersect.cpp:84).  volume == ID_NONE builds an
// unflipped tree (the global surface tree).  A surface that `volume` does not
// bound is an error, as in the reference (ray_tracer.cpp:98-100).
template <class MM>
SurfaceArrays flatten_surfaces(const MM& mm, const std::vector<MeshID>& surfaces, MeshID volume)
{
  SurfaceArrays a;
  for (MeshID surface : surfaces) {
    const auto senses = mm.surface_senses(surface);
    if (volume != ID_NONE && volume != senses.first && volume != senses.second)
      throw std::invalid_argument("volume " + std::to_string(volume) + " is not a parent of surface " +
                                  std::to_string(surface));
    const uint8_t rev = (volume != ID_NONE && volume != senses.first && volume == senses.second) ? 1 : 0;
    for (MeshID face : mm.get_surface_faces(surface)) {
      const auto v = mm.face_vertices(face);
      for (int k = 0; k < 3; ++k) {
        a.tris.push_back((int32_t)(a.verts.size() / 3));
        a.verts.insert(a.verts.end(), {v[k].x, v[k].y, v[k].z});
      }
      a.face_ids.push_back(face);
      a.surf_ids.push_back(surface);
      a.flip.push_back(rev);
    }
  }
  return a;
}

// The surface tree of one volume (create_surface_tree).
template <class MM>
SurfaceArrays flatten_volume_surfaces(const MM& mm, MeshID volume)
{
  return flatten_surfaces(mm, mm.get_volume_surfaces(volume), volume);
}

// The elements of `volumes` as one tree, in volume order; neighbours across
// the volumes' shared faces stay connected, neighbours outside are -1.
template <class MM>
ElementArrays flatten_elements(const MM& mm, const std::vector<MeshID>& volumes)
{
  ElementArrays a;
  for (MeshID volume : volumes) {
    const int32_t base = (int32_t)(a.verts.size() / 3);
    for (const auto& p : mm.get_volume_vertices(volume)) a.verts.insert(a.verts.end(), {p.x, p.y, p.z});
    const auto conn = mm.get_volume_connectivity(volume);
    const auto elems = mm.get_volume_elements(volume);
    if (conn.size() != 4 * elems.size())
      throw std::invalid_argument("volume " + std::to_string(volume) + ": connectivity is not 4 per element");
    for (int c : conn) a.tets.push_back(base + c);
    for (MeshID e : elems) {
      if (!a.index_of.emplace(e, (int32_t)a.elem_ids.size()).second)
        throw std::invalid_argument("element " + std::to_string(e) + " belongs to two volumes");
      a.elem_ids.push_back(e);
    }
  }
  a.adj.reserve(4 * a.elem_ids.size());
  for (MeshID e : a.elem_ids)
    for (int f = 0; f < 4; ++f) {
      const MeshID n = mm.adjacent_element(e, f);
      const auto it = n == ID_NONE ? a.index_of.end() : a.index_of.find(n);
      a.adj.push_back(it == a.index_of.end() ? -1 : it->second);
    }
  return a;
}

// Boundary face -> element index of `elements`, indexed by face id (the face
// ids of a surface tree over those faces; -1 where the owner is outside).
template <class MM>
std::vector<int32_t> boundary_face_elements(const MM& mm, const std::vector<MeshID>& faces,
                                            const ElementArrays& elements)
{
  MeshID top = -1;
  for (MeshID f : faces) {
    if (f < 0) throw std::invalid_argument("boundary face ids must be >= 0");
    top = std::max(top, f);
  }
  std::vector<int32_t> map((size_t)(top + 1), -1);
  for (MeshID f : faces) {
    const MeshID e = mm.get_boundary_face_element(f);
    const auto it = e == ID_NONE ? elements.index_of.end() : elements.index_of.find(e);
    map[(size_t)f] = it == elements.index_of.end() ? -1 : it->second;
  }
  return map;
}

// Build helpers on a HipRayTracer (tree ids of the device library).
inline TreeID build_surface_tree(HipRayTracer& rt, const SurfaceArrays& a, double bump)
{
  return rt.create_surface_tree(a.verts, a.tris, a.face_ids, a.surf_ids, a.flip, bump);
}

inline TreeID build_element_tree(HipRayTracer& rt, const ElementArrays& a, const std::array<int32_t, 12>& order)
{
  return rt.create_element_tree(a.verts, a.tets, a.adj, order);
}

}  // namespace xdg_hip

#endif  // XDG_HIP_MESH_HPP
