// Mesh managers written for the flattening tests, in the shape of
// xdg::MeshManager (the calls include/xdg_hip_mesh.hpp uses), and a
// brute-force tet containment for checking walks.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <set>
#include <utility>
#include <vector>

#include "xdg_hip_mesh.hpp"

using xdg_hip::ID_NONE;
using xdg_hip::MeshID;
using xdg_hip::MOAB_TET_FACES;

struct P3 {
  double x, y, z;
};

// barycentric containment with a small tolerance (test check, not the
// library's test): the signed volumes of p against each face
inline double orient(const P3& a, const P3& b, const P3& c, const P3& d)
{
  const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
  const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
  const double wx = d.x - a.x, wy = d.y - a.y, wz = d.z - a.z;
  return ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) + uz * (vx * wy - vy * wx);
}
inline bool contains(const std::array<P3, 4>& t, const P3& p, double tol)
{
  const double v = orient(t[0], t[1], t[2], t[3]);
  const double l0 = orient(p, t[1], t[2], t[3]) / v, l1 = orient(t[0], p, t[2], t[3]) / v;
  const double l2 = orient(t[0], t[1], p, t[3]) / v, l3 = orient(t[0], t[1], t[2], p) / v;
  return l0 >= -tol && l1 >= -tol && l2 >= -tol && l3 >= -tol;
}

// ---------------------------------------------------------------------------
// The mock box: [-2,5] x [-3,6] x [-4,7], 8 corners + centre; 12 faces (two
// per surface); 12 tets, each a boundary face and the centre, so face 0 of
// a tet (mock order) is its boundary face.
// ---------------------------------------------------------------------------
class BoxManager {
 public:
  explicit BoxManager(bool split = false) : split_(split)
  {
    const double x0 = -2, y0 = -3, z0 = -4, x1 = 5, y1 = 6, z1 = 7;
    v_ = {{x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}, {x0, y0, z1}, {x1, y0, z0},
          {x1, y1, z0}, {x0, y1, z0}, {x0, y0, z0}, {(x0 + x1) / 2, (y0 + y1) / 2, (z0 + z1) / 2}};
    t_ = {{0, 1, 2}, {0, 2, 3}, {4, 6, 5}, {4, 7, 6}, {0, 5, 1}, {0, 4, 5},
          {2, 6, 7}, {2, 7, 3}, {1, 5, 6}, {1, 6, 2}, {0, 7, 4}, {0, 3, 7}};
    // the mock's tets: a boundary triangle and the centre, listed in the
    // mock's own element order (the y = -3 pair before the y = 6 pair)
    k_ = {{0, 1, 2, 8}, {0, 2, 3, 8}, {4, 6, 5, 8}, {4, 7, 6, 8}, {0, 5, 1, 8}, {0, 4, 5, 8},
          {2, 6, 7, 8}, {2, 7, 3, 8}, {0, 7, 4, 8}, {0, 3, 7, 8}, {1, 5, 6, 8}, {1, 6, 2, 8}};
    const int a[12][4] = {{-1, 1, 4, 11}, {-1, 9, 0, 7}, {-1, 5, 3, 10}, {-1, 2, 8, 6},
                          {-1, 0, 5, 10}, {-1, 4, 8, 2}, {-1, 7, 11, 3}, {-1, 1, 6, 9},
                          {-1, 5, 9, 3}, {-1, 8, 1, 7}, {-1, 11, 4, 2}, {-1, 0, 10, 6}};
    for (int i = 0; i < 12; ++i)
      for (int f = 0; f < 4; ++f) adj_[i][f] = a[i][f];
  }
  static constexpr MeshID SURF0 = 10, FACE0 = 1000, ELEM0 = 100, VOL = 1, IPC = 2, VOL_B = 3;

  std::vector<MeshID> get_volume_surfaces(MeshID) const { return {10, 11, 12, 13, 14, 15}; }
  std::pair<MeshID, MeshID> surface_senses(MeshID) const { return {VOL, IPC}; }
  std::vector<MeshID> get_surface_faces(MeshID s) const
  {
    return {FACE0 + 2 * (s - SURF0), FACE0 + 2 * (s - SURF0) + 1};
  }
  std::array<P3, 3> face_vertices(MeshID f) const
  {
    const auto& t = t_[f - FACE0];
    return {v_[t[0]], v_[t[1]], v_[t[2]]};
  }
  std::vector<MeshID> get_volume_elements(MeshID vol) const
  {
    std::vector<MeshID> e;
    const int lo = (split_ && vol == VOL_B) ? 6 : 0;
    const int hi = (split_ && vol == VOL) ? 6 : 12;
    if (vol == IPC) return e;
    for (int i = lo; i < hi; ++i) e.push_back(ELEM0 + i);
    return e;
  }
  // the volume's own vertex list (local indices), as volume_local_mesh_data
  std::vector<P3> get_volume_vertices(MeshID vol) const
  {
    std::vector<P3> out;
    for (int i : local_verts(vol)) out.push_back(v_[i]);
    return out;
  }
  std::vector<int> get_volume_connectivity(MeshID vol) const
  {
    const auto lv = local_verts(vol);
    std::map<int, int> loc;
    for (size_t i = 0; i < lv.size(); ++i) loc[lv[i]] = (int)i;
    std::vector<int> c;
    for (MeshID e : get_volume_elements(vol))
      for (int k : k_[e - ELEM0]) c.push_back(loc[k]);
    return c;
  }
  MeshID adjacent_element(MeshID e, int f) const
  {
    const int n = adj_[e - ELEM0][f];
    return n < 0 ? ID_NONE : ELEM0 + n;
  }
  // the tet whose first three vertices are the face's
  MeshID get_boundary_face_element(MeshID face) const
  {
    const auto& t = t_[face - FACE0];
    for (int e = 0; e < 12; ++e)
      if (k_[e][0] == t[0] && k_[e][1] == t[1] && k_[e][2] == t[2]) return ELEM0 + e;
    return ID_NONE;
  }
  std::array<P3, 4> element_vertices(MeshID e) const
  {
    const auto& t = k_[e - ELEM0];
    return {v_[t[0]], v_[t[1]], v_[t[2]], v_[t[3]]};
  }

 private:
  std::vector<int> local_verts(MeshID vol) const
  {
    std::set<int> s;
    for (MeshID e : get_volume_elements(vol))
      for (int k : k_[e - ELEM0]) s.insert(k);
    return std::vector<int>(s.begin(), s.end());
  }
  bool split_;
  std::vector<P3> v_;
  std::vector<std::array<int, 3>> t_;
  std::vector<std::array<int, 4>> k_;
  int adj_[12][4];
};

// ---------------------------------------------------------------------------
// An n^3-hex Kuhn cube in [0,1]^3 (six tets per hex, positively oriented so
// that the MOAB faces point outward), element ids 5000.., adjacency found by
// matching faces in the MOAB face order.
// ---------------------------------------------------------------------------
class KuhnManager {
 public:
  explicit KuhnManager(int n) : n_(n)
  {
    const int m = n + 1;
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
          // a deterministic interior jitter keeps the mesh irregular
          const bool in = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
          const double h = 1.0 / n, a = in ? 0.15 * h : 0.0;
          v_.push_back({i * h + a * std::sin(1.3 * i + 2.1 * j + 0.7 * k), j * h + a * std::sin(0.4 * i + 1.7 * k),
                        k * h + a * std::cos(2.3 * j + 0.9 * i)});
        }
    auto vid = [&](int i, int j, int k) { return i + m * (j + m * k); };
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          for (const auto& p : perms) {
            int c[3] = {0, 0, 0};
            const int v0 = vid(i, j, k);
            c[p[0]] = 1;
            int v1 = vid(i + c[0], j + c[1], k + c[2]);
            c[p[1]] = 1;
            int v2 = vid(i + c[0], j + c[1], k + c[2]);
            const int v3 = vid(i + 1, j + 1, k + 1);
            const int inv = (p[0] > p[1]) + (p[0] > p[2]) + (p[1] > p[2]);
            if (inv % 2) std::swap(v1, v2);
            tets_.push_back({v0, v1, v2, v3});
          }
    std::map<std::array<int, 3>, std::pair<int, int>> open;
    adj_.assign(tets_.size(), {-1, -1, -1, -1});
    for (size_t e = 0; e < tets_.size(); ++e)
      for (int f = 0; f < 4; ++f) {
        std::array<int, 3> key = {tets_[e][MOAB_TET_FACES[3 * f]], tets_[e][MOAB_TET_FACES[3 * f + 1]],
                                  tets_[e][MOAB_TET_FACES[3 * f + 2]]};
        std::sort(key.begin(), key.end());
        auto it = open.find(key);
        if (it == open.end()) {
          open[key] = {(int)e, f};
        } else {
          adj_[e][f] = it->second.first;
          adj_[it->second.first][it->second.second] = (int)e;
          open.erase(it);
        }
      }
  }
  static constexpr MeshID ELEM0 = 5000;
  std::vector<MeshID> get_volume_elements(MeshID) const
  {
    std::vector<MeshID> e;
    for (size_t i = 0; i < tets_.size(); ++i) e.push_back(ELEM0 + (MeshID)i);
    return e;
  }
  std::vector<P3> get_volume_vertices(MeshID) const { return v_; }
  std::vector<int> get_volume_connectivity(MeshID) const
  {
    std::vector<int> c;
    for (const auto& t : tets_) c.insert(c.end(), t.begin(), t.end());
    return c;
  }
  MeshID adjacent_element(MeshID e, int f) const
  {
    const int a = adj_[e - ELEM0][f];
    return a < 0 ? ID_NONE : ELEM0 + a;
  }
  std::array<P3, 4> element_vertices(MeshID e) const
  {
    const auto& t = tets_[e - ELEM0];
    return {v_[t[0]], v_[t[1]], v_[t[2]], v_[t[3]]};
  }

 private:
  int n_;
  std::vector<P3> v_;
  std::vector<std::array<int, 4>> tets_;
  std::vector<std::array<int, 4>> adj_;
};

