// The MeshManager -> device-array path (include/xdg_hip_mesh.hpp) run
// against mesh managers written for this test in the shape of
// xdg::MeshManager, then queried through the C++ facade on the GPU:
//
//  * BoxManager: the reference's test-mock box (tests/mesh_mock.h), one
//    volume (1) and its implicit complement (2, the reverse parent of all six
//    surfaces), with ids that are not indices (surfaces 10.., faces 1000..,
//    elements 100..) so that every id remap is exercised;
//  * the reverse-sense flip: the complement's tree sees the box from
//    outside; the unflipped global tree does not;
//  * element trees with the mock face order, the boundary-face map and the
//    implicit-complement re-entry of segments, and a global element tree over
//    two volumes that split the mock's elements;
//  * KuhnManager: a 3^3-hex Kuhn cube with the MOAB face order; walks checked
//    against brute-force containment of each piece's midpoint, and the same
//    mesh with the wrong (mock) face order caught by that check.
//
// Exits 0 when every check passes; prints each failure.
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <vector>

#include "test_managers.h"
#include "xdg_hip_mesh.hpp"

using namespace xdg_hip;

static int failures = 0;
#define CHECK(cond)                                                 \
  do {                                                              \
    if (!(cond)) {                                                  \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      ++failures;                                                   \
    }                                                               \
  } while (0)

static double len(const P3& a, const P3& b)
{
  return std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z));
}

// every piece lies in its element (midpoint containment) and the pieces
// cover the chord; returns the number of pieces that fail containment
template <class MM>
static int check_walk(const MM& mm, const ElementArrays& ea, const std::vector<std::pair<MeshID, double>>& w,
                      const P3& a, const P3& b, double covered)
{
  const double L = len(a, b);
  const P3 u = {(b.x - a.x) / L, (b.y - a.y) / L, (b.z - a.z) / L};
  double s = 0.0, sum = 0.0;
  int bad = 0;
  for (const auto& [e, l] : w) {
    const double mid = s + 0.5 * l;
    const P3 p = {a.x + mid * u.x, a.y + mid * u.y, a.z + mid * u.z};
    if (l > 1e-9 && (e < 0 || !contains(mm.element_vertices(ea.elem_ids[(size_t)e]), p, 1e-9))) ++bad;
    s += l;
    sum += l;
  }
  if (std::fabs(sum - covered) > 1e-9) ++bad;
  return bad;
}

int main()
{
  HipRayTracer rt(0, Precision::MIXED);

  // ---- surfaces: flip and remapped ids --------------------------------------
  BoxManager box;
  const double bump = volume_bump(box, BoxManager::VOL);
  CHECK(std::fabs(bump - 1e-3) < 1e-15);  // chord ~ 14 -> 1.4e-5 < the 1e-3 floor
  const auto vol_arrays = flatten_volume_surfaces(box, BoxManager::VOL);
  const auto ipc_arrays = flatten_volume_surfaces(box, BoxManager::IPC);
  CHECK(vol_arrays.face_ids.size() == 12 && vol_arrays.flip[0] == 0 && ipc_arrays.flip[0] == 1);
  const TreeID vt = build_surface_tree(rt, vol_arrays, bump);
  const TreeID it = build_surface_tree(rt, ipc_arrays, volume_bump(box, BoxManager::IPC));
  std::vector<MeshID> all = {10, 11, 12, 13, 14, 15};
  const TreeID gt = build_surface_tree(rt, flatten_surfaces(box, all, ID_NONE), bump);

  // test_ray_fire.cpp distances from the centre, with the surface ids of the
  // manager and the hit face appended to the exclusion list
  std::vector<MeshID> hist;
  const auto r1 = rt.ray_fire(vt, {0, 0, 0}, {1, 0, 0}, INFTY, HitOrientation::EXITING, &hist);
  CHECK(std::fabs(r1.first - 5.0) < 1e-12 && r1.second == 12);  // faces 1004/1005: surface 12 (x = 5)
  CHECK(hist.size() == 1 && (hist[0] == 1004 || hist[0] == 1005));
  CHECK(std::fabs(rt.ray_fire(vt, {0, 0, 0}, {0, -1, 0}).first - 3.0) < 1e-12);
  CHECK(std::fabs(rt.ray_fire(vt, {0, 0, 0}, {0, 0, 1}).first - 7.0) < 1e-12);
  // the complement sees the box from outside: its faces are flipped
  const auto r2 = rt.ray_fire(it, {10, 0, 0}, {-1, 0, 0});
  CHECK(std::fabs(r2.first - 5.0) < 1e-12 && r2.second == 12);
  const Direction mx = {-1, 0, 0}, px = {1, 0, 0};
  CHECK(rt.point_in_volume(it, {10, 0, 0}, &mx) && !rt.point_in_volume(it, {0, 0, 0}, &px));
  CHECK(rt.point_in_volume(vt, {0, 0, 0}) && !rt.point_in_volume(vt, {10, 0, 0}));
  // the global tree is unflipped: the near face culls, the far one hits
  CHECK(std::fabs(rt.ray_fire(gt, {10, 0, 0}, {-1, 0, 0}).first - 12.0) < 1e-12);
  // a volume that does not bound a surface is rejected, as the reference's
  // fatal_error (ray_tracer.cpp:98-100)
  bool threw = false;
  try {
    flatten_surfaces(box, all, 7);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  CHECK(threw);

  // ---- elements (mock face order), boundary map, re-entry ----------------
  const auto ea = flatten_elements(box, {BoxManager::VOL});
  CHECK(ea.elem_ids.size() == 12 && ea.elem_ids[3] == 103 && ea.index_of.at(107) == 7);
  CHECK(ea.adj[4 * 0 + 0] == -1 && ea.adj[4 * 0 + 1] == 1 && ea.adj[4 * 11 + 3] == 6);
  const TreeID et = build_element_tree(rt, ea, MOCK_TET_FACES);
  std::vector<MeshID> ipc_faces;
  for (MeshID s : all)
    for (MeshID f : box.get_surface_faces(s)) ipc_faces.push_back(f);
  const auto bmap = boundary_face_elements(box, ipc_faces, ea);
  CHECK(bmap.size() == 1012 && bmap[1000] == 0 && bmap[1011] == 9 && bmap[5] == -1);
  rt.attach_boundary(et, it, bmap);
  // find_element: the lowest containing element index, by brute force
  const P3 probes[] = {{0.1, 0.2, 0.3}, {4.5, 1.0, 1.5}, {-1.5, -2.0, 6.0}, {1.5, 5.5, -3.5}, {10, 10, 10}};
  for (const auto& p : probes) {
    MeshID exp = ID_NONE;
    for (int i = 0; i < 12 && exp == ID_NONE; ++i)
      if (contains(box.element_vertices(100 + i), p, 1e-12)) exp = i;
    CHECK(rt.find_element(et, {p.x, p.y, p.z}) == exp);
  }
  // inside the box
  const P3 a = {-1.9, 0.1, 0.2}, b = {4.9, 0.3, -0.1};
  CHECK(check_walk(box, ea, rt.segments(et, {a.x, a.y, a.z}, {b.x, b.y, b.z}), a, b, len(a, b)) == 0);
  // from outside: the start is not in the mesh, the complement's ray enters
  // through the x = -2 face into its boundary element (xdg.cpp:164-185)
  const P3 c = {-3.0, 0.5, 0.5}, d = {4.0, 0.5, 0.5};
  const auto wc = rt.segments(et, {c.x, c.y, c.z}, {d.x, d.y, d.z});
  CHECK(!wc.empty());
  if (!wc.empty()) {
    const MeshID first = ea.elem_ids[(size_t)wc.front().first];
    CHECK(first == 106 || first == 107);  // faces 1006/1007 lie on x = -2
    double sum = 0.0;
    for (const auto& pr : wc) sum += pr.second;
    CHECK(std::fabs(sum - 6.0) < 1e-9);  // from the entry face at x = -2 to the end
  }

  // ---- global element tree over two volumes that split the elements ------
  BoxManager split(true);
  const auto ga = flatten_elements(split, {BoxManager::VOL, BoxManager::VOL_B});
  CHECK(ga.elem_ids.size() == 12);
  const TreeID gtree = build_element_tree(rt, ga, MOCK_TET_FACES);
  const auto w1 = rt.segments(et, {a.x, a.y, a.z}, {b.x, b.y, b.z});
  const auto w2 = rt.segments(gtree, {a.x, a.y, a.z}, {b.x, b.y, b.z});
  CHECK(w1.size() == w2.size());
  for (size_t i = 0; i < std::min(w1.size(), w2.size()); ++i)
    CHECK(ea.elem_ids[(size_t)w1[i].first] == ga.elem_ids[(size_t)w2[i].first] && w1[i].second == w2[i].second);
  // each half alone stops at the interface: its neighbours there are -1
  const auto half = flatten_elements(split, {BoxManager::VOL});
  int cut = 0;
  for (int32_t x : half.adj) cut += x < 0;
  CHECK(half.elem_ids.size() == 6 && cut > 6);
  for (const auto& p : probes) {
    MeshID exp = ID_NONE;
    for (int i = 0; i < 12 && exp == ID_NONE; ++i)
      if (contains(box.element_vertices(100 + i), p, 1e-12)) exp = i;
    const MeshID got = rt.find_element(gtree, {p.x, p.y, p.z});
    CHECK((exp == ID_NONE) == (got == ID_NONE));
    if (exp != ID_NONE && got != ID_NONE) CHECK(contains(box.element_vertices(ga.elem_ids[(size_t)got]), p, 1e-9));
  }

  // ---- MOAB face order on a Kuhn cube, and the wrong order caught --------
  KuhnManager kuhn(3);
  const auto ka = flatten_elements(kuhn, {1});
  CHECK(ka.elem_ids.size() == 162 && ka.elem_ids[0] == 5000);
  const TreeID kt = build_element_tree(rt, ka, MOAB_TET_FACES);
  const TreeID kbad = build_element_tree(rt, ka, MOCK_TET_FACES);
  const P3 segs[][2] = {{{0.05, 0.1, 0.2}, {0.95, 0.9, 0.7}},
                        {{0.5, 0.02, 0.5}, {0.4, 0.97, 0.55}},
                        {{0.9, 0.8, 0.05}, {0.1, 0.15, 0.93}},
                        {{0.33, 0.61, 0.12}, {0.71, 0.29, 0.88}}};
  int bad_good = 0, bad_wrong = 0;
  for (const auto& sg : segs) {
    const P3& s0 = sg[0];
    const P3& s1 = sg[1];
    bad_good += check_walk(kuhn, ka, rt.segments(kt, {s0.x, s0.y, s0.z}, {s1.x, s1.y, s1.z}), s0, s1, len(s0, s1));
    bad_wrong += check_walk(kuhn, ka, rt.segments(kbad, {s0.x, s0.y, s0.z}, {s1.x, s1.y, s1.z}), s0, s1,
                            len(s0, s1));
  }
  CHECK(bad_good == 0);
  CHECK(bad_wrong > 0);

  if (failures == 0) std::printf("mesh flattening: all checks passed\n");
  return failures == 0 ? 0 : 1;
}
