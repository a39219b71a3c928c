// C++ facade (include/xdg_hip.hpp) on the reference's MeshMock box: the
// known answers of the reference's own tests (tests/test_ray_fire.cpp,
// test_point_in_volume.cpp, test_occluded.cpp, test_closest.cpp) through the
// RayTracer-shaped calls, plus the facade's error behaviour (unknown trees)
// and find_element / find_volume inside and outside the box.  Exits 0 when
// every check passes; prints each failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "xdg_hip.hpp"

using namespace xdg_hip;

static int failures = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
      ++failures;                                                          \
    }                                                                      \
  } while (0)
static bool near(double a, double b) { return std::fabs(a - b) <= 1e-6; }

int main()
{
  // MeshMock (tests/mesh_mock.h): box [-2,5] x [-3,6] x [-4,7], 12 faces,
  // six surfaces of two faces, one volume; 12 tets around the centre.
  const double x0 = -2, y0 = -3, z0 = -4, x1 = 5, y1 = 6, z1 = 7;
  std::vector<double> verts = {x1, y0, z1, x1, y1, z1, x0, y1, z1, x0, y0, z1, x1, y0, z0, x1, y1, z0,
                               x0, y1, z0, x0, y0, z0, (x0 + x1) / 2, (y0 + y1) / 2, (z0 + z1) / 2};
  std::vector<int32_t> tris = {0, 1, 2, 0, 2, 3, 4, 6, 5, 4, 7, 6, 0, 5, 1, 0, 4, 5,
                               2, 6, 7, 2, 7, 3, 1, 5, 6, 1, 6, 2, 0, 7, 4, 0, 3, 7};
  std::vector<int32_t> faces(12), surfs(12);
  for (int i = 0; i < 12; ++i) {
    faces[i] = i;
    surfs[i] = i / 2;
  }
  std::vector<uint8_t> flip(12, 0);
  std::vector<int32_t> tets = {0, 1, 2, 8, 0, 2, 3, 8, 4, 6, 5, 8, 4, 7, 6, 8, 0, 5, 1, 8, 0, 4, 5, 8,
                               2, 6, 7, 8, 2, 7, 3, 8, 0, 7, 4, 8, 0, 3, 7, 8, 1, 5, 6, 8, 1, 6, 2, 8};
  std::vector<int32_t> adj = {-1, 1, 4, 11, -1, 9, 0, 7, -1, 5, 3, 10, -1, 2, 8, 6, -1, 0, 5, 10, -1, 4, 8, 2,
                              -1, 7, 11, 3, -1, 1, 6, 9, -1, 5, 9, 3, -1, 8, 1, 7, -1, 11, 4, 2, -1, 0, 10, 6};
  const std::array<int32_t, 12> mock_faces = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};

  for (Precision prec : {Precision::PLAIN, Precision::MIXED}) {
    HipRayTracer rt(0, prec);
    rt.init();
    const TreeID vt = rt.create_surface_tree(verts, tris, faces, surfs, flip, 1e-3);
    const TreeID et = rt.create_element_tree(verts, tets, adj, mock_faces);
    CHECK(rt.num_registered_surface_trees() == 1 && rt.num_registered_element_trees() == 1);

    // ray_fire (test_ray_fire.cpp)
    Position o{0, 0, 0};
    CHECK(near(rt.ray_fire(vt, o, {1, 0, 0}).first, 5.0));
    CHECK(near(rt.ray_fire(vt, o, {-1, 0, 0}).first, 2.0));
    CHECK(near(rt.ray_fire(vt, o, {0, 1, 0}).first, 6.0));
    CHECK(near(rt.ray_fire(vt, o, {0, -1, 0}).first, 3.0));
    CHECK(near(rt.ray_fire(vt, o, {0, 0, 1}).first, 7.0));
    CHECK(near(rt.ray_fire(vt, o, {0, 0, -1}).first, 4.0));
    CHECK(near(rt.ray_fire(vt, {-10, 0, 0}, {1, 0, 0}).first, 15.0));
    CHECK(near(rt.ray_fire(vt, {10, 0, 0}, {-1, 0, 0}).first, 12.0));
    CHECK(near(rt.ray_fire(vt, {-10, 0, 0}, {1, 0, 0}, INFTY, HitOrientation::ENTERING).first, 8.0));
    CHECK(near(rt.ray_fire(vt, {10, 0, 0}, {-1, 0, 0}, INFTY, HitOrientation::ENTERING).first, 5.0));
    CHECK(rt.ray_fire(vt, o, {1, 0, 0}, 4.5).second == ID_NONE);
    auto h = rt.ray_fire(vt, o, {1, 0, 0}, 5.1);
    CHECK(h.second != ID_NONE && near(h.first, 5.0));
    std::vector<MeshID> ex;
    CHECK(near(rt.ray_fire(vt, o, {1, 0, 0}, INFTY, HitOrientation::EXITING, &ex).first, 5.0));
    CHECK(ex.size() == 1);
    CHECK(rt.ray_fire(vt, o, {1, 0, 0}, INFTY, HitOrientation::EXITING, &ex).second == ID_NONE);
    CHECK(ex.size() == 1);

    // point_in_volume (test_point_in_volume.cpp:40-78)
    CHECK(rt.point_in_volume(vt, {0, 0, 0}));
    CHECK(rt.point_in_volume(vt, {0.0, 0.0, 1000.0}) == false);
    CHECK(rt.point_in_volume(vt, {4.0 - 1e-6, 0, 0}));
    CHECK(rt.point_in_volume(vt, {5.001, 0, 0}) == false);
    Direction dx{1, 0, 0}, mdx{-1, 0, 0};
    CHECK(rt.point_in_volume(vt, {5.0, 0, 0}, &dx));
    CHECK(rt.point_in_volume(vt, {5.1, 0, 0}, &dx) == false);
    CHECK(rt.point_in_volume(vt, {5.1, 0, 0}, &mdx) == false);

    // occluded (test_occluded.cpp:20-32): hit -> dist == -INFTY
    double dist = 0.0;
    CHECK(rt.occluded(vt, {-100, 0, 0}, {1, 0, 0}, dist) && dist == -INFTY);
    CHECK(rt.occluded(vt, {-100, 0, 0}, {-1, 0, 0}, dist) == false && dist == INFTY);

    // closest (test_closest.cpp:26-42): distance and the triangle id
    auto c = rt.closest(vt, {0, 0, 0});
    CHECK(near(c.first, 2.0) && c.second >= 0 && c.second < 12);
    CHECK(near(rt.closest(vt, {4, 0, 0}).first, 1.0));
    CHECK(near(rt.closest(vt, {10, 0, 0}).first, 5.0));

    // find_element (test_find_element.cpp): inside -> some tet, outside -> none
    CHECK(rt.find_element(et, {0, 0, 0}) >= 0);
    CHECK(rt.find_element(et, {100, 0, 0}) == ID_NONE);
    CHECK(rt.find_element(et + 5, {0, 0, 0}) == ID_NONE);  // unknown tree: warning + ID_NONE

    // unknown surface tree: std::out_of_range as std::map::at
    bool threw = false;
    try {
      rt.ray_fire(vt + 7, o, {1, 0, 0});
    } catch (const std::out_of_range&) {
      threw = true;
    }
    CHECK(threw);

    // find_volume (xdg.cpp:113-128) with the single volume (id 0) and ipc 1
    CHECK(rt.find_volume({vt}, {0}, 1, {0, 0, 0}, {1, 0, 0}) == 0);
    CHECK(rt.find_volume({vt}, {0}, 1, {50, 0, 0}, {1, 0, 0}) == 1);
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
