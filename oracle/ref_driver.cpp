// ref_driver.cpp -- golden-vector driver around the reference's own sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
// reference sources where they lie (include/xdg/geometry/plucker.h,
// face_common.h, src/geometry/closest.cpp, src/mesh_manager_interface.cpp,
// src/element_face_accessor.cpp with tests/mesh_mock.h) into oracle/_ref/.
// It only runs in the development container (the reference tree does not
// travel); tests/golden/make_golden.py stores its outputs as fixtures.
//
// Protocol: ref_driver <mode> <in.bin> <out.bin>, raw little-endian doubles.
//   plucker : 19 doubles/case (v0 v1 v2 o d tmax tmin use_orient orient)
//             -> 2 doubles (hit, t)
//   edge    : 12 doubles/case (va vb ray ray_normal) -> 1 double
//   normal  : 9 doubles/case -> 3 doubles (triangle_normal, face_common.h)
//   closest : 12 doubles/case (v0 v1 v2 p) -> 4 doubles (region, x, y, z)
//   mock_next : 7 doubles/case (elem r u) on MeshMock -> 2 (next, dist)
//   mock_walk : 8 doubles/case (elem r u distance) -> 1 + 2*64 doubles
//               (count, then up to 64 (elem, len) pairs, zero padded)
//   mock_walk_se : 7 doubles/case (elem start end) -> 1 + 2*64 doubles,
//               MeshManager::walk_elements(elem, start, end)
//   mock_face_normal : 1 double/case (face id) -> 3 doubles
//   mock_tracks : 1 double (count) -> 6 doubles/track: the std::mt19937(42)
//               start/end points of tests/test_tracks.cpp:170-180
#include <cstdio>
#include <random>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xdg/geometry/closest.h"
#include "xdg/geometry/face_common.h"
#include "xdg/geometry/plucker.h"
#include "xdg/testing/mesh_mock.h"

using namespace xdg;

static std::vector<double> read_all(const char* path)
{
  FILE* f = std::fopen(path, "rb");
  if (!f) { std::perror(path); std::exit(2); }
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<double> buf(n / sizeof(double));
  if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
    std::fprintf(stderr, "short read\n");
    std::exit(2);
  }
  std::fclose(f);
  return buf;
}

static void write_all(const char* path, const std::vector<double>& out)
{
  FILE* f = std::fopen(path, "wb");
  if (!f) { std::perror(path); std::exit(2); }
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
}

static Vec3da v3(const double* p) { return Vec3da(p[0], p[1], p[2]); }

int main(int argc, char** argv)
{
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s mode in out\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  std::vector<double> in = read_all(argv[2]);
  std::vector<double> out;

  if (mode == "plucker") {
    for (size_t i = 0; i + 19 <= in.size(); i += 19) {
      const double* c = &in[i];
      Vec3da tri[3] = {v3(c), v3(c + 3), v3(c + 6)};
      auto r = plucker_ray_tri_intersect(tri, v3(c + 9), v3(c + 12), c[15],
                                         c[16], c[17] != 0.0, (int)c[18]);
      out.push_back(r.hit ? 1.0 : 0.0);
      out.push_back(r.t);
    }
  } else if (mode == "edge") {
    for (size_t i = 0; i + 12 <= in.size(); i += 12) {
      const double* c = &in[i];
      out.push_back(plucker_edge_test(v3(c), v3(c + 3), v3(c + 6), v3(c + 9)));
    }
  } else if (mode == "normal") {
    for (size_t i = 0; i + 9 <= in.size(); i += 9) {
      const double* c = &in[i];
      std::array<Vertex, 3> tri = {v3(c), v3(c + 3), v3(c + 6)};
      Direction n = triangle_normal(tri);
      out.push_back(n.x);
      out.push_back(n.y);
      out.push_back(n.z);
    }
  } else if (mode == "closest") {
    for (size_t i = 0; i + 12 <= in.size(); i += 12) {
      const double* c = &in[i];
      std::array<Position, 3> tri = {v3(c), v3(c + 3), v3(c + 6)};
      Position p = v3(c + 9);
      Region reg = determine_region(tri, p);
      Position q = closest_location_on_triangle(tri, p);
      out.push_back((double)static_cast<int>(reg));
      out.push_back(q.x);
      out.push_back(q.y);
      out.push_back(q.z);
    }
  } else if (mode == "mock_next") {
    auto mm = std::make_shared<MeshMock>();
    for (size_t i = 0; i + 7 <= in.size(); i += 7) {
      const double* c = &in[i];
      auto r = mm->next_element((MeshID)c[0], v3(c + 1), v3(c + 4));
      out.push_back((double)r.first);
      out.push_back(r.second);
    }
  } else if (mode == "mock_walk") {
    auto mm = std::make_shared<MeshMock>();
    for (size_t i = 0; i + 8 <= in.size(); i += 8) {
      const double* c = &in[i];
      auto segs = mm->walk_elements((MeshID)c[0], v3(c + 1), v3(c + 4), c[7]);
      out.push_back((double)segs.size());
      for (size_t j = 0; j < 64; ++j) {
        if (j < segs.size()) {
          out.push_back((double)segs[j].first);
          out.push_back(segs[j].second);
        } else {
          out.push_back(0.0);
          out.push_back(0.0);
        }
      }
    }
  } else if (mode == "mock_walk_se") {
    auto mm = std::make_shared<MeshMock>();
    for (size_t i = 0; i + 7 <= in.size(); i += 7) {
      const double* c = &in[i];
      auto segs = mm->walk_elements((MeshID)c[0], v3(c + 1), v3(c + 4));
      out.push_back((double)segs.size());
      for (size_t j = 0; j < 64; ++j) {
        if (j < segs.size()) {
          out.push_back((double)segs[j].first);
          out.push_back(segs[j].second);
        } else {
          out.push_back(0.0);
          out.push_back(0.0);
        }
      }
    }
  } else if (mode == "mock_face_normal") {
    auto mm = std::make_shared<MeshMock>();
    for (size_t i = 0; i < in.size(); ++i) {
      Direction n = mm->face_normal((MeshID)in[i]);
      out.push_back(n.x);
      out.push_back(n.y);
      out.push_back(n.z);
    }
  } else if (mode == "mock_tracks") {
    auto mm = std::make_shared<MeshMock>();
    auto bbox = mm->bounding_box();
    std::mt19937 gen(42);
    std::uniform_real_distribution<double> x_dist(bbox.min_x, bbox.max_x);
    std::uniform_real_distribution<double> y_dist(bbox.min_y, bbox.max_y);
    std::uniform_real_distribution<double> z_dist(bbox.min_z, bbox.max_z);
    int n = in.empty() ? 0 : (int)in[0];
    for (int i = 0; i < n; ++i) {
      Position start = {x_dist(gen), y_dist(gen), z_dist(gen)};
      Position end = {x_dist(gen), y_dist(gen), z_dist(gen)};
      out.push_back(start.x);
      out.push_back(start.y);
      out.push_back(start.z);
      out.push_back(end.x);
      out.push_back(end.y);
      out.push_back(end.z);
    }
  } else {
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  write_all(argv[3], out);
  return 0;
}
