// A .h5m file to device trees: H5mMeshManager (include/xdg_h5m.hpp) ->
// flattening (include/xdg_hip_mesh.hpp) -> element tree, the implicit
// complement's surface tree and the boundary-face map -> one batch of
// XDG::segments (xdg_hip_tally_segments) on the GPU.  Writes each segment's
// element count (int32) and sequence hash (uint64) for
// tests/test_gpu_h5m.py to compare with the oracle on the original mesh.
//   usage: h5m_device_walk <file.h5m> <segments.f64: n x 6> <out.bin>
#include <cstdio>
#include <string>
#include <vector>

#include "xdg_h5m.hpp"

static std::vector<double> read_doubles(const char* path)
{
  std::vector<double> v;
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot read ") + path);
  double x;
  while (std::fread(&x, 8, 1, f) == 1) v.push_back(x);
  std::fclose(f);
  return v;
}

int main(int argc, char** argv)
{
  using namespace xdg_hip;
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <file.h5m> <segments.f64> <out.bin>\n", argv[0]);
    return 2;
  }
  try {
    H5mMeshManager mm(argv[1]);
    const MeshID vol = mm.volumes().front(), ipc = mm.implicit_complement();
    HipRayTracer rt(0, Precision::MIXED);
    const auto ea = flatten_elements(mm, {vol});
    const TreeID et = build_element_tree(rt, ea, MOAB_TET_FACES);
    const TreeID st = build_surface_tree(rt, flatten_volume_surfaces(mm, ipc), volume_bump(mm, ipc));
    std::vector<MeshID> faces;
    for (MeshID s : mm.get_volume_surfaces(ipc))
      for (MeshID f : mm.get_surface_faces(s)) faces.push_back(f);
    rt.attach_boundary(et, st, boundary_face_elements(mm, faces, ea));

    const std::vector<double> seg = read_doubles(argv[2]);
    const int64_t n = (int64_t)seg.size() / 6;
    std::vector<double> p0((size_t)(3 * n)), p1((size_t)(3 * n));
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        p0[(size_t)(3 * i + k)] = seg[(size_t)(6 * i + k)];
        p1[(size_t)(3 * i + k)] = seg[(size_t)(6 * i + 3 + k)];
      }
    xdg_hip_ctx* c = rt.handle();
    void *d0 = nullptr, *d1 = nullptr, *dc = nullptr, *dh = nullptr;
    auto ok = [&](int rc) {
      if (rc != XDG_HIP_OK) throw std::runtime_error(xdg_hip_last_error(c));
    };
    ok(xdg_hip_malloc(c, (size_t)(24 * n), &d0));
    ok(xdg_hip_malloc(c, (size_t)(24 * n), &d1));
    ok(xdg_hip_malloc(c, (size_t)(4 * n), &dc));
    ok(xdg_hip_malloc(c, (size_t)(8 * n), &dh));
    ok(xdg_hip_memcpy_htod(c, d0, p0.data(), (size_t)(24 * n)));
    ok(xdg_hip_memcpy_htod(c, d1, p1.data(), (size_t)(24 * n)));
    ok(xdg_hip_tally_segments(c, et, n, (const double*)d0, (const double*)d1, nullptr, (int32_t*)dc, (uint64_t*)dh,
                              nullptr, nullptr, nullptr, nullptr));
    ok(xdg_hip_synchronize(c, nullptr));
    std::vector<int32_t> cnt((size_t)n);
    std::vector<uint64_t> hsh((size_t)n);
    ok(xdg_hip_memcpy_dtoh(c, cnt.data(), dc, (size_t)(4 * n)));
    ok(xdg_hip_memcpy_dtoh(c, hsh.data(), dh, (size_t)(8 * n)));
    for (void* p : {d0, d1, dc, dh}) xdg_hip_free(c, p);
    uint32_t flags = 0;
    ok(xdg_hip_status(c, &flags));
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) throw std::runtime_error("cannot write output");
    std::fwrite(cnt.data(), 4, cnt.size(), f);
    std::fwrite(hsh.data(), 8, hsh.size(), f);
    std::fclose(f);
    std::printf("h5m device walk: %lld segments, status %u\n", (long long)n, flags);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "h5m_device_walk: %s\n", e.what());
    return 1;
  }
  return 0;
}
