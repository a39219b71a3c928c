// Reads a .h5m file through xdg_hip::H5mMeshManager (include/xdg_h5m.hpp),
// flattens its first volume and the implicit complement with
// include/xdg_hip_mesh.hpp, and writes the arrays as raw little-endian
// files into the output directory for tests/test_h5m_roundtrip.py to check
// (CPU only: no device calls).
//   usage: h5m_flatten_dump <file.h5m> <outdir>
#include <cstdio>
#include <string>
#include <vector>

#include "xdg_h5m.hpp"

template <class T>
static void dump(const std::string& path, const std::vector<T>& v)
{
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <file.h5m> <outdir>\n", argv[0]);
    return 2;
  }
  try {
    xdg_hip::H5mMeshManager mm(argv[1]);
    const std::string out = argv[2];
    const xdg_hip::MeshID vol = mm.volumes().front(), ipc = mm.implicit_complement();
    const auto ea = xdg_hip::flatten_elements(mm, {vol});
    const auto sv = xdg_hip::flatten_volume_surfaces(mm, vol);
    const auto si = xdg_hip::flatten_volume_surfaces(mm, ipc);
    std::vector<xdg_hip::MeshID> ipc_faces;
    for (auto s : mm.get_volume_surfaces(ipc))
      for (auto f : mm.get_surface_faces(s)) ipc_faces.push_back(f);
    dump(out + "/tets.i32", ea.tets);
    dump(out + "/adj.i32", ea.adj);
    dump(out + "/elem_ids.i32", ea.elem_ids);
    dump(out + "/verts.f64", ea.verts);
    dump(out + "/bmap.i32", xdg_hip::boundary_face_elements(mm, ipc_faces, ea));
    dump(out + "/ipc_faces.i32", ipc_faces);
    dump(out + "/vol_flip.u8", sv.flip);
    dump(out + "/ipc_flip.u8", si.flip);
    dump(out + "/vol_faces.i32", sv.face_ids);
    std::vector<double> bumps = {xdg_hip::volume_bump(mm, vol), xdg_hip::volume_bump(mm, ipc)};
    dump(out + "/bumps.f64", bumps);
    std::vector<int32_t> ids = {vol, ipc};
    dump(out + "/vol_ids.i32", ids);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "h5m_flatten_dump: %s\n", e.what());
    return 1;
  }
  return 0;
}
