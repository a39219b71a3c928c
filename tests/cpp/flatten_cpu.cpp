// CPU-only checks of include/xdg_hip_mesh.hpp (no device calls): the
// arrays flattened from the test managers.  Exits 0 when every check passes.
#include <cstdio>

#include "test_managers.h"

using namespace xdg_hip;

static int failures = 0;
#define CHECK(cond)                                                 \
  do {                                                              \
    if (!(cond)) {                                                  \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      ++failures;                                                   \
    }                                                               \
  } while (0)

int main()
{
  BoxManager box;
  // bump: volume box [-2,5]x[-3,6]x[-4,7] (already contains the origin),
  // chord sqrt(7^2+9^2+11^2) * 1e-6 < 1e-3
  CHECK(volume_bump(box, BoxManager::VOL) == 1e-3);
  const auto sv = flatten_volume_surfaces(box, BoxManager::VOL);
  const auto si = flatten_volume_surfaces(box, BoxManager::IPC);
  const auto sg = flatten_surfaces(box, {10, 11, 12, 13, 14, 15}, ID_NONE);
  CHECK(sv.face_ids.size() == 12 && sv.tris.size() == 36 && sv.verts.size() == 108);
  bool flips_ok = true;
  for (int i = 0; i < 12; ++i) {
    flips_ok &= sv.flip[i] == 0 && si.flip[i] == 1 && sg.flip[i] == 0;
    flips_ok &= sv.face_ids[i] == 1000 + i && sv.surf_ids[i] == 10 + i / 2 && si.face_ids[i] == 1000 + i;
    flips_ok &= sv.tris[3 * i] == 3 * i && sv.tris[3 * i + 2] == 3 * i + 2;
  }
  CHECK(flips_ok);
  // face 1004 = mock triangle 4 = vertices 0, 5, 1: (5,-3,7) (5,6,-4) (5,6,7)
  CHECK(sv.verts[36] == 5 && sv.verts[37] == -3 && sv.verts[38] == 7 && sv.verts[40] == 6 && sv.verts[41] == -4);
  bool threw = false;
  try {
    flatten_surfaces(box, {10}, 42);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  CHECK(threw);

  const auto ea = flatten_elements(box, {BoxManager::VOL});
  CHECK(ea.elem_ids.size() == 12 && ea.tets.size() == 48 && ea.verts.size() == 27);
  bool adj_ok = true;
  for (int i = 0; i < 12; ++i)
    for (int f = 0; f < 4; ++f) {
      const MeshID n = box.adjacent_element(100 + i, f);
      adj_ok &= ea.adj[4 * i + f] == (n == ID_NONE ? -1 : n - 100);
    }
  CHECK(adj_ok);
  CHECK(ea.tets[0] == 0 && ea.tets[3] == 8);  // tet 0 = mock triangle 0 + centre
  // tets 8..11 are the y = -3 pair, then the y = 6 pair (the mock's order)
  CHECK(ea.tets[4 * 8] == 0 && ea.tets[4 * 8 + 1] == 7 && ea.tets[4 * 10] == 1 && ea.tets[4 * 10 + 1] == 5);
  // every neighbour pair shares a face: three common vertices
  bool shared = true;
  for (int e = 0; e < 12; ++e)
    for (int f = 0; f < 4; ++f) {
      const int n = ea.adj[4 * e + f];
      if (n < 0) continue;
      int common = 0;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) common += ea.tets[4 * e + i] == ea.tets[4 * n + j];
      shared &= common == 3;
    }
  CHECK(shared);
  // two volumes: duplicated vertex blocks, connectivity offset by the first
  BoxManager split(true);
  const auto ga = flatten_elements(split, {BoxManager::VOL, BoxManager::VOL_B});
  const auto ha = flatten_elements(split, {BoxManager::VOL});
  CHECK(ga.elem_ids.size() == 12 && ha.elem_ids.size() == 6);
  CHECK(ga.verts.size() == 3 * (split.get_volume_vertices(1).size() + split.get_volume_vertices(3).size()));
  bool same = true;  // the same tets, coordinate for coordinate
  for (int e = 0; e < 12; ++e)
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c)
        same &= ga.verts[3 * ga.tets[4 * e + k] + c] == ea.verts[3 * ea.tets[4 * e + k] + c];
  CHECK(same);
  CHECK(ga.adj == ea.adj);
  int cut = 0;
  for (int32_t x : ha.adj) cut += x < 0;
  CHECK(cut > 6);
  bool dup = false;
  try {
    flatten_elements(box, {BoxManager::VOL, BoxManager::VOL});
  } catch (const std::invalid_argument&) {
    dup = true;
  }
  CHECK(dup);

  std::vector<MeshID> faces;
  for (int f = 1000; f < 1012; ++f) faces.push_back(f);
  const auto bm = boundary_face_elements(box, faces, ea);
  CHECK(bm.size() == 1012 && bm[0] == -1 && bm[1000] == 0 && bm[1007] == 7);
  CHECK(bm[1008] == 10 && bm[1009] == 11 && bm[1010] == 8 && bm[1011] == 9);
  const auto bh = boundary_face_elements(split, faces, ha);
  CHECK(bh[1005] == 5 && bh[1006] == -1);  // face 1006's owner is in the other volume

  KuhnManager kuhn(3);
  const auto ka = flatten_elements(kuhn, {1});
  CHECK(ka.elem_ids.size() == 162);
  int boundary = 0;
  for (int32_t x : ka.adj) boundary += x < 0;
  CHECK(boundary == 6 * 9 * 2);  // two triangles per boundary square, 9 squares per side

  if (failures == 0) std::printf("flatten (cpu): all checks passed\n");
  return failures == 0 ? 0 : 1;
}
