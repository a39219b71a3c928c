// ray_order.hip -- coherence ordering of a ray batch before traversal.
//
// Incoherent rays leave most of a wave idle: a wave runs the node loop
// until its slowest lane is done, and rays of one batch differ in tree,
// start cell and direction.  Sorting the batch by (tree, direction octant,
// [coarse direction,] Morton code of the origin inside the tree's bounds)
// puts rays that walk similar paths into the same wave.  Traversal then
// reads rays through the permutation and writes results at the caller's
// index, so results are identical with and without ordering.
//
// No counterpart in the reference: Embree traces rays one at a time on the
// CPU (rtcIntersect1 per call, src/embree/ray_tracer.cpp:290-330).

#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"

using namespace xdgd;

namespace xdgh {
namespace {

__device__ __forceinline__ uint32_t spread3(uint32_t v)  // 10 bits -> every third bit of 30
{
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t quant10(double x, float lo, float hi)
{
  const double w = (double)hi - (double)lo;
  double q = w > 0.0 ? (x - (double)lo) / w * 1024.0 : 0.0;
  q = q >= 0.0 ? (q > 1023.0 ? 1023.0 : q) : 0.0;  // NaN -> 0
  return (uint32_t)q;
}

// key = [tree: tree_bits][octant: 3][coarse direction: dir_bits][origin Morton: rest]
__global__ void k_ray_keys(RayOrderArgs a, int tree_bits, int dir_bits, uint32_t* keys, int32_t* idx)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int32_t t = a.tree_per_ray ? a.tree_per_ray[i] : a.tree;
  if (t < 1 || t > a.ntrees) t = 0;
  const double ox = a.org[3 * i], oy = a.org[3 * i + 1], oz = a.org[3 * i + 2];
  const double dx = a.dir[3 * i], dy = a.dir[3 * i + 1], dz = a.dir[3 * i + 2];
  uint32_t oct = (dx < 0.0 ? 1u : 0u) | (dy < 0.0 ? 2u : 0u) | (dz < 0.0 ? 4u : 0u);
  uint32_t m = 0;
  if (t != 0) {
    const SurfTreeDev& T = a.table[t];
    m = spread3(quant10(ox, T.box_lo[0], T.box_hi[0])) | (spread3(quant10(oy, T.box_lo[1], T.box_hi[1])) << 1) |
        (spread3(quant10(oz, T.box_lo[2], T.box_hi[2])) << 2);
  }
  const int morton_bits = 29 - tree_bits - 3 * dir_bits;
  uint32_t key = ((uint32_t)t << (32 - tree_bits)) | (oct << (29 - tree_bits));
  if (dir_bits > 0) {
    // |d| quantised per axis into 2^dir_bits bins within the octant
    const double s = (double)(1u << dir_bits);
    uint32_t qx = (uint32_t)fmin(fabs(dx) * s, s - 1.0);
    uint32_t qy = (uint32_t)fmin(fabs(dy) * s, s - 1.0);
    uint32_t qz = (uint32_t)fmin(fabs(dz) * s, s - 1.0);
    const uint32_t q = (qx << (2 * dir_bits)) | (qy << dir_bits) | qz;
    key |= q << morton_bits;
  }
  key |= m >> (30 - morton_bits);
  keys[i] = key;
  idx[i] = (int32_t)i;
}

}  // namespace

size_t ray_order_temp_bytes(int64_t n) { return sizeof(uint32_t) * (size_t)radix_sort_hist_words(n); }

const int32_t* launch_ray_order(const RayOrderArgs& a, int mode, const RayOrderScratch& w, hipStream_t s)
{
  if (a.n <= 0) return nullptr;
  int tree_bits = 1;
  while ((1 << tree_bits) <= a.ntrees && tree_bits < 12) ++tree_bits;
  const int dir_bits = mode >= 2 ? 1 : 0;
  const unsigned grid = (unsigned)((a.n + 255) / 256);
  hipLaunchKernelGGL(k_ray_keys, dim3(grid), dim3(256), 0, s, a, tree_bits, dir_bits, w.keys_in, w.idx_in);
  // stable sort: rays with equal keys keep the caller's order
  radix_sort_pairs(w.keys_in, w.idx_in, w.keys_out, w.idx_out, a.n, 32, (uint32_t*)w.temp, s);
  return w.idx_in;
}

}  // namespace xdgh
