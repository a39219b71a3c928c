// bvh_build.hip -- device LBVH builder.
//
// Replaces the Embree SAH build of EmbreeRayTracer::create_surface_tree /
// create_element_tree (src/embree/ray_tracer.cpp:50-188).  The reference
// dilates primitive boxes (triangle_intersect.cpp:36-50,
// tetrahedron_contain.cpp:47-63) so FP32 traversal stays conservative; this
// builder receives those dilated FP64 boxes and emits
//   * a BVH2 with both child boxes inline per node, FP64 (plain kernels) and
//     outward-rounded FP32 (mixed kernels), identical topology;
//   * the leaf-ordered primitive index array.
// Pipeline: centroid bounds (wave reduction + ordered-int atomics) ->
// 30-bit Morton codes -> rocPRIM radix sort of (code, index) pairs (stable,
// hence deterministic) -> Karras (HPG 2012) radix-tree hierarchy over the
// (code, index) keys -> bottom-up refit with atomic arrival flags -> node
// emission, collapsing subtrees of <= leaf_max primitives into leaves.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>

#include "internal.h"

namespace xdgh {

using namespace xdgd;

void* dev_alloc(size_t bytes, std::vector<DevBuf>& owned)
{
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  XDG_HIP_CHECK(hipMalloc(&p, bytes));
  owned.push_back(DevBuf{p, bytes});
  return p;
}

void dev_free_all(std::vector<DevBuf>& owned)
{
  for (auto& b : owned)
    if (b.p) (void)hipFree(b.p);
  owned.clear();
}

namespace {

constexpr int BLK = 256;

// monotone map double -> uint64 (ordered-int trick)
__device__ __forceinline__ unsigned long long ord_enc(double x)
{
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__device__ __forceinline__ double ord_dec(unsigned long long u)
{
  unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffULL) : ~u;
  return __longlong_as_double((long long)b);
}

__global__ void k_init_bounds(unsigned long long* cb)
{
  int i = threadIdx.x;
  if (i < 3) cb[i] = ULLONG_MAX;
  else if (i < 6) cb[i] = 0ULL;
}

__global__ void k_centroid_bounds(const double* __restrict__ box6, int64_t n, unsigned long long* cb)
{
  double mn[3] = {INFTY, INFTY, INFTY};
  double mx[3] = {-INFTY, -INFTY, -INFTY};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double* b = box6 + 6 * i;
    for (int k = 0; k < 3; ++k) {
      double c = 0.5 * (b[k] + b[3 + k]);
      mn[k] = fmin(mn[k], c);
      mx[k] = fmax(mx[k], c);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin(mn[k], __shfl_xor(mn[k], off));
      mx[k] = fmax(mx[k], __shfl_xor(mx[k], off));
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; ++k) {
      atomicMin(cb + k, ord_enc(mn[k]));
      atomicMax(cb + 3 + k, ord_enc(mx[k]));
    }
  }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_morton(const double* __restrict__ box6, int64_t n, const unsigned long long* cb,
                         uint32_t* keys, int32_t* vals)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double lo[3], ext[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = ord_dec(cb[k]);
    ext[k] = ord_dec(cb[3 + k]) - lo[k];
  }
  const double* b = box6 + 6 * i;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    double c = 0.5 * (b[k] + b[3 + k]);
    double f = ext[k] > 0.0 ? (c - lo[k]) / ext[k] : 0.0;
    double s = fmin(fmax(f * 1024.0, 0.0), 1023.0);
    q[k] = (uint32_t)s;
  }
  keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  vals[i] = (int32_t)i;
}

// Karras 2012 prefix length over the augmented keys (code, position)
__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int64_t n, int64_t i, int64_t j)
{
  if (j < 0 || j >= n) return -1;
  uint32_t a = k[i], b = k[j];
  if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
  return __clz((int)(a ^ b));
}

// children: internal node index, or ~leaf_position
__global__ void k_karras(const uint32_t* __restrict__ keys, int64_t n, int32_t* left, int32_t* right,
                         int32_t* rfirst, int32_t* rlast, int32_t* parent_int, int32_t* parent_leaf)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int64_t s = 0;
  int64_t div = 2;
  while (true) {
    int64_t t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
    div *= 2;
  }
  int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  int64_t lo = i < j ? i : j;
  int64_t hi = i < j ? j : i;
  int32_t lc, rc;
  if (lo == gamma) {
    lc = ~(int32_t)gamma;
    parent_leaf[gamma] = (int32_t)i;
  } else {
    lc = (int32_t)gamma;
    parent_int[gamma] = (int32_t)i;
  }
  if (hi == gamma + 1) {
    rc = ~(int32_t)(gamma + 1);
    parent_leaf[gamma + 1] = (int32_t)i;
  } else {
    rc = (int32_t)(gamma + 1);
    parent_int[gamma + 1] = (int32_t)i;
  }
  left[i] = lc;
  right[i] = rc;
  rfirst[i] = (int32_t)lo;
  rlast[i] = (int32_t)hi;
  if (i == 0) parent_int[0] = -1;
}

__device__ __forceinline__ void child_box(int32_t c, const double* __restrict__ box6, const int32_t* __restrict__ vals,
                                          const volatile double* ibox, double* out)
{
  if (c < 0) {
    const double* b = box6 + 6 * (int64_t)vals[~c];
    for (int k = 0; k < 6; ++k) out[k] = b[k];
  } else {
    for (int k = 0; k < 6; ++k) out[k] = ibox[6 * (int64_t)c + k];
  }
}

__global__ void k_refit(const double* __restrict__ box6, const int32_t* __restrict__ vals, int64_t n,
                        const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                        const int32_t* __restrict__ parent_int, const int32_t* __restrict__ parent_leaf,
                        unsigned int* flags, double* ibox)
{
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  int32_t p = parent_leaf[s];
  for (int guard = 0; p >= 0 && p < n - 1 && guard < 128; ++guard) {
    __threadfence();
    unsigned int old = atomicAdd(flags + p, 1u);
    if (old == 0) return;
    __threadfence();
    double a[6], b[6];
    child_box(left[p], box6, vals, ibox, a);
    child_box(right[p], box6, vals, ibox, b);
    volatile double* dst = ibox + 6 * (int64_t)p;
    for (int k = 0; k < 3; ++k) {
      dst[k] = fmin(a[k], b[k]);
      dst[3 + k] = fmax(a[3 + k], b[3 + k]);
    }
    p = parent_int[p];
  }
}

__device__ __forceinline__ float f32_down(double x)
{
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __forceinline__ float f32_up(double x)
{
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

__global__ void k_emit(const double* __restrict__ box6, const int32_t* __restrict__ vals, int64_t n, int leaf_max,
                       const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                       const int32_t* __restrict__ rfirst, const int32_t* __restrict__ rlast,
                       const double* __restrict__ ibox, NodeD* nd, NodeF* nf)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int64_t size = (int64_t)rlast[i] - rfirst[i] + 1;
  if (i != 0 && size <= leaf_max) return;  // inside a collapsed leaf
  NodeD d;
  NodeF f;
  int32_t ch[2] = {left[i], right[i]};
  for (int c = 0; c < 2; ++c) {
    double b[6];
    int32_t code;
    int32_t x = ch[c];
    if (x < 0) {
      const double* pb = box6 + 6 * (int64_t)vals[~x];
      for (int k = 0; k < 6; ++k) b[k] = pb[k];
      code = leaf_code(~x, 1);
    } else {
      for (int k = 0; k < 6; ++k) b[k] = ibox[6 * (int64_t)x + k];
      int64_t cs = (int64_t)rlast[x] - rfirst[x] + 1;
      code = cs <= leaf_max ? leaf_code(rfirst[x], (int32_t)cs) : x;
    }
    double* db = c == 0 ? d.b0 : d.b1;
    float* fb = c == 0 ? f.b0 : f.b1;
    for (int k = 0; k < 3; ++k) {
      db[k] = b[k];
      db[3 + k] = b[3 + k];
      fb[k] = f32_down(b[k]);
      fb[3 + k] = f32_up(b[3 + k]);
    }
    if (c == 0) {
      d.c0 = code;
      f.c0 = code;
    } else {
      d.c1 = code;
      f.c1 = code;
    }
  }
  d.pad0 = d.pad1 = 0;
  f.pad0 = f.pad1 = 0;
  nd[i] = d;
  nf[i] = f;
}

// depth of every emitted node = length of its parent chain
__global__ void k_depth(int64_t n, int leaf_max, const int32_t* __restrict__ rfirst,
                        const int32_t* __restrict__ rlast, const int32_t* __restrict__ parent_int, int* depth)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int64_t size = (int64_t)rlast[i] - rfirst[i] + 1;
  if (i != 0 && size <= leaf_max) return;
  int dd = 1;
  int32_t p = parent_int[i];
  while (p >= 0 && p < n - 1 && dd < 256) {
    ++dd;
    p = parent_int[p];
  }
  atomicMax(depth, dd);
}

__global__ void k_maxabs(const double* __restrict__ box6, int64_t n, unsigned long long* out)
{
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 6 * n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(box6[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_fill_u32(unsigned int* p, int64_t n, unsigned int v)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned grid_for(int64_t n, int blk = BLK)
{
  int64_t g = (n + blk - 1) / blk;
  return (unsigned)std::max<int64_t>(1, g);
}

// Host-side structural check of a freshly built tree, run before any
// traversal kernel can read it: every internal index in range and reached
// once, every leaf range inside [0, n), every primitive covered exactly once,
// the leaf order a permutation, depth within the 64-entry traversal stack.
void validate_tree(const LbvhOut& out, int64_t n)
{
  std::vector<int32_t> order((size_t)n);
  XDG_HIP_CHECK(hipMemcpy(order.data(), out.order, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  std::vector<uint8_t> seen((size_t)n, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (order[i] < 0 || order[i] >= n || seen[order[i]]) throw HipError{"lbvh: leaf order is not a permutation"};
    seen[order[i]] = 1;
  }
  std::fill(seen.begin(), seen.end(), 0);
  std::vector<NodeF> nodes((size_t)std::max<int64_t>(out.n_nodes, 0));
  if (out.n_nodes > 0)
    XDG_HIP_CHECK(hipMemcpy(nodes.data(), out.nodes_f, sizeof(NodeF) * out.n_nodes, hipMemcpyDeviceToHost));
  std::vector<uint8_t> visited(nodes.size(), 0);
  std::vector<std::pair<int32_t, int>> stack{{out.root, 1}};
  int64_t covered = 0;
  int maxd = 0;
  while (!stack.empty()) {
    auto [c, d] = stack.back();
    stack.pop_back();
    if (c < 0) {
      int64_t f = leaf_first(c), k = leaf_count(c);
      if (f < 0 || f + k > n) throw HipError{"lbvh: leaf range out of bounds"};
      for (int64_t j = f; j < f + k; ++j) {
        if (seen[j]) throw HipError{"lbvh: primitive covered twice"};
        seen[j] = 1;
      }
      covered += k;
      continue;
    }
    if (c >= out.n_nodes || visited[c]) throw HipError{"lbvh: bad or repeated internal node"};
    visited[c] = 1;
    maxd = std::max(maxd, d);
    if (d > 64) throw HipError{"lbvh: tree deeper than 64"};
    stack.push_back({nodes[c].c0, d + 1});
    stack.push_back({nodes[c].c1, d + 1});
  }
  if (covered != n) throw HipError{"lbvh: not every primitive is reachable"};
  if (out.root >= 0 && maxd != out.depth) throw HipError{"lbvh: depth mismatch"};
}

}  // namespace

void lbvh_build(const double* box6, int64_t n, int leaf_max, hipStream_t s, std::vector<DevBuf>& owned,
                LbvhOut& out)
{
  if (leaf_max < 1) leaf_max = 1;
  if (leaf_max > 8) leaf_max = 8;
  std::vector<DevBuf> tmp;
  try {
    out.order = (int32_t*)dev_alloc(sizeof(int32_t) * std::max<int64_t>(n, 1), owned);
    int64_t nn = std::max<int64_t>(n - 1, 1);
    out.nodes_d = (NodeD*)dev_alloc(sizeof(NodeD) * nn, owned);
    out.nodes_f = (NodeF*)dev_alloc(sizeof(NodeF) * nn, owned);
    out.n_nodes = n > 1 ? n - 1 : 0;
    if (n == 0) {
      out.root = leaf_code(0, 1);
      out.depth = 0;
      return;
    }
    // max |coordinate| (mixed-traversal safety limit)
    auto* d_max = (unsigned long long*)dev_alloc(sizeof(unsigned long long), tmp);
    XDG_HIP_CHECK(hipMemsetAsync(d_max, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_maxabs, dim3(std::min<unsigned>(grid_for(6 * n), 4096)), dim3(BLK), 0, s, box6, n, d_max);

    auto* cb = (unsigned long long*)dev_alloc(6 * sizeof(unsigned long long), tmp);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, cb);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(BLK), 0, s, box6, n,
                       cb);
    auto* keys = (uint32_t*)dev_alloc(sizeof(uint32_t) * n, tmp);
    auto* vals = (int32_t*)dev_alloc(sizeof(int32_t) * n, tmp);
    auto* keys2 = (uint32_t*)dev_alloc(sizeof(uint32_t) * n, tmp);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(BLK), 0, s, box6, n, cb, keys, vals);
    size_t tbytes = 0;
    XDG_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tbytes, keys, keys2, vals, out.order, (unsigned)n, 0, 30, s));
    void* tstore = dev_alloc(tbytes, tmp);
    XDG_HIP_CHECK(rocprim::radix_sort_pairs(tstore, tbytes, keys, keys2, vals, out.order, (unsigned)n, 0, 30, s));

    unsigned long long hmax = 0;
    if (n == 1) {
      out.root = leaf_code(0, 1);
      out.depth = 0;
    } else {
      int64_t ni = n - 1;
      auto* left = (int32_t*)dev_alloc(sizeof(int32_t) * ni, tmp);
      auto* right = (int32_t*)dev_alloc(sizeof(int32_t) * ni, tmp);
      auto* rfirst = (int32_t*)dev_alloc(sizeof(int32_t) * ni, tmp);
      auto* rlast = (int32_t*)dev_alloc(sizeof(int32_t) * ni, tmp);
      auto* pint = (int32_t*)dev_alloc(sizeof(int32_t) * ni, tmp);
      auto* pleaf = (int32_t*)dev_alloc(sizeof(int32_t) * n, tmp);
      auto* flags = (unsigned int*)dev_alloc(sizeof(unsigned int) * ni, tmp);
      auto* ibox = (double*)dev_alloc(sizeof(double) * 6 * ni, tmp);
      auto* d_depth = (int*)dev_alloc(sizeof(int), tmp);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ni)), dim3(BLK), 0, s, flags, ni, 0u);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ni)), dim3(BLK), 0, s, (unsigned int*)pint, ni, 0xffffffffu);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLK), 0, s, (unsigned int*)pleaf, n, 0xffffffffu);
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(BLK), 0, s, (unsigned int*)d_depth, 1, 0u);
      hipLaunchKernelGGL(k_karras, dim3(grid_for(ni)), dim3(BLK), 0, s, keys2, n, left, right, rfirst, rlast, pint,
                         pleaf);
      hipLaunchKernelGGL(k_refit, dim3(grid_for(n)), dim3(BLK), 0, s, box6, out.order, n, left, right, pint, pleaf,
                         flags, ibox);
      hipLaunchKernelGGL(k_emit, dim3(grid_for(ni)), dim3(BLK), 0, s, box6, out.order, n, leaf_max, left, right,
                         rfirst, rlast, ibox, out.nodes_d, out.nodes_f);
      hipLaunchKernelGGL(k_depth, dim3(grid_for(ni)), dim3(BLK), 0, s, n, leaf_max, rfirst, rlast, pint, d_depth);
      int depth = 0;
      XDG_HIP_CHECK(hipMemcpyAsync(&depth, d_depth, sizeof(int), hipMemcpyDeviceToHost, s));
      XDG_HIP_CHECK(hipMemcpyAsync(&hmax, d_max, sizeof(hmax), hipMemcpyDeviceToHost, s));
      XDG_HIP_CHECK(hipStreamSynchronize(s));
      out.depth = depth;
      out.root = n <= leaf_max ? leaf_code(0, (int32_t)n) : 0;
      if (n <= leaf_max) out.depth = 0;
    }
    if (n == 1) {
      XDG_HIP_CHECK(hipMemcpyAsync(&hmax, d_max, sizeof(hmax), hipMemcpyDeviceToHost, s));
      XDG_HIP_CHECK(hipStreamSynchronize(s));
    }
    double m;
    static_assert(sizeof(m) == sizeof(hmax), "size");
    memcpy(&m, &hmax, sizeof m);
    out.max_abs = m;
    XDG_HIP_CHECK(hipGetLastError());
    const char* ve = std::getenv("XDG_HIP_VALIDATE");
    if (!(ve && ve[0] == '0')) validate_tree(out, n);
  } catch (...) {
    dev_free_all(tmp);
    throw;
  }
  dev_free_all(tmp);
}

}  // namespace xdgh

// ---------------------------------------------------------------------------
// BVH4 collapse (host): each 4-wide node takes a binary node's two children
// and repeatedly opens the internal child with the largest surface area until
// it holds four children (or only leaves remain).  Nodes are emitted
// breadth-first so that the children of a node are contiguous.  Leaf codes
// are unchanged, so the BVH4 addresses the same leaf-ordered primitives.
// ---------------------------------------------------------------------------
namespace xdgh {

namespace {

struct Cand {
  int32_t code;
  double box[6];
  float fbox[6];
};

double half_area(const double* b)
{
  const double dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

}  // namespace

void bvh4_collapse(const LbvhOut& bin, int64_t n, std::vector<DevBuf>& owned, Bvh4Out& out)
{
  using namespace xdgd;
  const int32_t EMPTY = INT32_MIN;
  if (bin.root < 0 || bin.n_nodes == 0) {  // a single leaf: no nodes
    out.root = bin.root;
    out.stack = 0;
    out.n_nodes = 0;
    out.nodes_d = (Node4D*)dev_alloc(sizeof(Node4D), owned);
    out.nodes_f = (Node4F*)dev_alloc(sizeof(Node4F), owned);
    return;
  }
  std::vector<NodeD> bd((size_t)bin.n_nodes);
  std::vector<NodeF> bf((size_t)bin.n_nodes);
  XDG_HIP_CHECK(hipMemcpy(bd.data(), bin.nodes_d, sizeof(NodeD) * bin.n_nodes, hipMemcpyDeviceToHost));
  XDG_HIP_CHECK(hipMemcpy(bf.data(), bin.nodes_f, sizeof(NodeF) * bin.n_nodes, hipMemcpyDeviceToHost));

  std::vector<Node4D> od;
  std::vector<Node4F> of;
  od.reserve((size_t)bin.n_nodes / 2 + 1);
  of.reserve((size_t)bin.n_nodes / 2 + 1);
  // work item: binary node index -> BVH4 slot to fill, and the stack entries
  // pushed above it on the worst path
  struct Item {
    int32_t bnode;
    int32_t slot;
    int32_t above;
  };
  std::vector<Item> queue;
  od.emplace_back();
  of.emplace_back();
  queue.push_back({bin.root, 0, 0});
  int32_t max_stack = 0;
  for (size_t q = 0; q < queue.size(); ++q) {
    const Item it = queue[q];
    const NodeD& b = bd[it.bnode];
    const NodeF& f = bf[it.bnode];
    std::vector<Cand> cs(2);
    cs[0].code = b.c0;
    cs[1].code = b.c1;
    for (int k = 0; k < 6; ++k) {
      cs[0].box[k] = b.b0[k];
      cs[1].box[k] = b.b1[k];
      cs[0].fbox[k] = f.b0[k];
      cs[1].fbox[k] = f.b1[k];
    }
    while (cs.size() < 4) {
      int best = -1;
      double area = -1.0;
      for (int j = 0; j < (int)cs.size(); ++j)
        if (cs[j].code >= 0 && half_area(cs[j].box) > area) {
          area = half_area(cs[j].box);
          best = j;
        }
      if (best < 0) break;
      const NodeD& cb = bd[cs[best].code];
      const NodeF& cf = bf[cs[best].code];
      Cand a, c;
      a.code = cb.c0;
      c.code = cb.c1;
      for (int k = 0; k < 6; ++k) {
        a.box[k] = cb.b0[k];
        c.box[k] = cb.b1[k];
        a.fbox[k] = cf.b0[k];
        c.fbox[k] = cf.b1[k];
      }
      cs[best] = a;
      cs.push_back(c);
    }
    const int32_t nc = (int32_t)cs.size();
    const int32_t above = it.above + (nc - 1);
    max_stack = std::max(max_stack, above);
    Node4D nd;
    Node4F nf;
    std::memset((void*)&nd, 0, sizeof nd);
    std::memset((void*)&nf, 0, sizeof nf);
    for (int j = 0; j < 4; ++j) {
      if (j < nc) {
        nd.lox[j] = cs[j].box[0];
        nd.loy[j] = cs[j].box[1];
        nd.loz[j] = cs[j].box[2];
        nd.hix[j] = cs[j].box[3];
        nd.hiy[j] = cs[j].box[4];
        nd.hiz[j] = cs[j].box[5];
        nf.lox[j] = cs[j].fbox[0];
        nf.loy[j] = cs[j].fbox[1];
        nf.loz[j] = cs[j].fbox[2];
        nf.hix[j] = cs[j].fbox[3];
        nf.hiy[j] = cs[j].fbox[4];
        nf.hiz[j] = cs[j].fbox[5];
      } else {
        // inverted box: the near-plane entry distance is +inf for every ray
        nd.lox[j] = nd.loy[j] = nd.loz[j] = HUGE_VAL;
        nd.hix[j] = nd.hiy[j] = nd.hiz[j] = -HUGE_VAL;
        nf.lox[j] = nf.loy[j] = nf.loz[j] = HUGE_VALF;
        nf.hix[j] = nf.hiy[j] = nf.hiz[j] = -HUGE_VALF;
        nd.child[j] = nf.child[j] = EMPTY;
      }
    }
    for (int j = 0; j < nc; ++j) {
      if (cs[j].code >= 0) {
        const int32_t slot = (int32_t)od.size();
        od.emplace_back();
        of.emplace_back();
        queue.push_back({cs[j].code, slot, above});
        nd.child[j] = nf.child[j] = slot;
      } else {
        nd.child[j] = nf.child[j] = cs[j].code;
      }
    }
    od[it.slot] = nd;
    of[it.slot] = nf;
  }
  // structural check before any kernel reads the BVH4: child indices in
  // range, every primitive covered exactly once
  {
    std::vector<uint8_t> seen((size_t)n, 0);
    int64_t covered = 0;
    for (const Node4D& nd : od) {
      for (int j = 0; j < 4; ++j) {
        const int32_t c = nd.child[j];
        if (c == EMPTY) {
          if (!(nd.lox[j] == HUGE_VAL && nd.hix[j] == -HUGE_VAL)) throw HipError{"bvh4: empty slot not inverted"};
          continue;
        }
        if (c >= 0) {
          if ((size_t)c >= od.size()) throw HipError{"bvh4: child index out of range"};
          continue;
        }
        const int64_t f = leaf_first(c), k = leaf_count(c);
        if (f < 0 || f + k > n) throw HipError{"bvh4: leaf range out of bounds"};
        for (int64_t p = f; p < f + k; ++p) {
          if (seen[p]) throw HipError{"bvh4: primitive covered twice"};
          seen[p] = 1;
        }
        covered += k;
      }
    }
    if (covered != n) throw HipError{"bvh4: not every primitive is reachable"};
  }
  // the stack never holds more than the pushes along one root-to-leaf path
  out.stack = std::max(max_stack, 1);
  out.root = 0;
  out.n_nodes = (int64_t)od.size();
  out.nodes_d = (Node4D*)dev_alloc(sizeof(Node4D) * od.size(), owned);
  out.nodes_f = (Node4F*)dev_alloc(sizeof(Node4F) * of.size(), owned);
  XDG_HIP_CHECK(hipMemcpy(out.nodes_d, od.data(), sizeof(Node4D) * od.size(), hipMemcpyHostToDevice));
  XDG_HIP_CHECK(hipMemcpy(out.nodes_f, of.data(), sizeof(Node4F) * of.size(), hipMemcpyHostToDevice));
  (void)n;
}

}  // namespace xdgh
