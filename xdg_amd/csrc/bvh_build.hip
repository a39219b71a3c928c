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
// Pipeline, all on the device: centroid bounds (wave reduction +
// ordered-int atomics) -> 63-bit Morton codes -> stable LSD radix sort of
// (code, index) pairs (4-bit digits, hand-written below; stable, hence
// deterministic) -> Karras (HPG 2012) radix-tree hierarchy over the
// (code, index) keys -> bottom-up refit with atomic arrival flags -> node
// emission, collapsing subtrees of <= leaf_max primitives into leaves ->
// structural check by breadth-first levels.  The BVH4 the traversal kernels
// walk is collapsed from that tree level by level on the device as well.
// Only counts and the max |coordinate| travel to the host.
#include <cstring>

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


// spread the low 21 bits of v to every third bit
__device__ __forceinline__ uint64_t expand21(uint64_t v)
{
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// 63-bit Morton code of each box centroid over the centroid bounds
__global__ void k_morton(const double* __restrict__ box6, int64_t n, const unsigned long long* cb, uint64_t* keys,
                         int32_t* vals)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double lo[3], ext[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = ord_dec(cb[k]);
    ext[k] = ord_dec(cb[3 + k]) - lo[k];
  }
  const double* b = box6 + 6 * i;
  uint64_t q[3];
  for (int k = 0; k < 3; ++k) {
    double c = 0.5 * (b[k] + b[3 + k]);
    double f = ext[k] > 0.0 ? (c - lo[k]) / ext[k] : 0.0;
    double s = fmin(fmax(f * 2097152.0, 0.0), 2097151.0);
    q[k] = (uint64_t)s;
  }
  keys[i] = (expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]);
  vals[i] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint64 key, int32 value) pairs, 4-bit digits.
// Per pass: per-tile digit counts, one exclusive scan over (digit, tile),
// then a stable scatter: a tile is staged in LDS, each thread ranks its 8
// consecutive items against the counts of all lower threads (16-bit fields
// packed 4 to a word, scanned across the workgroup with wave shuffles).
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 16;

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                       uint32_t* hist, int64_t ntiles)
{
  __shared__ uint32_t c[RS_BINS];
  if (threadIdx.x < RS_BINS) c[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = blockIdx.x * (int64_t)RS_TILE;
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int64_t i = base + k * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&c[(keys[i] >> shift) & 15u], 1u);
  }
  __syncthreads();
  if (threadIdx.x < RS_BINS) hist[threadIdx.x * ntiles + blockIdx.x] = c[threadIdx.x];
}

// exclusive scan of a[0..m) in place (one workgroup of 1024), total in *total
__global__ void __launch_bounds__(1024) k_scan_excl(uint32_t* a, int64_t m, uint32_t* total)
{
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t base = 0; base < m; base += 1024 * 4) {
    uint32_t v[4], sum = 0;
    for (int k = 0; k < 4; ++k) {
      const int64_t i = base + threadIdx.x * 4 + k;
      v[k] = i < m ? a[i] : 0u;
      sum += v[k];
    }
    uint32_t x = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (wv == 0) {
      uint32_t w = lane < 16 ? wsum[lane] : 0u;
      for (int off = 1; off < 16; off <<= 1) {
        const uint32_t y = __shfl_up(w, off);
        if (lane >= off) w += y;
      }
      if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    uint32_t e = carry + (wv ? wsum[wv - 1] : 0u) + x - sum;
    for (int k = 0; k < 4; ++k) {
      const int64_t i = base + threadIdx.x * 4 + k;
      if (i < m) a[i] = e;
      e += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[15];
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__device__ __forceinline__ uint32_t field16(const uint64_t* w, uint32_t d)
{
  return (uint32_t)((w[d >> 2] >> (16 * (d & 3))) & 0xffffull);
}

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K* __restrict__ kin, const int32_t* __restrict__ vin,
                                                          K* kout, int32_t* vout, int64_t n, int shift,
                                                          const uint32_t* __restrict__ off, int64_t ntiles)
{
  __shared__ K sk[RS_TILE];
  __shared__ int32_t sv[RS_TILE];
  __shared__ uint64_t wtot[RS_THREADS / 64][4];
  const int64_t base = blockIdx.x * (int64_t)RS_TILE;
  const int64_t cnt = n - base < RS_TILE ? n - base : RS_TILE;
  for (int k = 0; k < RS_ITEMS; ++k) {  // coalesced loads, staged in LDS
    const int j = k * RS_THREADS + threadIdx.x;
    if (j < cnt) {
      sk[j] = kin[base + j];
      sv[j] = vin[base + j];
    }
  }
  __syncthreads();
  // this thread's items: tile positions t*8 .. t*8+7, in order
  const int first = threadIdx.x * RS_ITEMS;
  uint64_t c[4] = {0, 0, 0, 0};
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int j = first + k;
    if (j < cnt) {
      const uint32_t d = (uint32_t)((sk[j] >> shift) & 15u);
      c[d >> 2] += 1ull << (16 * (d & 3));
    }
  }
  // exclusive scan of the packed counts over the workgroup's threads
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t x[4] = {c[0], c[1], c[2], c[3]};
  for (int off2 = 1; off2 < 64; off2 <<= 1)
    for (int w = 0; w < 4; ++w) {
      const uint64_t y = __shfl_up(x[w], off2);
      if (lane >= off2) x[w] += y;
    }
  if (lane == 63)
    for (int w = 0; w < 4; ++w) wtot[wv][w] = x[w];
  __syncthreads();
  uint64_t ex[4];
  for (int w = 0; w < 4; ++w) {
    uint64_t before = 0;
    for (int q = 0; q < wv; ++q) before += wtot[q][w];
    ex[w] = before + x[w] - c[w];
  }
  uint64_t run[4] = {0, 0, 0, 0};
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int j = first + k;
    if (j >= cnt) break;
    const K key = sk[j];
    const uint32_t d = (uint32_t)((key >> shift) & 15u);
    const int64_t pos = (int64_t)off[d * ntiles + blockIdx.x] + field16(ex, d) + field16(run, d);
    run[d >> 2] += 1ull << (16 * (d & 3));
    kout[pos] = key;
    vout[pos] = sv[j];
  }
}

template <class K>
void radix_sort_impl(K* keys, int32_t* vals, K* k2, int32_t* v2, int64_t n, int bits, uint32_t* hist, hipStream_t s)
{
  if (n <= 1) return;
  const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  int passes = (bits + 3) / 4;
  if (passes & 1) ++passes;  // even: the result lands back in keys/vals
  K *ka = keys, *kb = k2;
  int32_t *va = vals, *vb = v2;
  for (int p = 0; p < passes; ++p) {
    const int shift = 4 * p;
    hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka, n, shift, hist, ntiles);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, hist, (int64_t)RS_BINS * ntiles, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka, va, kb, vb, n, shift,
                       hist, ntiles);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
}

// Karras 2012 prefix length over the augmented keys (code, position)
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j)
{
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
  return __clzll((long long)(a ^ b));
}

// children: internal node index, or ~leaf_position
__global__ void k_karras(const uint64_t* __restrict__ keys, int64_t n, int32_t* left, int32_t* right,
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

// ---------------------------------------------------------------------------
// Binned-SAH topology, the alternative to the Karras hierarchy: top down, one
// level of tasks at a time.  A task is an internal node whose primitives
// occupy the positions [start, start + cnt) of the order array; one
// workgroup per task bins the centroids into SAH_BINS bins per axis (counts
// and FP32 primitive bounds, LDS atomics, so the result does not depend on
// the order of the atomics), takes the split of least
// area(left) * n_left + area(right) * n_right (ties: lower axis, then lower
// bin) and partitions its range stably in place.  Children of two or more
// primitives are the next level's tasks, and node ids are handed out in task
// order by a scan, so the tree is deterministic.  Tasks of more than
// SAH_MORTON_MIN primitives split where the Karras tree would, at the first
// key of their range with its highest differing Morton bit set: one workgroup
// would otherwise bin a whole top-level range alone, and the SAH near the
// root changes little.  Those ranges are never reordered, so they still
// hold the sorted keys.  Tasks whose centroids no bin boundary separates, and
// every task from SAH_MEDIAN_LEVEL on, split at the object median of their
// current order (which keeps the depth within the 64 levels the traversal
// stacks allow).  The result takes the Karras
// kernel's form -- children as internal ids or ~position, ranges, parents --
// so refit, emission, validation and the BVH4 collapse are shared.
// ---------------------------------------------------------------------------
constexpr int SAH_BINS = 32;
constexpr int SAH_THREADS = 256;
constexpr int SAH_MEDIAN_LEVEL = 36;
constexpr int SAH_MORTON_MIN = 32768;

struct SahTask {
  int32_t node, start, cnt, pad;
};

// monotone map float -> uint32 (ordered-int trick)
__device__ __forceinline__ unsigned int ordf(float x)
{
  const unsigned int b = __float_as_uint(x);
  return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unordf(unsigned int u)
{
  const unsigned int b = (u >> 31) ? (u & 0x7fffffffu) : ~u;
  return __uint_as_float(b);
}

__device__ __forceinline__ int sah_bin(double c, double lo, double scale)
{
  const int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b >= SAH_BINS ? SAH_BINS - 1 : b);
}

__device__ __forceinline__ float half_area(const float* b)
{
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

__global__ void k_sah_root(SahTask* tasks, int32_t* rfirst, int32_t* rlast, int32_t* pint, int64_t n)
{
  tasks[0] = SahTask{0, 0, (int32_t)n, 0};
  rfirst[0] = 0;
  rlast[0] = (int32_t)(n - 1);
  pint[0] = -1;
}

// Split position of the Morton-sorted range [a, b) (b - a >= 2): the first
// key with the highest bit in which keys[a] and keys[b - 1] differ, or the
// middle when they are equal.
__device__ int32_t morton_split(const uint64_t* __restrict__ keys, int32_t a, int32_t b)
{
  const uint64_t x = keys[a] ^ keys[b - 1];
  if (x == 0) return a + (b - a) / 2;
  const uint64_t bit = 1ull << (63 - __clzll((long long)x));
  int32_t lo = a + 1, hi = b - 1;  // keys[a] has the bit clear, keys[b - 1] set
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] & bit)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

__global__ __launch_bounds__(SAH_THREADS) void k_sah_split(const double* __restrict__ box6,
                                                          const uint64_t* __restrict__ keys, int32_t* idx,
                                                          int32_t* tmp, const SahTask* __restrict__ tasks,
                                                          int32_t* nleft, int median)
{
  __shared__ double red[6][SAH_THREADS / 64];
  __shared__ double cb[6];
  __shared__ unsigned int lcnt[3][SAH_BINS];
  __shared__ unsigned int lbox[3][SAH_BINS][6];
  __shared__ float cost[SAH_THREADS];
  __shared__ int cand[SAH_THREADS];
  __shared__ unsigned int wsum[SAH_THREADS / 64];
  __shared__ int s_axis, s_k, s_nl;
  const SahTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (t.cnt <= 2 || median) {
    if (tid == 0) nleft[blockIdx.x] = t.cnt / 2;
    return;
  }
  if (t.cnt > SAH_MORTON_MIN) {
    if (tid == 0) nleft[blockIdx.x] = morton_split(keys, t.start, t.start + t.cnt) - t.start;
    return;
  }
  // centroid bounds of the task
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int p = tid; p < t.cnt; p += SAH_THREADS) {
    const double* b = box6 + 6 * (int64_t)idx[t.start + p];
    for (int k = 0; k < 3; ++k) {
      const double c = 0.5 * (b[k] + b[3 + k]);
      mn[k] = fmin(mn[k], c);
      mx[k] = fmax(mx[k], c);
    }
  }
  for (int k = 0; k < 3; ++k)
    for (int off = 32; off > 0; off >>= 1) {
      mn[k] = fmin(mn[k], __shfl_xor(mn[k], off));
      mx[k] = fmax(mx[k], __shfl_xor(mx[k], off));
    }
  if (lane == 0)
    for (int k = 0; k < 3; ++k) {
      red[k][wv] = mn[k];
      red[3 + k][wv] = mx[k];
    }
  for (int i = tid; i < 3 * SAH_BINS; i += SAH_THREADS) {
    lcnt[i / SAH_BINS][i % SAH_BINS] = 0u;
    for (int j = 0; j < 3; ++j) {
      lbox[i / SAH_BINS][i % SAH_BINS][j] = ordf(INFINITY);
      lbox[i / SAH_BINS][i % SAH_BINS][3 + j] = ordf(-INFINITY);
    }
  }
  __syncthreads();
  if (tid < 6) {
    double v = red[tid][0];
    for (int w = 1; w < SAH_THREADS / 64; ++w) v = tid < 3 ? fmin(v, red[tid][w]) : fmax(v, red[tid][w]);
    cb[tid] = v;
  }
  __syncthreads();
  double lo[3], scale[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = cb[k];
    const double ext = cb[3 + k] - cb[k];
    scale[k] = ext > 0.0 ? SAH_BINS / ext : 0.0;
  }
  // bin counts and FP32 bounds (the cost needs no more)
  for (int p = tid; p < t.cnt; p += SAH_THREADS) {
    const double* b = box6 + 6 * (int64_t)idx[t.start + p];
    unsigned int fb[6];
    for (int j = 0; j < 6; ++j) fb[j] = ordf((float)b[j]);
    for (int k = 0; k < 3; ++k) {
      const int bin = sah_bin(0.5 * (b[k] + b[3 + k]), lo[k], scale[k]);
      atomicAdd(&lcnt[k][bin], 1u);
      for (int j = 0; j < 3; ++j) {
        atomicMin(&lbox[k][bin][j], fb[j]);
        atomicMax(&lbox[k][bin][3 + j], fb[3 + j]);
      }
    }
  }
  __syncthreads();
  // one candidate per thread: axis ax, left = bins [0, k)
  float cst = INFINITY;
  if (tid < 3 * (SAH_BINS - 1)) {
    const int ax = tid / (SAH_BINS - 1), k = tid % (SAH_BINS - 1) + 1;
    if (scale[ax] > 0.0) {
      unsigned int nl = 0, nr = 0;
      float l[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      for (int b = 0; b < SAH_BINS; ++b) {
        const unsigned int c = lcnt[ax][b];
        if (!c) continue;
        float* d = b < k ? l : r;
        if (b < k)
          nl += c;
        else
          nr += c;
        for (int j = 0; j < 3; ++j) {
          d[j] = fminf(d[j], unordf(lbox[ax][b][j]));
          d[3 + j] = fmaxf(d[3 + j], unordf(lbox[ax][b][3 + j]));
        }
      }
      if (nl && nr) cst = half_area(l) * (float)nl + half_area(r) * (float)nr;
    }
  }
  cost[tid] = cst;
  cand[tid] = tid;
  __syncthreads();
  for (int st = SAH_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      const float a = cost[tid], b = cost[tid + st];
      if (b < a || (b == a && cand[tid + st] < cand[tid])) {
        cost[tid] = b;
        cand[tid] = cand[tid + st];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (cost[0] < INFINITY) {
      const int ax = cand[0] / (SAH_BINS - 1), k = cand[0] % (SAH_BINS - 1) + 1;
      unsigned int nl = 0;
      for (int b = 0; b < k; ++b) nl += lcnt[ax][b];
      s_axis = ax;
      s_k = k;
      s_nl = (int)nl;
    } else {
      s_axis = -1;
      s_k = 0;
      s_nl = t.cnt / 2;
    }
    nleft[blockIdx.x] = s_nl;
  }
  __syncthreads();
  const int ax = s_axis, ks = s_k, nl = s_nl;
  if (ax < 0) return;  // no bin boundary separates the centroids: the object median, order kept
  // stable partition of the range into tmp, then back
  int lrun = 0;
  for (int base = 0; base < t.cnt; base += SAH_THREADS) {
    const int pr = base + tid;
    int prim = 0, f = 0;
    if (pr < t.cnt) {
      prim = idx[t.start + pr];
      const double* b = box6 + 6 * (int64_t)prim;
      f = sah_bin(0.5 * (b[ax] + b[3 + ax]), lo[ax], scale[ax]) < ks ? 1 : 0;
    }
    const unsigned long long bal = __ballot(f);
    const int wpre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = (unsigned int)__popcll(bal);
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < SAH_THREADS / 64; ++w) {
      if (w < wv) before += (int)wsum[w];
      tot += (int)wsum[w];
    }
    if (pr < t.cnt) {
      const int lpre = lrun + before + wpre;
      tmp[t.start + (f ? lpre : nl + (pr - lpre))] = prim;
    }
    lrun += tot;
    __syncthreads();
  }
  for (int pr = tid; pr < t.cnt; pr += SAH_THREADS) idx[t.start + pr] = tmp[t.start + pr];
}

__global__ void k_sah_count(const SahTask* __restrict__ tasks, int64_t nt, const int32_t* __restrict__ nleft,
                            uint32_t* ci)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int32_t nl = nleft[i], nr = tasks[i].cnt - nl;
  ci[i] = (nl >= 2 ? 1u : 0u) + (nr >= 2 ? 1u : 0u);
}

__global__ void k_sah_link(const SahTask* __restrict__ tasks, int64_t nt, const int32_t* __restrict__ nleft,
                           const uint32_t* __restrict__ off, int32_t next_id, int32_t* left, int32_t* right,
                           int32_t* rfirst, int32_t* rlast, int32_t* pint, int32_t* pleaf, SahTask* out)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const SahTask t = tasks[i];
  const int32_t nl = nleft[i], nr = t.cnt - nl;
  uint32_t o = off[i];
  int32_t id = next_id + (int32_t)o;
  if (nl >= 2) {
    left[t.node] = id;
    pint[id] = t.node;
    rfirst[id] = t.start;
    rlast[id] = t.start + nl - 1;
    out[o++] = SahTask{id, t.start, nl, 0};
    ++id;
  } else {
    left[t.node] = ~t.start;
    pleaf[t.start] = t.node;
  }
  const int32_t rs = t.start + nl;
  if (nr >= 2) {
    right[t.node] = id;
    pint[id] = t.node;
    rfirst[id] = rs;
    rlast[id] = rs + nr - 1;
    out[o] = SahTask{id, rs, nr, 0};
  } else {
    right[t.node] = ~rs;
    pleaf[rs] = t.node;
  }
}

// Binned-SAH hierarchy over the n >= 2 primitives listed in order (Morton
// order on entry, with their sorted keys; the leaf order on return), into
// the topology arrays t.
void sah_topology(const double* box6, const uint64_t* keys, int64_t n, int32_t* order, const LbvhTopo& t,
                  hipStream_t s, std::vector<DevBuf>& tmp)
{
  const int64_t cap = n / 2 + 1;  // tasks hold >= 2 primitives each, disjoint
  auto* ta = (SahTask*)dev_alloc(sizeof(SahTask) * cap, tmp);
  auto* tb = (SahTask*)dev_alloc(sizeof(SahTask) * cap, tmp);
  auto* nl = (int32_t*)dev_alloc(sizeof(int32_t) * cap, tmp);
  auto* ci = (uint32_t*)dev_alloc(sizeof(uint32_t) * cap, tmp);
  auto* tot = (uint32_t*)dev_alloc(sizeof(uint32_t), tmp);
  auto* scratch = (int32_t*)dev_alloc(sizeof(int32_t) * n, tmp);
  hipLaunchKernelGGL(k_sah_root, dim3(1), dim3(1), 0, s, ta, t.rfirst, t.rlast, t.pint, n);
  int64_t nt = 1;
  int32_t next_id = 1;
  for (int level = 0; nt > 0; ++level) {
    hipLaunchKernelGGL(k_sah_split, dim3((unsigned)nt), dim3(SAH_THREADS), 0, s, box6, keys, order, scratch, ta,
                       nl, level >= SAH_MEDIAN_LEVEL ? 1 : 0);
    hipLaunchKernelGGL(k_sah_count, dim3(grid_for(nt)), dim3(BLK), 0, s, ta, nt, nl, ci);
    launch_scan_excl(ci, nt, tot, s);
    hipLaunchKernelGGL(k_sah_link, dim3(grid_for(nt)), dim3(BLK), 0, s, ta, nt, nl, ci, next_id, t.left, t.right,
                       t.rfirst, t.rlast, t.pint, t.pleaf, tb);
    uint32_t h = 0;
    XDG_HIP_CHECK(hipMemcpyAsync(&h, tot, sizeof(h), hipMemcpyDeviceToHost, s));
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    if ((int64_t)h > cap || (int64_t)next_id + h > n - 1) throw HipError{"sah: task count out of range"};
    next_id += (int32_t)h;
    nt = (int64_t)h;
    std::swap(ta, tb);
  }
  if (next_id != n - 1) throw HipError{"sah: internal node count differs from n - 1"};
}


// ---------------------------------------------------------------------------
// Structural check of the emitted binary tree, breadth-first from the root:
// every child index in range, every primitive covered exactly once by a
// reachable leaf, depth (root = 1) within the traversal stack, the leaf
// order a permutation.  err bits: 1 index out of range, 2 too deep.
// ---------------------------------------------------------------------------
struct BItem {
  int32_t node;
  int32_t depth;
};

__global__ void k_b2_level(const NodeD* __restrict__ nd, int64_t n_nodes, int64_t n, const BItem* __restrict__ in,
                           unsigned int nin, BItem* out, unsigned int* nout, unsigned int* seen, int* depth,
                           unsigned int* err)  // out holds n_nodes items
{
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin) return;
  const BItem it = in[i];
  if (it.node < 0 || it.node >= n_nodes) {
    atomicOr(err, 1u);
    return;
  }
  atomicMax(depth, it.depth);
  if (it.depth > 64) {
    atomicOr(err, 2u);
    return;
  }
  const int32_t ch[2] = {nd[it.node].c0, nd[it.node].c1};
  for (int c = 0; c < 2; ++c) {
    const int32_t x = ch[c];
    if (x >= 0) {
      const unsigned int slot = atomicAdd(nout, 1u);
      if ((int64_t)slot < n_nodes)
        out[slot] = BItem{x, it.depth + 1};
      else
        atomicOr(err, 1u);  // more items than internal nodes: a node is reached twice
    } else {
      const int64_t f = leaf_first(x), k = leaf_count(x);
      if (f < 0 || f + k > n) {
        atomicOr(err, 1u);
        continue;
      }
      for (int64_t p = f; p < f + k; ++p) atomicAdd(seen + p, 1u);
    }
  }
}

// every entry of seen must be 1; order must be a permutation of [0, n)
__global__ void k_cover_check(const unsigned int* __restrict__ seen, const int32_t* __restrict__ order,
                              unsigned int* perm, int64_t n, unsigned int* err)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (seen[i] != 1u) atomicOr(err, 4u);
  if (order) {
    const int32_t o = order[i];
    if (o < 0 || o >= n) atomicOr(err, 8u);
    else if (atomicAdd(perm + o, 1u) != 0u) atomicOr(err, 8u);
  }
}

// returns the depth (root = 1)
int validate_tree(const LbvhOut& out, int64_t n, hipStream_t s)
{
  if (out.root < 0 || out.n_nodes == 0) return 0;  // a single leaf
  std::vector<DevBuf> tmp;
  try {
    auto* lv0 = (BItem*)dev_alloc(sizeof(BItem) * (size_t)out.n_nodes, tmp);
    auto* lv1 = (BItem*)dev_alloc(sizeof(BItem) * (size_t)out.n_nodes, tmp);
    auto* seen = (unsigned int*)dev_alloc(sizeof(unsigned int) * (size_t)n, tmp);
    auto* perm = (unsigned int*)dev_alloc(sizeof(unsigned int) * (size_t)n, tmp);
    auto* misc = (unsigned int*)dev_alloc(sizeof(unsigned int) * 4, tmp);  // nout, err, depth
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLK), 0, s, seen, n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLK), 0, s, perm, n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(BLK), 0, s, misc, 4, 0u);
    const BItem root{out.root, 1};
    XDG_HIP_CHECK(hipMemcpyAsync(lv0, &root, sizeof root, hipMemcpyHostToDevice, s));
    unsigned int nin = 1;
    for (int level = 0; nin > 0; ++level) {
      if (level > 64) throw HipError{"lbvh: tree deeper than 64"};
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, 1, 0u);
      hipLaunchKernelGGL(k_b2_level, dim3(grid_for(nin)), dim3(BLK), 0, s, out.nodes_d, out.n_nodes, n, lv0, nin,
                         lv1, misc, seen, (int*)(misc + 2), misc + 1);
      unsigned int h2[2] = {0, 0};
      XDG_HIP_CHECK(hipMemcpyAsync(h2, misc, sizeof h2, hipMemcpyDeviceToHost, s));
      XDG_HIP_CHECK(hipStreamSynchronize(s));
      nin = h2[0];
      if (h2[1] != 0u || nin > (unsigned int)out.n_nodes) throw HipError{"lbvh: malformed tree (bad child index, node reached twice or too deep)"};
      std::swap(lv0, lv1);
    }
    hipLaunchKernelGGL(k_cover_check, dim3(grid_for(n)), dim3(BLK), 0, s, seen, out.order, perm, n, misc + 1);
    unsigned int h[3] = {0, 0, 0};
    XDG_HIP_CHECK(hipMemcpyAsync(h, misc, sizeof h, hipMemcpyDeviceToHost, s));
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    if (h[1] & 1u) throw HipError{"lbvh: child index out of range"};
    if (h[1] & 2u) throw HipError{"lbvh: tree deeper than 64"};
    if (h[1] & 4u) throw HipError{"lbvh: a primitive is not covered exactly once"};
    if (h[1] & 8u) throw HipError{"lbvh: leaf order is not a permutation"};
    dev_free_all(tmp);
    return (int)h[2];
  } catch (...) {
    dev_free_all(tmp);
    throw;
  }
}


}  // namespace

int64_t radix_sort_hist_words(int64_t n) { return (int64_t)RS_BINS * std::max<int64_t>((n + RS_TILE - 1) / RS_TILE, 1); }

void radix_sort_pairs(uint64_t* keys, int32_t* vals, uint64_t* k2, int32_t* v2, int64_t n, int bits, uint32_t* hist,
                      hipStream_t s)
{
  radix_sort_impl(keys, vals, k2, v2, n, bits, hist, s);
}

void radix_sort_pairs(uint32_t* keys, int32_t* vals, uint32_t* k2, int32_t* v2, int64_t n, int bits, uint32_t* hist,
                      hipStream_t s)
{
  radix_sort_impl(keys, vals, k2, v2, n, bits, hist, s);
}

void launch_scan_excl(uint32_t* a, int64_t m, uint32_t* total, hipStream_t s)
{
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, a, m, total);
}

void lbvh_build(const double* box6, int64_t n, int leaf_max, hipStream_t s, std::vector<DevBuf>& owned,
                LbvhOut& out, bool keep_topology, bool sah)
{
  if (leaf_max < 1) leaf_max = 1;
  if (leaf_max > 8) leaf_max = 8;
  out.leaf_max = leaf_max;
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
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (unsigned int*)d_max, 2, 0u);
    hipLaunchKernelGGL(k_maxabs, dim3(std::min<unsigned>(grid_for(6 * n), 4096)), dim3(BLK), 0, s, box6, n, d_max);

    auto* cb = (unsigned long long*)dev_alloc(6 * sizeof(unsigned long long), tmp);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, cb);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(BLK), 0, s, box6, n,
                       cb);
    auto* keys = (uint64_t*)dev_alloc(sizeof(uint64_t) * n, tmp);
    auto* keys2 = (uint64_t*)dev_alloc(sizeof(uint64_t) * n, tmp);
    auto* vals2 = (int32_t*)dev_alloc(sizeof(int32_t) * n, tmp);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(BLK), 0, s, box6, n, cb, keys, out.order);
    auto* hist = (uint32_t*)dev_alloc(sizeof(uint32_t) * (size_t)radix_sort_hist_words(n), tmp);
    radix_sort_pairs(keys, out.order, keys2, vals2, n, 63, hist, s);

    unsigned long long hmax = 0;
    if (n == 1) {
      out.root = leaf_code(0, 1);
      out.depth = 0;
    } else {
      int64_t ni = n - 1;
      std::vector<DevBuf>& keep = keep_topology ? owned : tmp;
      LbvhTopo t;
      t.left = (int32_t*)dev_alloc(sizeof(int32_t) * ni, keep);
      t.right = (int32_t*)dev_alloc(sizeof(int32_t) * ni, keep);
      t.rfirst = (int32_t*)dev_alloc(sizeof(int32_t) * ni, keep);
      t.rlast = (int32_t*)dev_alloc(sizeof(int32_t) * ni, keep);
      t.pint = (int32_t*)dev_alloc(sizeof(int32_t) * ni, keep);
      t.pleaf = (int32_t*)dev_alloc(sizeof(int32_t) * n, keep);
      t.flags = (unsigned int*)dev_alloc(sizeof(unsigned int) * ni, keep);
      t.ibox = (double*)dev_alloc(sizeof(double) * 6 * ni, keep);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ni)), dim3(BLK), 0, s, t.flags, ni, 0u);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ni)), dim3(BLK), 0, s, (unsigned int*)t.pint, ni, 0xffffffffu);
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLK), 0, s, (unsigned int*)t.pleaf, n, 0xffffffffu);
      if (sah)
        sah_topology(box6, keys, n, out.order, t, s, tmp);
      else
        hipLaunchKernelGGL(k_karras, dim3(grid_for(ni)), dim3(BLK), 0, s, keys, n, t.left, t.right, t.rfirst,
                           t.rlast, t.pint, t.pleaf);
      hipLaunchKernelGGL(k_refit, dim3(grid_for(n)), dim3(BLK), 0, s, box6, out.order, n, t.left, t.right, t.pint,
                         t.pleaf, t.flags, t.ibox);
      hipLaunchKernelGGL(k_emit, dim3(grid_for(ni)), dim3(BLK), 0, s, box6, out.order, n, leaf_max, t.left, t.right,
                         t.rfirst, t.rlast, t.ibox, out.nodes_d, out.nodes_f);
      out.root = n <= leaf_max ? leaf_code(0, (int32_t)n) : 0;
      if (keep_topology) out.topo = t;
    }
    XDG_HIP_CHECK(hipMemcpyAsync(&hmax, d_max, sizeof(hmax), hipMemcpyDeviceToHost, s));
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    double m;
    static_assert(sizeof(m) == sizeof(hmax), "size");
    memcpy(&m, &hmax, sizeof m);
    out.max_abs = m;
    XDG_HIP_CHECK(hipGetLastError());
    // depth and structure from a breadth-first pass over the emitted tree
    out.depth = validate_tree(out, n, s);
  } catch (...) {
    dev_free_all(tmp);
    throw;
  }
  dev_free_all(tmp);
}

void lbvh_refit(const double* box6, int64_t n, LbvhOut& out, hipStream_t s)
{
  if (n <= 1 || out.root < 0) return;
  const LbvhTopo& t = out.topo;
  if (!t.left) throw HipError{"lbvh_refit: the tree was built without its topology"};
  const int64_t ni = n - 1;
  std::vector<DevBuf> tmp;
  try {
    auto* d_max = (unsigned long long*)dev_alloc(sizeof(unsigned long long), tmp);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (unsigned int*)d_max, 2, 0u);
    hipLaunchKernelGGL(k_maxabs, dim3(std::min<unsigned>(grid_for(6 * n), 4096)), dim3(BLK), 0, s, box6, n, d_max);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ni)), dim3(BLK), 0, s, t.flags, ni, 0u);
    hipLaunchKernelGGL(k_refit, dim3(grid_for(n)), dim3(BLK), 0, s, box6, out.order, n, t.left, t.right, t.pint,
                       t.pleaf, t.flags, t.ibox);
    hipLaunchKernelGGL(k_emit, dim3(grid_for(ni)), dim3(BLK), 0, s, box6, out.order, n, out.leaf_max, t.left,
                       t.right, t.rfirst, t.rlast, t.ibox, out.nodes_d, out.nodes_f);
    unsigned long long hmax = 0;
    XDG_HIP_CHECK(hipMemcpyAsync(&hmax, d_max, sizeof(hmax), hipMemcpyDeviceToHost, s));
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    memcpy(&out.max_abs, &hmax, sizeof hmax);
    XDG_HIP_CHECK(hipGetLastError());
  } catch (...) {
    dev_free_all(tmp);
    throw;
  }
  dev_free_all(tmp);
}

// ---------------------------------------------------------------------------
// BVH4 collapse on the device, level by level from the root.  Each 4-wide
// node takes a binary node's two children and repeatedly opens the internal
// child with the largest surface area (first such child on ties) until it
// holds four children or only leaves remain.  A level's items count their
// internal children, an exclusive scan numbers them, and the next level is
// written in that order, so node numbering is deterministic and the
// children of a node are contiguous.  Leaf codes are unchanged, so the BVH4
// addresses the same leaf-ordered primitives.  The pass also checks the
// result: children in range, every primitive covered once by a reachable
// leaf, the stack bound (pushes along any root-to-leaf path).
// ---------------------------------------------------------------------------
namespace {

struct C4Item {
  int32_t bnode;  // binary node this BVH4 node is built from
  int32_t slot;   // BVH4 node index
  int32_t above;  // stack entries pushed above it on its path
};

__device__ __forceinline__ double half_area_d(const double* b)
{
  const double dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ const double* src_box(const NodeD* nd, int32_t src)
{
  return (src & 1) ? nd[src >> 1].b1 : nd[src >> 1].b0;
}

// children of the BVH4 node over binary node b: codes and box sources
__device__ __forceinline__ int c4_children(const NodeD* __restrict__ nd, int64_t n_nodes, int32_t b, int32_t* code,
                                           int32_t* src, bool* bad)
{
  code[0] = nd[b].c0;
  src[0] = 2 * b;
  code[1] = nd[b].c1;
  src[1] = 2 * b + 1;
  int nc = 2;
  while (nc < 4) {
    int best = -1;
    double area = -1.0;
    for (int j = 0; j < nc; ++j)
      if (code[j] >= 0) {
        const double a = half_area_d(src_box(nd, src[j]));
        if (a > area) {
          area = a;
          best = j;
        }
      }
    if (best < 0) break;
    const int32_t x = code[best];
    if (x >= n_nodes) {
      *bad = true;
      return 0;
    }
    code[best] = nd[x].c0;
    src[best] = 2 * x;
    code[nc] = nd[x].c1;
    src[nc] = 2 * x + 1;
    ++nc;
  }
  return nc;
}

__global__ void k_c4_count(const NodeD* __restrict__ nd, int64_t n_nodes, const C4Item* __restrict__ items,
                           unsigned int nitems, uint32_t* cnt, unsigned int* err)
{
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  int32_t code[4], src[4];
  bool bad = items[i].bnode < 0 || items[i].bnode >= n_nodes;
  int nc = bad ? 0 : c4_children(nd, n_nodes, items[i].bnode, code, src, &bad);
  uint32_t k = 0;
  for (int j = 0; j < nc; ++j) k += code[j] >= 0 ? 1u : 0u;
  if (bad) {
    atomicOr(err, 1u);
    k = 0;
  }
  cnt[i] = k;
}

__global__ void k_c4_emit(const NodeD* __restrict__ nd, const NodeF* __restrict__ nf, int64_t n_nodes, int64_t n,
                          const C4Item* __restrict__ items, unsigned int nitems, const uint32_t* __restrict__ offs,
                          int32_t next_base, int64_t cap, Node4D* o4d, Node4F* o4f, int32_t* src4, C4Item* next,
                          unsigned int* seen, int* max_stack, unsigned int* err)
{
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const C4Item it = items[i];
  if (it.slot < 0 || it.slot >= cap) {
    atomicOr(err, 1u);
    return;
  }
  int32_t code[4], src[4];
  bool bad = it.bnode < 0 || it.bnode >= n_nodes;
  const int nc = bad ? 0 : c4_children(nd, n_nodes, it.bnode, code, src, &bad);
  if (bad) {
    atomicOr(err, 1u);
    return;
  }
  const int32_t above = it.above + (nc - 1);
  atomicMax(max_stack, above);
  Node4D d4;
  Node4F f4;
  memset((void*)&d4, 0, sizeof d4);
  memset((void*)&f4, 0, sizeof f4);
  uint32_t rank = 0;
  for (int j = 0; j < 4; ++j) {
    if (j < nc) {
      const double* b = src_box(nd, src[j]);
      const float* bf = (src[j] & 1) ? nf[src[j] >> 1].b1 : nf[src[j] >> 1].b0;
      d4.lox[j] = b[0];
      d4.loy[j] = b[1];
      d4.loz[j] = b[2];
      d4.hix[j] = b[3];
      d4.hiy[j] = b[4];
      d4.hiz[j] = b[5];
      f4.lox[j] = bf[0];
      f4.loy[j] = bf[1];
      f4.loz[j] = bf[2];
      f4.hix[j] = bf[3];
      f4.hiy[j] = bf[4];
      f4.hiz[j] = bf[5];
      if (code[j] >= 0) {
        const int32_t slot = next_base + (int32_t)(offs[i] + rank);
        next[offs[i] + rank] = C4Item{code[j], slot, above};
        ++rank;
        d4.child[j] = f4.child[j] = slot;
      } else {
        const int64_t f = leaf_first(code[j]), k = leaf_count(code[j]);
        if (f < 0 || f + k > n) {
          atomicOr(err, 1u);
        } else {
          for (int64_t p = f; p < f + k; ++p) atomicAdd(seen + p, 1u);
        }
        d4.child[j] = f4.child[j] = code[j];
      }
      if (src4) src4[4 * (int64_t)it.slot + j] = src[j];
    } else {
      // inverted box: the near-plane entry distance is +inf for every ray
      d4.lox[j] = d4.loy[j] = d4.loz[j] = HUGE_VAL;
      d4.hix[j] = d4.hiy[j] = d4.hiz[j] = -HUGE_VAL;
      f4.lox[j] = f4.loy[j] = f4.loz[j] = HUGE_VALF;
      f4.hix[j] = f4.hiy[j] = f4.hiz[j] = -HUGE_VALF;
      d4.child[j] = f4.child[j] = INT32_MIN;
      if (src4) src4[4 * (int64_t)it.slot + j] = -1;
    }
  }
  o4d[it.slot] = d4;
  o4f[it.slot] = f4;
}

__global__ void k_c4_refit(const NodeD* __restrict__ nd, const NodeF* __restrict__ nf, const int32_t* __restrict__ src4,
                           int64_t n4, Node4D* o4d, Node4F* o4f)
{
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= 4 * n4) return;
  const int64_t node = q >> 2;
  const int j = (int)(q & 3);
  const int32_t src = src4[q];
  if (src < 0) return;  // empty slot keeps its inverted box
  const double* b = src_box(nd, src);
  const float* bf = (src & 1) ? nf[src >> 1].b1 : nf[src >> 1].b0;
  Node4D& d4 = o4d[node];
  Node4F& f4 = o4f[node];
  d4.lox[j] = b[0];
  d4.loy[j] = b[1];
  d4.loz[j] = b[2];
  d4.hix[j] = b[3];
  d4.hiy[j] = b[4];
  d4.hiz[j] = b[5];
  f4.lox[j] = bf[0];
  f4.loy[j] = bf[1];
  f4.loz[j] = bf[2];
  f4.hix[j] = bf[3];
  f4.hiy[j] = bf[4];
  f4.hiz[j] = bf[5];
}

__global__ void k_copy_bytes16(const uint4* __restrict__ src, uint4* dst, int64_t n16)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

}  // namespace

void bvh4_collapse(const LbvhOut& bin, int64_t n, hipStream_t s, std::vector<DevBuf>& owned, Bvh4Out& out,
                   bool keep_sources)
{
  if (bin.root < 0 || bin.n_nodes == 0) {  // a single leaf: no nodes
    out.root = bin.root;
    out.stack = 0;
    out.n_nodes = 0;
    out.nodes_d = (Node4D*)dev_alloc(sizeof(Node4D), owned);
    out.nodes_f = (Node4F*)dev_alloc(sizeof(Node4F), owned);
    return;
  }
  std::vector<DevBuf> tmp;
  try {
    // a 4-wide node absorbs at least one binary node: n_nodes bounds the count
    const int64_t cap = bin.n_nodes;
    auto* t4d = (Node4D*)dev_alloc(sizeof(Node4D) * (size_t)cap, tmp);
    auto* t4f = (Node4F*)dev_alloc(sizeof(Node4F) * (size_t)cap, tmp);
    int32_t* tsrc = keep_sources ? (int32_t*)dev_alloc(sizeof(int32_t) * 4 * (size_t)cap, tmp) : nullptr;
    auto* lv0 = (C4Item*)dev_alloc(sizeof(C4Item) * (size_t)cap, tmp);
    auto* lv1 = (C4Item*)dev_alloc(sizeof(C4Item) * (size_t)cap, tmp);
    auto* cnt = (uint32_t*)dev_alloc(sizeof(uint32_t) * (size_t)cap, tmp);
    auto* seen = (unsigned int*)dev_alloc(sizeof(unsigned int) * (size_t)n, tmp);
    auto* misc = (unsigned int*)dev_alloc(sizeof(unsigned int) * 4, tmp);  // total, err, max_stack
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLK), 0, s, seen, n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, misc, 4, 0u);
    const C4Item root{bin.root, 0, 0};
    XDG_HIP_CHECK(hipMemcpyAsync(lv0, &root, sizeof root, hipMemcpyHostToDevice, s));
    unsigned int nin = 1;
    int64_t slots = 1;
    for (int level = 0; nin > 0; ++level) {
      if (level > 64) throw HipError{"bvh4: deeper than 64 levels"};
      hipLaunchKernelGGL(k_c4_count, dim3(grid_for(nin)), dim3(BLK), 0, s, bin.nodes_d, bin.n_nodes, lv0, nin, cnt,
                         misc + 1);
      hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, cnt, (int64_t)nin, misc);
      unsigned int total = 0;
      XDG_HIP_CHECK(hipMemcpyAsync(&total, misc, sizeof total, hipMemcpyDeviceToHost, s));
      XDG_HIP_CHECK(hipStreamSynchronize(s));
      if (slots + (int64_t)total > cap) throw HipError{"bvh4: more nodes than the binary tree has"};
      hipLaunchKernelGGL(k_c4_emit, dim3(grid_for(nin)), dim3(BLK), 0, s, bin.nodes_d, bin.nodes_f, bin.n_nodes, n,
                         lv0, nin, cnt, (int32_t)slots, cap, t4d, t4f, tsrc, lv1, seen, (int*)(misc + 2), misc + 1);
      slots += total;
      nin = total;
      std::swap(lv0, lv1);
    }
    hipLaunchKernelGGL(k_cover_check, dim3(grid_for(n)), dim3(BLK), 0, s, seen, (const int32_t*)nullptr,
                       (unsigned int*)nullptr, n, misc + 1);
    unsigned int h[3] = {0, 0, 0};
    XDG_HIP_CHECK(hipMemcpyAsync(h, misc, sizeof h, hipMemcpyDeviceToHost, s));
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    if (h[1] & 1u) throw HipError{"bvh4: child index or leaf range out of bounds"};
    if (h[1] & 4u) throw HipError{"bvh4: a primitive is not covered exactly once"};
    // right-sized final arrays, copied on the device
    out.n_nodes = slots;
    out.nodes_d = (Node4D*)dev_alloc(sizeof(Node4D) * (size_t)slots, owned);
    out.nodes_f = (Node4F*)dev_alloc(sizeof(Node4F) * (size_t)slots, owned);
    XDG_HIP_CHECK(hipMemcpyAsync(out.nodes_d, t4d, sizeof(Node4D) * (size_t)slots, hipMemcpyDeviceToDevice, s));
    XDG_HIP_CHECK(hipMemcpyAsync(out.nodes_f, t4f, sizeof(Node4F) * (size_t)slots, hipMemcpyDeviceToDevice, s));
    if (keep_sources) {
      out.src = (int32_t*)dev_alloc(sizeof(int32_t) * 4 * (size_t)slots, owned);
      XDG_HIP_CHECK(hipMemcpyAsync(out.src, tsrc, sizeof(int32_t) * 4 * (size_t)slots, hipMemcpyDeviceToDevice, s));
    }
    XDG_HIP_CHECK(hipStreamSynchronize(s));
    // the stack never holds more than the pushes along one root-to-leaf path
    out.stack = std::max((int32_t)h[2], 1);
    out.root = 0;
  } catch (...) {
    dev_free_all(tmp);
    throw;
  }
  dev_free_all(tmp);
}

void bvh4_refit(const LbvhOut& bin, const Bvh4Out& b4, hipStream_t s)
{
  if (b4.n_nodes == 0) return;
  if (!b4.src) throw HipError{"bvh4_refit: the tree was built without its box sources"};
  hipLaunchKernelGGL(k_c4_refit, dim3(grid_for(4 * b4.n_nodes)), dim3(BLK), 0, s, bin.nodes_d, bin.nodes_f, b4.src,
                     b4.n_nodes, b4.nodes_d, b4.nodes_f);
  XDG_HIP_CHECK(hipGetLastError());
}

}  // namespace xdgh
