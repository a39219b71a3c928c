// capi.hip -- extern "C" boundary of libxdg_hip.so (declared in
// include/xdg_hip.h): context, tree registry, argument validation, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "xdg_hip.h"

using namespace xdgh;
using namespace xdgd;

namespace {

struct SurfTreeHost {
  SurfTreeDev dev{};
  std::vector<DevBuf> bufs;
  bool released = false;
  int32_t min_face = 0, max_face = -1;
  int64_t n_nodes = 0;
  int64_t n_nodes4 = 0;
  int64_t bytes = 0;
  // kept for xdg_hip_refit_surface_tree (trees built with the "refit" option)
  bool refittable = false;
  LbvhOut lb;
  Bvh4Out b4;
  int32_t* tri_idx = nullptr;  // 3 vertex indices per triangle, caller's order
  int64_t n_verts = 0;
};

// One registered transport problem: device copies of the topology, the
// tuned launch configuration and the buffers a launch needs (so a launch
// allocates nothing and can be captured in a graph).
struct TransportHost {
  TransportDev dev{};
  std::vector<DevBuf> bufs;
  int block = 128, lds_stack = 0, chunk = 1, refill = 1;
  int stack_need = 1;
  int32_t* spill = nullptr;
  size_t spill_entries = 0;
  unsigned long long* d_next = nullptr;
  uint32_t* d_live = nullptr;
  void* stage = nullptr;  // start state of one slice (staged form)
  int64_t stage_cap = 0;
  int stage_items = 0;
  int64_t start_grid[2] = {0, 0};  // plain, mixed: resident grid of the start stage's PIV kernel
  std::vector<int32_t> host_trees;
  uint64_t generation = 1;  // bumped when the spill area or the start buffer moves
};

// An instantiated transport graph and what it was captured against: a
// replay after the context's or the transport's buffers were freed or moved
// is refused instead of reading freed memory.
struct GraphHandle {
  hipGraphExec_t exec = nullptr;
  const void* ctx = nullptr;
  uint64_t ctx_generation = 0;
  int32_t transport = 0;
  uint64_t transport_generation = 0;
};

struct ElemTreeHost {
  ElemTreeDev dev{};
  std::vector<DevBuf> bufs;
  int64_t n_nodes = 0;
  int64_t bytes = 0;
  int64_t n_bfaces = 0;
};

struct InvalidArg {
  std::string msg;
};
struct NotFound {
  std::string msg;
};

int env_int(const char* name, int dflt)
{
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

int64_t bufs_bytes(const std::vector<DevBuf>& b)
{
  int64_t s = 0;
  for (auto& x : b) s += (int64_t)x.bytes;
  return s;
}

}  // namespace

struct xdg_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  std::vector<SurfTreeHost> surf;  // tree id = index + 1
  std::vector<ElemTreeHost> elem;
  SurfTreeDev* d_surf = nullptr;
  size_t d_surf_cap = 0;
  ElemTreeDev* d_elem = nullptr;
  size_t d_elem_cap = 0;
  int32_t* d_small = nullptr;  // find_volume tree/id lists
  size_t small_cap = 0;
  std::vector<int32_t> small_cache;
  uint32_t* d_status = nullptr;
  int32_t* d_fb_list = nullptr;  // mixed ray_fire: rays deferred to the FP64 pass
  size_t fb_cap = 0;
  unsigned int* d_fb_count = nullptr;
  unsigned long long* d_next = nullptr;  // persistent kernels' work counters: [0] ray_fire, [1] find_volume
  char* d_mw = nullptr;  // wavefront mesh walk scratch
  int64_t mw_cap = 0;
  MeshWalkScratch mw_w{};
  void* d_order = nullptr;  // ray ordering: keys, indices and sort storage
  int64_t order_cap = 0;
  RayOrderScratch order_w{};
  std::vector<TransportHost*> transports;  // handle = index + 1 (nullptr once destroyed)
  // bumped whenever device memory a captured graph may point at is freed or
  // moved (tree tables regrown, trees released); graphs remember the value
  uint64_t generation = 1;
  // the shared scratch above (d_next, d_fb_list/d_fb_count, d_order, d_mw)
  // serves one launch at a time: the last stream that used it and an event
  // after that use, so that a call on another stream waits for it
  hipStream_t scratch_stream = nullptr;
  hipEvent_t scratch_ev = nullptr;
  LaunchCfg cfg;
  int leaf_max = 2;
  int build_sah = 1;  // trees built next: binned-SAH hierarchy (1) or the Karras radix tree (0)
  int refit = 0;  // surface trees built next keep what a refit needs
  int max_surf_depth = 1;
  int max_surf_stack4 = 1;
  int max_elem_depth = 1;
};

namespace {

void sync_tables(xdg_hip_ctx* c)
{
  XDG_HIP_CHECK(hipDeviceSynchronize());
  ++c->generation;
  size_t ns = c->surf.size() + 1;
  if (ns > c->d_surf_cap) {
    if (c->d_surf) XDG_HIP_CHECK(hipFree(c->d_surf));
    c->d_surf_cap = std::max<size_t>(ns * 2, 16);
    XDG_HIP_CHECK(hipMalloc(&c->d_surf, c->d_surf_cap * sizeof(SurfTreeDev)));
  }
  std::vector<SurfTreeDev> hs(ns);
  std::memset((void*)hs.data(), 0, ns * sizeof(SurfTreeDev));
  for (size_t i = 0; i < c->surf.size(); ++i) hs[i + 1] = c->surf[i].dev;
  XDG_HIP_CHECK(hipMemcpy(c->d_surf, hs.data(), ns * sizeof(SurfTreeDev), hipMemcpyHostToDevice));

  size_t ne = c->elem.size() + 1;
  if (ne > c->d_elem_cap) {
    if (c->d_elem) XDG_HIP_CHECK(hipFree(c->d_elem));
    c->d_elem_cap = std::max<size_t>(ne * 2, 16);
    XDG_HIP_CHECK(hipMalloc(&c->d_elem, c->d_elem_cap * sizeof(ElemTreeDev)));
  }
  std::vector<ElemTreeDev> he(ne);
  std::memset((void*)he.data(), 0, ne * sizeof(ElemTreeDev));
  for (size_t i = 0; i < c->elem.size(); ++i) he[i + 1] = c->elem[i].dev;
  XDG_HIP_CHECK(hipMemcpy(c->d_elem, he.data(), ne * sizeof(ElemTreeDev), hipMemcpyHostToDevice));
}

void require(bool ok, const char* msg)
{
  if (!ok) throw InvalidArg{msg};
}

const SurfTreeHost& surf_tree(xdg_hip_ctx* c, int32_t tree)
{
  if (tree < 1 || (size_t)tree > c->surf.size() || c->surf[tree - 1].released)
    throw NotFound{"surface tree " + std::to_string(tree) + " is not registered"};
  return c->surf[tree - 1];
}

const ElemTreeHost& elem_tree(xdg_hip_ctx* c, int32_t tree)
{
  if (tree < 1 || (size_t)tree > c->elem.size())
    throw NotFound{"element tree " + std::to_string(tree) + " is not registered"};
  return c->elem[tree - 1];
}

hipStream_t pick(xdg_hip_ctx* c, void* s) { return s ? (hipStream_t)s : c->stream; }

// FP32 box test error for origin o: at most (3u max|x| + 5u |o|) / |d_i| in
// ray-parameter units (u = 2^-24) against a margin of bump / |d_i|, so
// bump * 2^20 - max|x| keeps it below 5/16 of the margin.  The 2^26 cap
// keeps fl32(o) times the clamped reciprocal 1e30 of a near-axis direction
// finite.  Rounded down to a float.
// FP32 bounds rounded outward: the box holds every vertex exactly (the
// transport's volume prefilter relies on that, whatever the coordinates'
// magnitude)
float f32_lo(double x)
{
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
float f32_hi(double x)
{
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

float mixed_limit(double bump, double max_abs)
{
  const double lim = std::min(bump * 1048576.0 - max_abs, 67108864.0);
  float f = lim > 0.0 ? (float)lim : 0.0f;
  if ((double)f > lim) f = nextafterf(f, 0.0f);
  return f;
}

bool capturing(hipStream_t s)
{
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  XDG_HIP_CHECK(hipStreamIsCapturing(s, &cs));
  return cs != hipStreamCaptureStatusNone;
}

// Before a launch that uses the context's shared scratch on stream s: wait
// for the last such launch if it went to another stream.  (Not under graph
// capture: a captured launch is ordered by the caller.)
void scratch_acquire(xdg_hip_ctx* c, hipStream_t s)
{
  if (c->scratch_stream && c->scratch_stream != s && !capturing(s))
    XDG_HIP_CHECK(hipStreamWaitEvent(s, c->scratch_ev, 0));
}

// After that launch: later users on other streams wait for this point.
void scratch_release(xdg_hip_ctx* c, hipStream_t s)
{
  if (capturing(s)) return;
  XDG_HIP_CHECK(hipEventRecord(c->scratch_ev, s));
  c->scratch_stream = s;
}

__global__ void k_tally_guard(const long long* __restrict__ tally, int64_t n, uint32_t* status, uint32_t bit)
{
  const long long lim = 1LL << 62;
  bool near = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const long long v = tally[i];
    near |= v >= lim || v < -lim;
  }
  if (near) atomicOr(status, bit);
}

// ---- refit of a surface tree (xdg_hip_refit_surface_tree) ----------------
__global__ void k_gather_v9(const double* __restrict__ verts, const int32_t* __restrict__ tri_idx, int64_t n,
                            double* v9)
{
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int j = 0; j < 3; ++j) {
    const int64_t v = tri_idx[3 * p + j];
    for (int k = 0; k < 3; ++k) v9[9 * p + 3 * j + k] = verts[3 * v + k];
  }
}

// new vertices into the leaf-ordered records; face, surface and flip kept,
// the edge endpoint orders recomputed from the new coordinates
__global__ void k_refit_tris(const double* __restrict__ v9, const int32_t* __restrict__ order, int64_t n, TriRec* tris)
{
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t p = order[s];
  TriRec r = tris[s];
  for (int k = 0; k < 9; ++k) r.v[k] = v9[9 * p + k];
  const unsigned fb = tri_first_bits(d3(r.v[0], r.v[1], r.v[2]), d3(r.v[3], r.v[4], r.v[5]), d3(r.v[6], r.v[7], r.v[8]));
  r.surf2 = (r.surf2 & ~0xE) | (int32_t)(fb << 1);
  tris[s] = r;
}

__device__ __forceinline__ unsigned long long ord_key(double x)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double ord_val(unsigned long long u)
{
  const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffULL) : ~u;
  double d;
  memcpy(&d, &b, sizeof d);
  return d;
}

// min / max vertex coordinate per axis (ordered-integer atomics); cb[0..2]
// start at ~0, cb[3..5] at 0
__global__ void k_v9_bounds(const double* __restrict__ v9, int64_t n, unsigned long long* cb)
{
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int k = 0; k < 3; ++k) {
    const double a = v9[9 * p + k], b = v9[9 * p + 3 + k], c = v9[9 * p + 6 + k];
    atomicMin(cb + k, ord_key(fmin(fmin(a, b), c)));
    atomicMax(cb + 3 + k, ord_key(fmax(fmax(a, b), c)));
  }
}

__global__ void k_init_minmax(unsigned long long* cb)
{
  const int i = threadIdx.x;
  if (i < 3) cb[i] = ~0ULL;
  else if (i < 6) cb[i] = 0ULL;
}

// Sets `bit` of the status word when an int64 fixed-point tally has reached
// 2^62 in magnitude (half its range; the sums are of non-negative lengths,
// so a wrapped sum reads negative and is caught as well).
void tally_guard(xdg_hip_ctx* c, const void* tally, int64_t n, uint32_t bit, hipStream_t s)
{
  if (!tally || n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
  k_tally_guard<<<dim3((unsigned)blocks), dim3(256), 0, s>>>((const long long*)tally, n, c->d_status, bit);
}

// traversal stack entries a surface query needs: one tree, or every tree
// (per-ray trees)
int surf_need(xdg_hip_ctx* c, const SurfTreeHost* t)
{
  if (c->cfg.wide) return t ? t->dev.stack4 : c->max_surf_stack4;
  return t ? t->dev.depth : c->max_surf_depth;
}

LaunchCfg cfg_for(xdg_hip_ctx* c, int depth)
{
  LaunchCfg k = c->cfg;
  k.stack = std::max(k.stack, std::max(depth, 1));
  if (k.stack > 64) throw InvalidArg{"tree deeper than the 64-entry traversal stack"};
  // keep the LDS stack within 64 KiB per workgroup
  while ((size_t)k.stack * k.block * 4 > 65536 && k.block > 64) k.block /= 2;
  return k;
}

// ray ordering buffers for n rays (grown on demand, kept by the context)
const RayOrderScratch& order_scratch(xdg_hip_ctx* c, int64_t n)
{
  if (n > c->order_cap) {
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (c->d_order) XDG_HIP_CHECK(hipFree(c->d_order));
    c->d_order = nullptr;
    c->order_cap = 0;
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    const size_t tb = ray_order_temp_bytes(n);
    char* p = nullptr;
    XDG_HIP_CHECK(hipMalloc((void**)&p, 4 * a + tb));
    c->d_order = p;
    c->order_cap = n;
    c->order_w = RayOrderScratch{(uint32_t*)p, (uint32_t*)(p + a), (int32_t*)(p + 2 * a), (int32_t*)(p + 3 * a),
                                 (void*)(p + 4 * a), tb};
  }
  return c->order_w;
}

// the mixed kernels' deferral list, grown to n entries on demand
int32_t* fb_list_for(xdg_hip_ctx* c, int64_t n)
{
  if ((size_t)n > c->fb_cap) {
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (c->d_fb_list) XDG_HIP_CHECK(hipFree(c->d_fb_list));
    c->d_fb_list = nullptr;
    c->fb_cap = 0;
    XDG_HIP_CHECK(hipMalloc(&c->d_fb_list, sizeof(int32_t) * (size_t)n));
    c->fb_cap = (size_t)n;
  }
  return c->d_fb_list;
}

// scratch of the wavefront mesh walk for n particles (grown on demand)
const MeshWalkScratch& mw_scratch_for(xdg_hip_ctx* c, int64_t n)
{
  if (n > c->mw_cap) {
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (c->d_mw) XDG_HIP_CHECK(hipFree(c->d_mw));
    c->d_mw = nullptr;
    c->mw_cap = 0;
    const size_t sv = ((size_t)n * sizeof(MwSave) + 255) & ~(size_t)255;
    const size_t sl = ((size_t)n * sizeof(int32_t) + 255) & ~(size_t)255;
    XDG_HIP_CHECK(hipMalloc((void**)&c->d_mw, sv + 2 * sl + 256));
    c->mw_cap = n;
    c->mw_w.save = (MwSave*)c->d_mw;
    c->mw_w.list[0] = (int32_t*)(c->d_mw + sv);
    c->mw_w.list[1] = (int32_t*)(c->d_mw + sv + sl);
    c->mw_w.count = (unsigned int*)(c->d_mw + sv + 2 * sl);
    c->mw_w.claim = (unsigned long long*)(c->mw_w.count + 2 * MW_ROUNDS + 2);  // 8-byte aligned, within the 256 bytes
  }
  return c->mw_w;
}

// rays below this count are traced in the caller's order
constexpr int64_t ORDER_MIN_RAYS = 1 << 16;

// ---- scratch of the allocation-free ray_fire --------------------------------
// One array per GPU, in the code object, held until the process ends: the
// persistent kernels' work counter and the mixed kernel's deferral list of
// up to XDG_HIP_NOALLOC_MAX_RAYS rays.  Every allocation-free call on a GPU
// shares it, so two such calls that overlap on one GPU race.
struct NoAllocScratch {
  unsigned long long next[4];  // [0]: next unclaimed ray of the persistent kernels
  unsigned int fb_count[4];    // [0]: rays deferred to the FP64 pass
  int32_t fb_list[XDG_HIP_NOALLOC_MAX_RAYS];
};

}  // namespace

static __device__ __align__(256) NoAllocScratch scratch;

namespace {

// Device address of `scratch` on each GPU (host code never uses the array's
// name as an address: it is a host-side stand-in).  Filled once per device
// by xdg_hip_create with that device current.
std::mutex g_scratch_mu;
std::vector<NoAllocScratch*> g_scratch_ptr;

hipError_t scratch_setup(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = (NoAllocScratch*)p;
  return hipSuccess;
}

NoAllocScratch* scratch_for(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

template <class F>
int guarded(xdg_hip_ctx* c, F&& f)
{
  if (!c) return XDG_HIP_ERR_INVALID;
  std::lock_guard<std::mutex> lk(c->mu);
  try {
    XDG_HIP_CHECK(hipSetDevice(c->device));
    return f();
  } catch (const InvalidArg& e) {
    c->err = e.msg;
    return XDG_HIP_ERR_INVALID;
  } catch (const NotFound& e) {
    c->err = e.msg;
    return XDG_HIP_ERR_NOT_FOUND;
  } catch (const HipError& e) {
    c->err = e.msg;
    return XDG_HIP_ERR_HIP;
  } catch (const std::exception& e) {
    c->err = e.what();
    return XDG_HIP_ERR_HIP;
  } catch (...) {
    c->err = "unknown error";
    return XDG_HIP_ERR_HIP;
  }
}

}  // namespace

extern "C" {

int xdg_hip_create(xdg_hip_ctx** out, int device)
{
  if (!out) return XDG_HIP_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return XDG_HIP_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return XDG_HIP_ERR_NO_DEVICE;
  auto* c = new xdg_hip_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_status, sizeof(uint32_t) * 4) != hipSuccess ||
      hipMalloc(&c->d_fb_count, sizeof(unsigned int) * 4) != hipSuccess ||
      hipMalloc(&c->d_next, sizeof(unsigned long long) * 2) != hipSuccess ||
      hipMemset(c->d_status, 0, sizeof(uint32_t) * 4) != hipSuccess) {  // status flags start clear
    if (c->d_next) (void)hipFree(c->d_next);
    if (c->d_fb_count) (void)hipFree(c->d_fb_count);
    if (c->d_status) (void)hipFree(c->d_status);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return XDG_HIP_ERR_HIP;
  }
  if (hipEventCreateWithFlags(&c->scratch_ev, hipEventDisableTiming) != hipSuccess ||
      scratch_setup(device) != hipSuccess) {
    if (c->scratch_ev) (void)hipEventDestroy(c->scratch_ev);
    (void)hipFree(c->d_next);
    (void)hipFree(c->d_fb_count);
    (void)hipFree(c->d_status);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return XDG_HIP_ERR_HIP;
  }
  // environment overrides of the defaults in LaunchCfg
  c->cfg.block = env_int("XDG_HIP_BLOCK", c->cfg.block);
  c->cfg.stack = env_int("XDG_HIP_STACK", 1);
  c->leaf_max = env_int("XDG_HIP_LEAF", c->leaf_max);
  c->build_sah = env_int("XDG_HIP_BUILD", c->build_sah);
  c->cfg.defer = env_int("XDG_HIP_DEFER", c->cfg.defer) != 0;
  c->cfg.wide = env_int("XDG_HIP_WIDE", c->cfg.wide) != 0;
  c->cfg.spec = env_int("XDG_HIP_SPEC", c->cfg.spec) != 0;
  c->cfg.persistent = env_int("XDG_HIP_PERSIST", c->cfg.persistent) & 7;
  c->cfg.order = env_int("XDG_HIP_ORDER", c->cfg.order);
  c->cfg.refill = env_int("XDG_HIP_REFILL", c->cfg.refill);
  try {
    sync_tables(c);
  } catch (...) {
    xdg_hip_destroy(c);
    return XDG_HIP_ERR_HIP;
  }
  *out = c;
  return XDG_HIP_OK;
}

int xdg_hip_destroy(xdg_hip_ctx* c)
{
  if (!c) return XDG_HIP_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto& t : c->surf) dev_free_all(t.bufs);
    for (auto& t : c->elem) dev_free_all(t.bufs);
    for (auto* t : c->transports) {
      if (!t) continue;
      if (t->spill) (void)hipFree(t->spill);
      if (t->stage) (void)hipFree(t->stage);
      dev_free_all(t->bufs);
      delete t;
    }
    if (c->d_surf) (void)hipFree(c->d_surf);
    if (c->d_elem) (void)hipFree(c->d_elem);
    if (c->d_small) (void)hipFree(c->d_small);
    if (c->d_status) (void)hipFree(c->d_status);
    if (c->d_fb_list) (void)hipFree(c->d_fb_list);
    if (c->d_mw) (void)hipFree(c->d_mw);
    if (c->d_fb_count) (void)hipFree(c->d_fb_count);
    if (c->d_next) (void)hipFree(c->d_next);
    if (c->d_order) (void)hipFree(c->d_order);
    if (c->scratch_ev) (void)hipEventDestroy(c->scratch_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
  }
  delete c;
  return XDG_HIP_OK;
}

const char* xdg_hip_last_error(const xdg_hip_ctx* c) { return c ? c->err.c_str() : "null context"; }

int xdg_hip_set_launch(xdg_hip_ctx* c, int block, int stack)
{
  return guarded(c, [&]() {
    require(block >= 64 && block <= 1024 && block % 64 == 0, "block must be a multiple of 64 in [64, 1024]");
    require(stack >= 1 && stack <= 64, "stack must be in [1, 64]");
    c->cfg.block = block;
    c->cfg.stack = stack;
    return XDG_HIP_OK;
  });
}

int xdg_hip_set_traversal(xdg_hip_ctx* c, int wide, int defer, int persistent)
{
  return guarded(c, [&]() {
    c->cfg.wide = wide != 0;
    c->cfg.defer = defer != 0;
    c->cfg.persistent = persistent & 7;
    return XDG_HIP_OK;
  });
}

}  // extern "C"

namespace {

struct OptionRef {
  const char* name;
  int lo, hi;
  int* (*get)(xdg_hip_ctx*);
};

int* opt_block(xdg_hip_ctx* c) { return &c->cfg.block; }
int* opt_stack(xdg_hip_ctx* c) { return &c->cfg.stack; }
int* opt_persistent(xdg_hip_ctx* c) { return &c->cfg.persistent; }
int* opt_order(xdg_hip_ctx* c) { return &c->cfg.order; }
int* opt_spec(xdg_hip_ctx* c) { return &c->cfg.spec; }
int* opt_wide(xdg_hip_ctx* c) { return &c->cfg.wide; }
int* opt_defer(xdg_hip_ctx* c) { return &c->cfg.defer; }
int* opt_leaf(xdg_hip_ctx* c) { return &c->leaf_max; }
int* opt_build(xdg_hip_ctx* c) { return &c->build_sah; }
int* opt_refill(xdg_hip_ctx* c) { return &c->cfg.refill; }
int* opt_fv_refill(xdg_hip_ctx* c) { return &c->cfg.fv_refill; }
int* opt_adj_sort(xdg_hip_ctx* c) { return &c->cfg.adj_sort; }
int* opt_mw_wave(xdg_hip_ctx* c) { return &c->cfg.mw_wave; }
int* opt_mw_refill(xdg_hip_ctx* c) { return &c->cfg.mw_refill; }
int* opt_refit(xdg_hip_ctx* c) { return &c->refit; }
int* opt_seg_order(xdg_hip_ctx* c) { return &c->cfg.seg_order; }
int* opt_tr_start(xdg_hip_ctx* c) { return &c->cfg.tr_start; }
int* opt_tr_slice(xdg_hip_ctx* c) { return &c->cfg.tr_slice; }
int* opt_tr_items(xdg_hip_ctx* c) { return &c->cfg.tr_items; }
int* opt_tr_rounds(xdg_hip_ctx* c) { return &c->cfg.tr_rounds; }
int* opt_tr_events(xdg_hip_ctx* c) { return &c->cfg.tr_events; }

const OptionRef OPTIONS[] = {
    {"block", 64, 1024, opt_block},         {"stack", 1, 64, opt_stack},     {"persistent", 0, 7, opt_persistent},
    {"ray_order", 0, 2, opt_order},         {"speculative", 0, 1, opt_spec}, {"wide", 0, 1, opt_wide},
    {"defer", 0, 1, opt_defer},             {"leaf_size", 1, 8, opt_leaf},
    {"build", 0, 1, opt_build},
    {"refill", 1, 64, opt_refill},          {"mw_wave", 0, 1, opt_mw_wave},
    {"mw_refill", 0, 64, opt_mw_refill},
    {"refit", 0, 1, opt_refit},            {"segment_order", 0, 2, opt_seg_order},
    {"transport_start", 0, 2, opt_tr_start}, {"transport_slice", 6, 24, opt_tr_slice},
    {"transport_items", 1, 64, opt_tr_items}, {"transport_rounds", 0, 64, opt_tr_rounds},
    {"transport_events", 0, 1, opt_tr_events}, {"fv_refill", 0, 64, opt_fv_refill},
    {"adjoint_sort", 0, 1, opt_adj_sort},
};

const OptionRef& find_option(const char* name)
{
  require(name != nullptr, "null option name");
  for (const auto& o : OPTIONS)
    if (std::strcmp(o.name, name) == 0) return o;
  throw NotFound{std::string("unknown option '") + name + "'"};
}

}  // namespace

extern "C" {

int xdg_hip_set_option(xdg_hip_ctx* c, const char* name, int value)
{
  return guarded(c, [&]() {
    const OptionRef& o = find_option(name);
    if (value < o.lo || value > o.hi)
      throw InvalidArg{std::string("option '") + name + "' must be in [" + std::to_string(o.lo) + ", " +
                       std::to_string(o.hi) + "]"};
    if (std::strcmp(name, "block") == 0) require((value & 63) == 0, "block must be a multiple of 64");
    *o.get(c) = value;
    return XDG_HIP_OK;
  });
}

int xdg_hip_get_option(xdg_hip_ctx* c, const char* name, int* value)
{
  return guarded(c, [&]() {
    require(value != nullptr, "null output");
    *value = *find_option(name).get(c);
    return XDG_HIP_OK;
  });
}

int xdg_hip_get_launch(xdg_hip_ctx* c, int* block, int* stack)
{
  return guarded(c, [&]() {
    if (block) *block = c->cfg.block;
    if (stack) *stack = c->cfg.stack;
    return XDG_HIP_OK;
  });
}

int xdg_hip_synchronize(xdg_hip_ctx* c, void* stream)
{
  return guarded(c, [&]() {
    XDG_HIP_CHECK(hipStreamSynchronize(pick(c, stream)));
    return XDG_HIP_OK;
  });
}

int xdg_hip_build_surface_tree(xdg_hip_ctx* c, int64_t n_verts, const double* verts, int64_t n_tris,
                               const int32_t* tris, const int32_t* face_ids, const int32_t* surf_ids,
                               const uint8_t* flip, double bump, int32_t* tree_out)
{
  return guarded(c, [&]() {
    require(tree_out != nullptr, "tree_out is null");
    require(n_tris >= 0 && n_verts >= 0, "negative size");
    require(n_tris == 0 || (verts && tris && face_ids && surf_ids && flip), "null input array");
    require(n_tris < (1LL << 28), "too many triangles for one tree");
    require(std::isfinite(bump) && bump >= 0.0, "bump must be finite and >= 0");
    std::vector<double> v9((size_t)n_tris * 9);
    for (int64_t i = 0; i < n_tris; ++i) {
      for (int j = 0; j < 3; ++j) {
        int32_t vi = tris[3 * i + j];
        require(vi >= 0 && vi < n_verts, "triangle vertex index out of range");
        for (int k = 0; k < 3; ++k) v9[9 * i + 3 * j + k] = verts[3 * (int64_t)vi + k];
      }
      require(surf_ids[i] >= 0 && surf_ids[i] < (1 << 27), "surface id out of range [0, 2^27)");
    }
    int32_t min_face = 0, max_face = -1;
    if (n_tris > 0) {
      min_face = *std::min_element(face_ids, face_ids + n_tris);
      max_face = *std::max_element(face_ids, face_ids + n_tris);
    }
    hipStream_t s = c->stream;
    SurfTreeHost t;
    std::vector<DevBuf> tmp;
    try {
      size_t nt = (size_t)std::max<int64_t>(n_tris, 1);
      double* d_v9 = (double*)dev_alloc(nt * 9 * sizeof(double), tmp);
      int32_t* d_face = (int32_t*)dev_alloc(nt * sizeof(int32_t), tmp);
      int32_t* d_surf = (int32_t*)dev_alloc(nt * sizeof(int32_t), tmp);
      uint8_t* d_flip = (uint8_t*)dev_alloc(nt, tmp);
      double* d_box = (double*)dev_alloc(nt * 6 * sizeof(double), tmp);
      if (n_tris > 0) {
        XDG_HIP_CHECK(hipMemcpyAsync(d_v9, v9.data(), (size_t)n_tris * 9 * sizeof(double), hipMemcpyHostToDevice, s));
        XDG_HIP_CHECK(hipMemcpyAsync(d_face, face_ids, (size_t)n_tris * 4, hipMemcpyHostToDevice, s));
        XDG_HIP_CHECK(hipMemcpyAsync(d_surf, surf_ids, (size_t)n_tris * 4, hipMemcpyHostToDevice, s));
        XDG_HIP_CHECK(hipMemcpyAsync(d_flip, flip, (size_t)n_tris, hipMemcpyHostToDevice, s));
      }
      launch_tri_boxes(d_v9, n_tris, bump, d_box, s);
      const bool keep = c->refit != 0;
      LbvhOut lb;
      lbvh_build(d_box, n_tris, c->leaf_max, s, t.bufs, lb, keep, c->build_sah != 0);
      TriRec* d_tris = (TriRec*)dev_alloc(nt * sizeof(TriRec), t.bufs);
      launch_write_tris(d_v9, d_face, d_surf, d_flip, lb.order, n_tris, d_tris, s);
      double* d_nrm = (double*)dev_alloc(nt * 3 * sizeof(double), t.bufs);
      launch_tri_normals(d_tris, n_tris, d_nrm, s);
      t.dev.normals = d_nrm;
      XDG_HIP_CHECK(hipGetLastError());
      XDG_HIP_CHECK(hipStreamSynchronize(s));
      Bvh4Out b4;
      bvh4_collapse(lb, n_tris, s, t.bufs, b4, keep);
      if (keep) {
        t.refittable = true;
        t.lb = lb;
        t.b4 = b4;
        t.n_verts = n_verts;
        t.tri_idx = (int32_t*)dev_alloc(sizeof(int32_t) * 3 * nt, t.bufs);
        if (n_tris > 0)
          XDG_HIP_CHECK(hipMemcpyAsync(t.tri_idx, tris, sizeof(int32_t) * 3 * (size_t)n_tris, hipMemcpyHostToDevice, s));
      }
      t.dev.nodes4_d = b4.nodes_d;
      t.dev.nodes4_f = b4.nodes_f;
      t.dev.root4 = b4.root;
      t.dev.stack4 = b4.stack;
      t.n_nodes4 = b4.n_nodes;
      t.dev.nodes_d = lb.nodes_d;
      t.dev.nodes_f = lb.nodes_f;
      t.dev.tris = d_tris;
      t.dev.root = lb.root;
      t.dev.n_prims = (int32_t)n_tris;
      t.dev.depth = lb.depth;
      t.dev.bump = bump;
      for (int k = 0; k < 3; ++k) {
        double lo = 0.0, hi = 0.0;
        for (int64_t i = 0; i < n_tris * 3; ++i) {
          const double x = v9[3 * i + k];
          if (i == 0 || x < lo) lo = x;
          if (i == 0 || x > hi) hi = x;
        }
        t.dev.box_lo[k] = f32_lo(lo);
        t.dev.box_hi[k] = f32_hi(hi);
      }
      // FP32 box test error for origin o: at most (3u max|x| + 5u |o|) / |d_i|
      // in ray-parameter units (u = 2^-24) against a margin of bump / |d_i|,
      // so bump * 2^20 - max|x| keeps it below 5/16 of the margin.  The 2^26
      // cap keeps fl32(o) times the clamped reciprocal 1e30 of a near-axis
      // direction finite.
      t.dev.o_limit = mixed_limit(bump, lb.max_abs);
      t.n_nodes = lb.n_nodes;
      t.bytes = bufs_bytes(t.bufs);
      t.min_face = min_face;
      t.max_face = max_face;
    } catch (...) {
      dev_free_all(tmp);
      dev_free_all(t.bufs);
      throw;
    }
    dev_free_all(tmp);
    c->surf.push_back(std::move(t));
    c->max_surf_depth = std::max(c->max_surf_depth, c->surf.back().dev.depth);
    c->max_surf_stack4 = std::max(c->max_surf_stack4, c->surf.back().dev.stack4);
    try {
      sync_tables(c);
    } catch (...) {
      dev_free_all(c->surf.back().bufs);
      c->surf.pop_back();
      throw;
    }
    *tree_out = (int32_t)c->surf.size();
    return XDG_HIP_OK;
  });
}

int xdg_hip_build_element_tree(xdg_hip_ctx* c, int64_t n_verts, const double* verts, int64_t n_tets,
                               const int32_t* tets, const int32_t* adj, const int32_t* face_order,
                               int32_t* tree_out)
{
  return guarded(c, [&]() {
    require(tree_out != nullptr, "tree_out is null");
    require(n_tets >= 0 && n_verts >= 0, "negative size");
    require(n_tets == 0 || (verts && tets && adj && face_order), "null input array");
    require(n_tets < (1LL << 28), "too many elements for one tree");
    if (face_order) {
      for (int f = 0; f < 12; ++f) require(face_order[f] >= 0 && face_order[f] < 4, "face_order entries must be 0..3");
    }
    std::vector<ElemRec> recs((size_t)n_tets);
    for (int64_t e = 0; e < n_tets; ++e) {
      ElemRec& r = recs[e];
      for (int j = 0; j < 4; ++j) {
        int32_t vi = tets[4 * e + j];
        require(vi >= 0 && vi < n_verts, "tet vertex index out of range");
        if (j < 3) r.conn[j] = vi;
        for (int k = 0; k < 3; ++k) r.v[3 * j + k] = verts[3 * (int64_t)vi + k];
        int32_t a = adj[4 * e + j];
        require(a >= -1 && a < n_tets, "adjacency out of range");
        r.adj[j] = a;
      }
      // lexicographic endpoint order of the six edges (plucker.h:35-43)
      static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
      r.edge_first = 0;
      for (int k = 0; k < 6; ++k) {
        const double* a = r.v + 3 * pa[k];
        const double* b = r.v + 3 * pb[k];
        const bool f = a[0] < b[0] || (!(a[0] > b[0]) && (a[1] < b[1] || (!(a[1] > b[1]) && a[2] < b[2])));
        r.edge_first |= (f ? 1u : 0u) << k;
      }
    }
    hipStream_t s = c->stream;
    ElemTreeHost t;
    std::vector<DevBuf> tmp;
    try {
      size_t ne = (size_t)std::max<int64_t>(n_tets, 1);
      ElemRec* d_elems = (ElemRec*)dev_alloc(ne * sizeof(ElemRec), t.bufs);
      double* d_box = (double*)dev_alloc(ne * 6 * sizeof(double), tmp);
      if (n_tets > 0)
        XDG_HIP_CHECK(hipMemcpyAsync(d_elems, recs.data(), (size_t)n_tets * sizeof(ElemRec), hipMemcpyHostToDevice, s));
      launch_elem_boxes(d_elems, n_tets, d_box, s);
      LbvhOut lb;
      lbvh_build(d_box, n_tets, c->leaf_max, s, t.bufs, lb, false, c->build_sah != 0);
      XDG_HIP_CHECK(hipGetLastError());
      XDG_HIP_CHECK(hipStreamSynchronize(s));
      t.dev.nodes_d = lb.nodes_d;
      t.dev.nodes_f = lb.nodes_f;
      t.dev.order = lb.order;
      t.dev.elems = d_elems;
      t.dev.root = lb.root;
      t.dev.n_prims = (int32_t)n_tets;
      t.dev.depth = lb.depth;
      for (int f = 0; f < 12; ++f) t.dev.face_order[f] = face_order ? face_order[f] : 0;
      t.dev.ipc_tree = -1;
      for (int k = 0; k < 3; ++k) {
        double lo = 0.0, hi = 0.0;
        for (int64_t v = 0; v < n_verts; ++v) {
          const double x = verts[3 * v + k];
          if (v == 0 || x < lo) lo = x;
          if (v == 0 || x > hi) hi = x;
        }
        t.dev.box_lo[k] = f32_lo(lo);
        t.dev.box_hi[k] = f32_hi(hi);
      }
      t.dev.bface_elem = nullptr;
      t.n_nodes = lb.n_nodes;
      t.bytes = bufs_bytes(t.bufs);
    } catch (...) {
      dev_free_all(tmp);
      dev_free_all(t.bufs);
      throw;
    }
    dev_free_all(tmp);
    c->elem.push_back(std::move(t));
    c->max_elem_depth = std::max(c->max_elem_depth, c->elem.back().dev.depth);
    try {
      sync_tables(c);
    } catch (...) {
      dev_free_all(c->elem.back().bufs);
      c->elem.pop_back();
      throw;
    }
    *tree_out = (int32_t)c->elem.size();
    return XDG_HIP_OK;
  });
}

int xdg_hip_attach_boundary(xdg_hip_ctx* c, int32_t element_tree, int32_t surface_tree, int64_t n_faces,
                            const int32_t* bface_elem)
{
  return guarded(c, [&]() {
    elem_tree(c, element_tree);
    const SurfTreeHost& st = surf_tree(c, surface_tree);
    require(n_faces >= 0 && (n_faces == 0 || bface_elem), "null bface_elem");
    require(st.dev.n_prims == 0 || (st.min_face >= 0 && st.max_face < n_faces),
            "surface tree face ids must index bface_elem");
    ElemTreeHost& et = c->elem[element_tree - 1];
    for (int64_t i = 0; i < n_faces; ++i)
      require(bface_elem[i] >= -1 && bface_elem[i] < et.dev.n_prims, "bface_elem out of range");
    int32_t* d = (int32_t*)dev_alloc(sizeof(int32_t) * std::max<int64_t>(n_faces, 1), et.bufs);
    if (n_faces > 0) XDG_HIP_CHECK(hipMemcpy(d, bface_elem, sizeof(int32_t) * n_faces, hipMemcpyHostToDevice));
    et.dev.ipc_tree = surface_tree;
    et.dev.bface_elem = d;
    et.n_bfaces = n_faces;
    et.bytes = bufs_bytes(et.bufs);
    sync_tables(c);
    return XDG_HIP_OK;
  });
}

int xdg_hip_get_tree_info(xdg_hip_ctx* c, int32_t tree, int element, xdg_hip_tree_info* info)
{
  return guarded(c, [&]() {
    require(info != nullptr, "info is null");
    if (element) {
      const ElemTreeHost& t = elem_tree(c, tree);
      info->n_prims = t.dev.n_prims;
      info->n_nodes = t.n_nodes;
      info->depth = t.dev.depth;
      info->root = t.dev.root;
      info->bytes = t.bytes;
      info->bump = 0.0;
      info->o_limit = 0.0;
    } else {
      const SurfTreeHost& t = surf_tree(c, tree);
      info->n_prims = t.dev.n_prims;
      info->n_nodes = t.n_nodes;
      info->depth = t.dev.depth;
      info->root = t.dev.root;
      info->bytes = t.bytes;
      info->bump = t.dev.bump;
      info->o_limit = t.dev.o_limit;
      info->n_nodes4 = t.n_nodes4;
      info->stack4 = t.dev.stack4;
    }
    return XDG_HIP_OK;
  });
}

int xdg_hip_release_surface_tree(xdg_hip_ctx* c, int32_t tree)
{
  return guarded(c, [&]() {
    surf_tree(c, tree);
    SurfTreeHost& t = c->surf[tree - 1];
    XDG_HIP_CHECK(hipDeviceSynchronize());
    dev_free_all(t.bufs);
    t.dev = SurfTreeDev{};
    t.released = true;
    sync_tables(c);
    return XDG_HIP_OK;
  });
}

int xdg_hip_refit_surface_tree(xdg_hip_ctx* c, int32_t tree, int64_t n_verts, const double* verts, double bump)
{
  return guarded(c, [&]() {
    surf_tree(c, tree);
    SurfTreeHost& t = c->surf[tree - 1];
    require(t.refittable, "the tree was not built with the \"refit\" option set");
    require(n_verts == t.n_verts, "n_verts differs from the build");
    require(n_verts == 0 || verts != nullptr, "null verts");
    require(std::isfinite(bump), "bump must be finite");
    const double b = bump >= 0.0 ? bump : t.dev.bump;
    const int64_t n = t.dev.n_prims;
    hipStream_t s = c->stream;
    XDG_HIP_CHECK(hipDeviceSynchronize());  // no query may read the tree meanwhile
    if (n > 0) {
      std::vector<DevBuf> tmp;
      try {
        double* d_v9 = (double*)dev_alloc(sizeof(double) * 9 * (size_t)n, tmp);
        double* d_box = (double*)dev_alloc(sizeof(double) * 6 * (size_t)n, tmp);
        auto* cb = (unsigned long long*)dev_alloc(sizeof(unsigned long long) * 6, tmp);
        const unsigned g = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_gather_v9, dim3(g), dim3(256), 0, s, verts, t.tri_idx, n, d_v9);
        launch_tri_boxes(d_v9, n, b, d_box, s);
        hipLaunchKernelGGL(k_refit_tris, dim3(g), dim3(256), 0, s, d_v9, t.lb.order, n, (TriRec*)t.dev.tris);
        launch_tri_normals(t.dev.tris, n, (double*)t.dev.normals, s);
        hipLaunchKernelGGL(k_init_minmax, dim3(1), dim3(64), 0, s, cb);
        hipLaunchKernelGGL(k_v9_bounds, dim3(g), dim3(256), 0, s, d_v9, n, cb);
        lbvh_refit(d_box, n, t.lb, s);
        bvh4_refit(t.lb, t.b4, s);
        unsigned long long h[6];
        XDG_HIP_CHECK(hipMemcpyAsync(h, cb, sizeof h, hipMemcpyDeviceToHost, s));
        XDG_HIP_CHECK(hipStreamSynchronize(s));
        for (int k = 0; k < 3; ++k) {
          t.dev.box_lo[k] = f32_lo(ord_val(h[k]));
          t.dev.box_hi[k] = f32_hi(ord_val(h[3 + k]));
        }
        if (n == 1) {  // no internal nodes: the refit kept no max |x|
          double mx = 0.0;
          for (int k = 0; k < 6; ++k) mx = std::max(mx, std::fabs(ord_val(h[k])));
          t.lb.max_abs = mx + b;
        }
      } catch (...) {
        dev_free_all(tmp);
        throw;
      }
      dev_free_all(tmp);
    }
    t.dev.bump = b;
    t.dev.o_limit = mixed_limit(b, t.lb.max_abs);
    sync_tables(c);
    return XDG_HIP_OK;
  });
}

int xdg_hip_num_trees(xdg_hip_ctx* c, int32_t* surface_trees, int32_t* element_trees)
{
  return guarded(c, [&]() {
    if (surface_trees) *surface_trees = (int32_t)c->surf.size();
    if (element_trees) *element_trees = (int32_t)c->elem.size();
    return XDG_HIP_OK;
  });
}

int xdg_hip_ray_fire(xdg_hip_ctx* c, int32_t tree, const int32_t* tree_per_ray, int64_t n, const double* org,
                     const double* dir, const double* tmax, int orientation, const int32_t* excl, int k,
                     int precision, double* t_out, int32_t* surf_out, int32_t* face_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(orientation >= -1 && orientation <= 1, "unknown orientation");
    require(k >= 0 && k <= XDG_HIP_MAX_EXCLUDE, "k must be in [0, 1024]");
    int depth = tree_per_ray ? surf_need(c, nullptr) : surf_need(c, &surf_tree(c, tree));
    if (n == 0) return XDG_HIP_OK;
    require(org && dir && t_out && surf_out, "null array");
    RayFireArgs a{c->d_surf, (int32_t)c->surf.size(), tree, tree_per_ray, n, org, dir, tmax, orientation,
                  excl, excl ? k : 0, t_out, surf_out, face_out, nullptr, nullptr};
    if (precision == XDG_HIP_MIXED && c->cfg.defer) {
      a.fb_list = fb_list_for(c, n);
      a.fb_count = c->d_fb_count;
    }
    a.next_ray = c->d_next;
    const hipStream_t s = pick(c, stream);
    scratch_acquire(c, s);
    if (c->cfg.order > 0 && n >= ORDER_MIN_RAYS) {
      RayOrderArgs o{c->d_surf, (int32_t)c->surf.size(), tree, tree_per_ray, n, org, dir};
      a.order = launch_ray_order(o, c->cfg.order, order_scratch(c, n), s);
    }
    launch_ray_fire(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), s);
    XDG_HIP_CHECK(hipGetLastError());
    scratch_release(c, s);
    return XDG_HIP_OK;
  });
}

int xdg_hip_ray_fire_noalloc(xdg_hip_ctx* c, int32_t tree, const int32_t* tree_per_ray, int64_t n, const double* org,
                             const double* dir, const double* tmax, int orientation, const int32_t* excl, int k,
                             int precision, double* t_out, int32_t* surf_out, int32_t* face_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(n <= XDG_HIP_NOALLOC_MAX_RAYS,
            "xdg_hip_ray_fire_noalloc: n exceeds XDG_HIP_NOALLOC_MAX_RAYS (4194304 rays); split the batch");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(orientation >= -1 && orientation <= 1, "unknown orientation");
    require(k >= 0 && k <= XDG_HIP_MAX_EXCLUDE, "k must be in [0, 1024]");
    int depth = tree_per_ray ? surf_need(c, nullptr) : surf_need(c, &surf_tree(c, tree));
    if (n == 0) return XDG_HIP_OK;
    require(org && dir && t_out && surf_out, "null array");
    NoAllocScratch* sp = scratch_for(c->device);
    if (!sp) throw HipError{"allocation-free scratch is not set up on this device"};
    char* base = (char*)sp;  // device address: offsets only, never dereferenced here
    RayFireArgs a{c->d_surf, (int32_t)c->surf.size(), tree, tree_per_ray, n, org, dir, tmax, orientation,
                  excl, excl ? k : 0, t_out, surf_out, face_out, nullptr, nullptr};
    if (precision == XDG_HIP_MIXED && c->cfg.defer) {
      a.fb_list = (int32_t*)(base + offsetof(NoAllocScratch, fb_list));
      a.fb_count = (unsigned int*)(base + offsetof(NoAllocScratch, fb_count));
    }
    a.next_ray = (unsigned long long*)(base + offsetof(NoAllocScratch, next));
    // traced in the caller's order: the ray-ordering option needs buffers
    // sized by the batch, which this entry point does not allocate
    launch_ray_fire(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_ray_fire_stats(xdg_hip_ctx* c, int32_t tree, const int32_t* tree_per_ray, int64_t n, const double* org,
                           const double* dir, const double* tmax, int orientation, const int32_t* excl, int k,
                           int precision, uint64_t* counters)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(counters != nullptr, "counters is null");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(orientation >= -1 && orientation <= 1, "unknown orientation");
    require(k >= 0 && k <= XDG_HIP_MAX_EXCLUDE, "k must be in [0, 1024]");
    int depth = tree_per_ray ? surf_need(c, nullptr) : surf_need(c, &surf_tree(c, tree));
    for (int j = 0; j < XDG_HIP_NUM_COUNTERS; ++j) counters[j] = 0;
    if (n == 0) return XDG_HIP_OK;
    require(org && dir, "null array");
    std::vector<DevBuf> tmp;
    try {
      auto* d_cnt = (unsigned long long*)dev_alloc(XDG_HIP_NUM_COUNTERS * sizeof(unsigned long long), tmp);
      XDG_HIP_CHECK(hipMemcpy(d_cnt, counters, XDG_HIP_NUM_COUNTERS * sizeof(uint64_t), hipMemcpyHostToDevice));
      RayFireArgs a{c->d_surf, (int32_t)c->surf.size(), tree, tree_per_ray, n, org, dir, tmax, orientation,
                    excl, excl ? k : 0, nullptr, nullptr, nullptr};
      scratch_acquire(c, c->stream);
      if (c->cfg.order > 0 && n >= ORDER_MIN_RAYS) {
        RayOrderArgs o{c->d_surf, (int32_t)c->surf.size(), tree, tree_per_ray, n, org, dir};
        a.order = launch_ray_order(o, c->cfg.order, order_scratch(c, n), c->stream);
      }
      launch_ray_fire_stats(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), d_cnt, c->stream);
      XDG_HIP_CHECK(hipGetLastError());
      scratch_release(c, c->stream);
      XDG_HIP_CHECK(hipStreamSynchronize(c->stream));
      XDG_HIP_CHECK(hipMemcpy(counters, d_cnt, XDG_HIP_NUM_COUNTERS * sizeof(uint64_t), hipMemcpyDeviceToHost));
    } catch (...) {
      dev_free_all(tmp);
      throw;
    }
    dev_free_all(tmp);
    return XDG_HIP_OK;
  });
}

int xdg_hip_point_in_volume(xdg_hip_ctx* c, int32_t tree, int64_t n, const double* pts, const double* dir,
                            const int32_t* excl, int k, int precision, uint8_t* inside_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(k >= 0 && k <= XDG_HIP_MAX_EXCLUDE, "k must be in [0, 1024]");
    const SurfTreeHost& t = surf_tree(c, tree);
    if (n == 0) return XDG_HIP_OK;
    require(pts && inside_out, "null array");
    PivArgs a{};
    a.table = c->d_surf;
    a.ntrees = (int32_t)c->surf.size();
    a.tree = tree;
    a.n = n;
    a.pts = pts;
    a.dir = dir;
    // ray_tracer.cpp:268: {1. / std::sqrt(2.0), 1 / std::sqrt(2.0), 0.0}
    a.dflt[0] = 1. / std::sqrt(2.0);
    a.dflt[1] = 1 / std::sqrt(2.0);
    a.dflt[2] = 0.0;
    a.excl = excl;
    a.k = excl ? k : 0;
    a.inside = inside_out;
    launch_point_in_volume(a, precision == XDG_HIP_MIXED, cfg_for(c, surf_need(c, &t)), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

// Device copy of a (trees, ids) volume list, cached across calls with the
// same list; a changed list waits for queued work that may still read the
// old one.
static const int32_t* volume_lists(xdg_hip_ctx* c, int nvol, const int32_t* vol_trees, const int32_t* vol_ids)
{
  std::vector<int32_t> lists(vol_trees, vol_trees + nvol);
  lists.insert(lists.end(), vol_ids, vol_ids + nvol);
  if (lists != c->small_cache) {
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (lists.size() > c->small_cap) {
      if (c->d_small) XDG_HIP_CHECK(hipFree(c->d_small));
      c->d_small = nullptr;
      c->small_cap = std::max<size_t>(lists.size(), 64);
      XDG_HIP_CHECK(hipMalloc(&c->d_small, c->small_cap * sizeof(int32_t)));
    }
    if (!lists.empty())
      XDG_HIP_CHECK(hipMemcpy(c->d_small, lists.data(), lists.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    c->small_cache = lists;
  }
  return c->d_small;
}

int xdg_hip_find_volume(xdg_hip_ctx* c, int nvol, const int32_t* vol_trees, const int32_t* vol_ids,
                        int32_t ipc_volume, int64_t n, const double* pts, const double* dir, int precision,
                        int32_t* vol_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0 && nvol >= 0, "negative size");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(nvol == 0 || (vol_trees && vol_ids), "null volume lists");
    int depth = 1;
    for (int v = 0; v < nvol; ++v) {
      if (vol_ids[v] == ipc_volume) continue;
      depth = std::max(depth, surf_need(c, &surf_tree(c, vol_trees[v])));
    }
    if (n == 0) return XDG_HIP_OK;
    require(pts && vol_out, "null array");
    hipStream_t s = pick(c, stream);
    volume_lists(c, nvol, vol_trees, vol_ids);
    FindVolumeArgs a{};
    a.table = c->d_surf;
    a.ntrees = (int32_t)c->surf.size();
    a.nvol = nvol;
    a.vol_trees = c->d_small;
    a.vol_ids = c->d_small + nvol;
    a.ipc = ipc_volume;
    a.n = n;
    a.pts = pts;
    a.dir = dir;
    a.dflt[0] = 1. / std::sqrt(2.0);
    a.dflt[1] = 1 / std::sqrt(2.0);
    a.dflt[2] = 0.0;
    a.vol_out = vol_out;
    a.next = c->d_next + 1;
    if (precision == XDG_HIP_MIXED) {
      a.fb_list = fb_list_for(c, n);
      a.fb_count = c->d_fb_count;
    }
    scratch_acquire(c, s);
    launch_find_volume(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), s);
    XDG_HIP_CHECK(hipGetLastError());
    scratch_release(c, s);
    return XDG_HIP_OK;
  });
}

int xdg_hip_point_in_volumes(xdg_hip_ctx* c, int nvol, const int32_t* vol_trees, const int32_t* vol_ids, int64_t n,
                             const double* pts, const double* dir, int max_k, int precision, int32_t* count_out,
                             int32_t* ids_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0 && nvol >= 0, "negative size");
    require(max_k >= 1, "max_k must be >= 1");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(nvol == 0 || (vol_trees && vol_ids), "null volume lists");
    int depth = 1;
    for (int v = 0; v < nvol; ++v) depth = std::max(depth, surf_need(c, &surf_tree(c, vol_trees[v])));
    if (n == 0) return XDG_HIP_OK;
    require(pts && dir && count_out && ids_out, "null array");
    const int32_t* d_lists = volume_lists(c, nvol, vol_trees, vol_ids);
    VolumesArgs a{};
    a.table = c->d_surf;
    a.ntrees = (int32_t)c->surf.size();
    a.nvol = nvol;
    a.vol_trees = d_lists;
    a.vol_ids = d_lists + nvol;
    a.n = n;
    a.pts = pts;
    a.dir = dir;
    a.max_k = max_k;
    a.count_out = count_out;
    a.ids_out = ids_out;
    launch_point_in_volumes(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_edge_rays(xdg_hip_ctx* c, int nvol, const int32_t* vol_trees, const int32_t* vol_ids, int64_t n,
                      const double* org, const double* dir, const double* len, const int32_t* skip, int precision,
                      double* t_out, int32_t* surf_out, int32_t* vol_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0 && nvol >= 0, "negative size");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(nvol == 0 || (vol_trees && vol_ids), "null volume lists");
    int depth = 1;
    for (int v = 0; v < nvol; ++v) depth = std::max(depth, surf_need(c, &surf_tree(c, vol_trees[v])));
    if (n == 0) return XDG_HIP_OK;
    require(org && dir && len && skip && t_out && surf_out && vol_out, "null array");
    const int32_t* d_lists = volume_lists(c, nvol, vol_trees, vol_ids);
    VolumesArgs a{};
    a.table = c->d_surf;
    a.ntrees = (int32_t)c->surf.size();
    a.nvol = nvol;
    a.vol_trees = d_lists;
    a.vol_ids = d_lists + nvol;
    a.n = n;
    a.pts = org;
    a.dir = dir;
    a.len = len;
    a.skip = skip;
    a.t_out = t_out;
    a.surf_out = surf_out;
    a.vol_out = vol_out;
    launch_edge_rays(a, precision == XDG_HIP_MIXED, cfg_for(c, depth), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_occluded(xdg_hip_ctx* c, int32_t tree, int64_t n, const double* org, const double* dir,
                     int precision, uint8_t* hit_out, double* dist_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    const SurfTreeHost& t = surf_tree(c, tree);
    if (n == 0) return XDG_HIP_OK;
    require(org && dir && hit_out, "null array");
    OccludedArgs a{c->d_surf, (int32_t)c->surf.size(), tree, n, org, dir, hit_out, dist_out};
    launch_occluded(a, precision == XDG_HIP_MIXED, cfg_for(c, surf_need(c, &t)), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_closest(xdg_hip_ctx* c, int32_t tree, int64_t n, const double* pts, double* dist_out,
                    int32_t* face_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    const SurfTreeHost& t = surf_tree(c, tree);
    if (n == 0) return XDG_HIP_OK;
    require(pts && dist_out && face_out, "null array");
    ClosestArgs a{c->d_surf, (int32_t)c->surf.size(), tree, n, pts, dist_out, face_out};
    launch_closest(a, cfg_for(c, t.dev.depth), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_find_element(xdg_hip_ctx* c, int32_t element_tree, int64_t n, const double* pts, int32_t* elem_out,
                         void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    const ElemTreeHost& t = elem_tree(c, element_tree);
    if (n == 0) return XDG_HIP_OK;
    require(pts && elem_out, "null array");
    FindElementArgs a{c->d_elem, (int32_t)c->elem.size(), element_tree, n, pts, elem_out};
    launch_find_element(a, cfg_for(c, t.dev.depth), pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_next_element(xdg_hip_ctx* c, int32_t element_tree, int64_t n, const int32_t* elem, const double* r,
                         const double* u, int32_t* next_out, double* dist_out, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    elem_tree(c, element_tree);
    if (n == 0) return XDG_HIP_OK;
    require(elem && r && u && next_out && dist_out, "null array");
    NextElementArgs a{c->d_elem, (int32_t)c->elem.size(), element_tree, n, elem, r, u, next_out, dist_out};
    launch_next_element(a, pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_tally_segments(xdg_hip_ctx* c, int32_t element_tree, int64_t n, const double* p0, const double* p1,
                           int64_t* tally_fx, int32_t* count_out, uint64_t* hash_out, const int64_t* seq_off,
                           int32_t* seq_elem, double* seq_len, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    const ElemTreeHost& t = elem_tree(c, element_tree);
    int depth = t.dev.depth;
    if (t.dev.ipc_tree >= 1) depth = std::max(depth, surf_tree(c, t.dev.ipc_tree).dev.depth);
    if (n == 0) return XDG_HIP_OK;
    require(p0 && p1 && count_out && hash_out, "null array");
    require(!seq_off || (seq_elem && seq_len), "seq_off given without seq_elem/seq_len");
    SegmentsArgs a{};
    a.etable = c->d_elem;
    a.nelem_trees = (int32_t)c->elem.size();
    a.stable = c->d_surf;
    a.nsurf_trees = (int32_t)c->surf.size();
    a.tree = element_tree;
    a.n = n;
    a.p0 = p0;
    a.p1 = p1;
    a.tally_fx = (unsigned long long*)tally_fx;
    a.count = count_out;
    a.hash = hash_out;
    a.seq_off = seq_off;
    a.seq_elem = seq_elem;
    a.seq_len = seq_len;
    a.status = c->d_status;
    const hipStream_t s = pick(c, stream);
    const bool ordered = c->cfg.seg_order && n >= ORDER_MIN_RAYS;
    if (ordered) {
      scratch_acquire(c, s);
      a.order = launch_segment_order(a, c->cfg.seg_order, order_scratch(c, n), s);
    }
    launch_segments(a, face_order_kind(t.dev.face_order), cfg_for(c, depth), s);
    tally_guard(c, tally_fx, t.dev.n_prims, XDG_HIP_STATUS_TALLY_RANGE, s);
    XDG_HIP_CHECK(hipGetLastError());
    if (ordered) scratch_release(c, s);
    return XDG_HIP_OK;
  });
}

int xdg_hip_tally_segments_volume(xdg_hip_ctx* c, int32_t element_tree, int32_t surface_tree, int64_t n,
                                  const double* p0, const double* p1, int64_t* tally_fx, int32_t* count_out,
                                  uint64_t* hash_out, const int64_t* seq_off, int32_t* seq_elem, double* seq_len,
                                  void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    const ElemTreeHost& t = elem_tree(c, element_tree);
    int depth = t.dev.depth;
    depth = std::max(depth, surf_tree(c, surface_tree).dev.depth);
    if (n == 0) return XDG_HIP_OK;
    require(p0 && p1 && count_out && hash_out, "null array");
    require(!seq_off || (seq_elem && seq_len), "seq_off given without seq_elem/seq_len");
    SegmentsArgs a{};
    a.etable = c->d_elem;
    a.nelem_trees = (int32_t)c->elem.size();
    a.stable = c->d_surf;
    a.nsurf_trees = (int32_t)c->surf.size();
    a.tree = element_tree;
    a.n = n;
    a.p0 = p0;
    a.p1 = p1;
    a.tally_fx = (unsigned long long*)tally_fx;
    a.count = count_out;
    a.hash = hash_out;
    a.seq_off = seq_off;
    a.seq_elem = seq_elem;
    a.seq_len = seq_len;
    a.status = c->d_status;
    a.volume_mode = 1;
    a.surf_tree = surface_tree;
    launch_segments(a, face_order_kind(t.dev.face_order), cfg_for(c, depth), pick(c, stream));
    tally_guard(c, tally_fx, t.dev.n_prims, XDG_HIP_STATUS_TALLY_RANGE, pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_mesh_walk(xdg_hip_ctx* c, int32_t element_tree, uint64_t first, int64_t n, uint32_t seed,
                      uint32_t stream_id, const double* lo, const double* hi, double mfp, int32_t max_events,
                      double* distance_out, int32_t* events_out, uint8_t* status_out, uint64_t* hash_out,
                      int64_t* tally_fx, void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0, "negative n");
    require(lo && hi, "null start box");
    for (int k = 0; k < 3; ++k) require(std::isfinite(lo[k]) && std::isfinite(hi[k]), "start box not finite");
    require(std::isfinite(mfp) && mfp > 0.0, "mfp must be finite and > 0");
    require(max_events >= 1, "max_events must be >= 1");
    const ElemTreeHost& t = elem_tree(c, element_tree);
    int depth = t.dev.depth;
    if (t.dev.ipc_tree >= 1) depth = std::max(depth, surf_tree(c, t.dev.ipc_tree).dev.depth);
    if (n == 0) return XDG_HIP_OK;
    MeshWalkArgs a{};
    a.etable = c->d_elem;
    a.nelem_trees = (int32_t)c->elem.size();
    a.stable = c->d_surf;
    a.nsurf_trees = (int32_t)c->surf.size();
    a.tree = element_tree;
    a.first = first;
    a.n = n;
    a.seed = seed;
    a.stream = stream_id;
    for (int k = 0; k < 3; ++k) {
      a.lo[k] = lo[k];
      a.hi[k] = hi[k];
    }
    a.mfp = mfp;
    a.max_events = max_events;
    a.distance = distance_out;
    a.events = events_out;
    a.status = status_out;
    a.hash = hash_out;
    a.tally_fx = (unsigned long long*)tally_fx;
    const LaunchCfg cfg = cfg_for(c, depth);
    const bool wave = cfg.mw_wave && t.dev.ipc_tree >= 1;
    const hipStream_t s = pick(c, stream);
    if (wave) scratch_acquire(c, s);
    launch_mesh_walk(a, face_order_kind(t.dev.face_order), cfg, wave ? &mw_scratch_for(c, n) : nullptr, s);
    tally_guard(c, tally_fx, t.dev.n_prims, XDG_HIP_STATUS_TALLY_RANGE, s);
    XDG_HIP_CHECK(hipGetLastError());
    if (wave) scratch_release(c, s);
    return XDG_HIP_OK;
  });
}

int xdg_hip_status(xdg_hip_ctx* c, uint32_t* flags)
{
  return guarded(c, [&]() {
    require(flags != nullptr, "flags is null");
    XDG_HIP_CHECK(hipDeviceSynchronize());
    XDG_HIP_CHECK(hipMemcpy(flags, c->d_status, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t zero = 0;
    XDG_HIP_CHECK(hipMemcpy(c->d_status, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    return XDG_HIP_OK;
  });
}

int xdg_hip_eval_primitive(xdg_hip_ctx* c, int kind, int64_t n, const double* in, double* out, void* stream)
{
  return guarded(c, [&]() {
    require(kind >= 0 && kind <= 5, "unknown primitive kind");
    require(n >= 0, "negative n");
    if (n == 0) return XDG_HIP_OK;
    require(in && out, "null array");
    launch_eval(kind, n, in, out, pick(c, stream));
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_ray_distance_adjoint(xdg_hip_ctx* c, int64_t n, const double* org, const double* dir, const double* t,
                                 const int32_t* face, const double* dldt, int64_t nv, const double* verts,
                                 int64_t ntri, const int32_t* tris, double* dldv, double* dldo, double* dldd,
                                 void* stream)
{
  return guarded(c, [&]() {
    require(n >= 0 && nv >= 0 && ntri >= 0, "negative size");
    if (n == 0) return XDG_HIP_OK;
    require(org && dir && t && face && dldt, "null ray array");
    require(!dldv || (verts && tris), "dldv needs verts and tris");
    require(verts || !dldv, "null verts");
    require(ntri == 0 || tris, "null tris");
    AdjointArgs a{n, org, dir, t, face, dldt, nv, verts, ntri, tris, dldv, dldo, dldd};
    if (!verts) a.ntri = 0;  // every ray reads as a miss without the mesh
    const hipStream_t s = pick(c, stream);
    if (c->cfg.adj_sort && dldv && n >= ORDER_MIN_RAYS) {
      scratch_acquire(c, s);
      launch_adjoint_sorted(a, order_scratch(c, n), s);
      XDG_HIP_CHECK(hipGetLastError());
      scratch_release(c, s);
    } else {
      launch_adjoint(a, 256, s);
      XDG_HIP_CHECK(hipGetLastError());
    }
    return XDG_HIP_OK;
  });
}

int xdg_hip_malloc(xdg_hip_ctx* c, size_t bytes, void** ptr)
{
  return guarded(c, [&]() {
    require(ptr != nullptr, "ptr is null");
    XDG_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 16));
    return XDG_HIP_OK;
  });
}

int xdg_hip_free(xdg_hip_ctx* c, void* ptr)
{
  return guarded(c, [&]() {
    if (ptr) XDG_HIP_CHECK(hipFree(ptr));
    return XDG_HIP_OK;
  });
}

int xdg_hip_memcpy_htod(xdg_hip_ctx* c, void* dst, const void* src, size_t bytes)
{
  return guarded(c, [&]() {
    if (bytes) XDG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return XDG_HIP_OK;
  });
}

int xdg_hip_memcpy_dtoh(xdg_hip_ctx* c, void* dst, const void* src, size_t bytes)
{
  return guarded(c, [&]() {
    if (bytes) XDG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return XDG_HIP_OK;
  });
}

}  // extern "C"

// ---- batched particle random walk ------------------------------------------

namespace {

TransportHost& transport_of(xdg_hip_ctx* c, int32_t handle)
{
  if (handle < 1 || (size_t)handle > c->transports.size() || !c->transports[handle - 1])
    throw NotFound{"transport problem " + std::to_string(handle) + " is not registered"};
  return *c->transports[handle - 1];
}

// refresh the tree table pointer (trees may have been added since)
void transport_refresh(xdg_hip_ctx* c, TransportHost& t)
{
  t.dev.table = c->d_surf;
  t.dev.ntrees = (int32_t)c->surf.size();
  int need = 1;
  for (int32_t id : t.host_trees) need = std::max(need, surf_tree(c, id).dev.stack4);
  t.stack_need = need;
  if (t.lds_stack <= 0 || t.lds_stack > need) t.lds_stack = need;
}

int transport_mask_words(const TransportHost& t)
{
  const int w = (t.dev.nvol + 63) / 64;
  return w <= TRANSPORT_MASK_WORDS_MAX ? w : 0;
}

size_t transport_stage_bytes(int64_t cap, int items, int mask_words)
{
  return (size_t)cap * (TRANSPORT_STAGE_BYTES + 8 * (size_t)items + 8 * (size_t)mask_words) +
         4 * (size_t)radix_sort_hist_words(cap) + 256;
}

// histories per start slice: 2^transport_slice, and few enough that the
// 32-bit item counts cannot wrap (one item per history and volume at most)
int64_t transport_slice_cap(xdg_hip_ctx* c, const TransportHost& t)
{
  const int64_t by_count = (int64_t)(UINT32_MAX / (uint64_t)std::max(t.dev.nvol, 1));
  return std::max<int64_t>(1, std::min<int64_t>((int64_t)1 << c->cfg.tr_slice, by_count));
}

bool transport_staged(xdg_hip_ctx* c, const TransportHost& t)
{
  return c->cfg.tr_start != 0 && transport_start_block(t.stack_need) > 0;
}

// global stack share for the shape, and the start buffer of the staged form
// for n histories (outside any capture: may allocate)
TransportLaunch transport_prepare(xdg_hip_ctx* c, TransportHost& t, bool mixed, bool may_alloc, int64_t n)
{
  transport_refresh(c, t);
  const TransportLaunch L = transport_shape(mixed, t.block, t.lds_stack, t.dev.nvol);
  const size_t need = (size_t)(t.stack_need - t.lds_stack) * (size_t)(L.grid * L.block);
  if (need > t.spill_entries) {
    if (!may_alloc) throw InvalidArg{"transport launch needs a larger stack spill area; run it once before capture"};
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (t.spill) XDG_HIP_CHECK(hipFree(t.spill));
    t.spill = nullptr;
    t.spill_entries = 0;
    XDG_HIP_CHECK(hipMalloc(&t.spill, need * sizeof(int32_t)));
    t.spill_entries = need;
    ++t.generation;
  }
  const int64_t cap = std::min<int64_t>(n, transport_slice_cap(c, t));
  if (transport_staged(c, t) && (cap > t.stage_cap || c->cfg.tr_items > t.stage_items)) {
    if (!may_alloc) throw InvalidArg{"transport launch needs a larger start buffer; run it once before capture"};
    const int64_t nc = std::max(cap, t.stage_cap);
    const int ni = std::max(c->cfg.tr_items, t.stage_items);
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (t.stage) XDG_HIP_CHECK(hipFree(t.stage));
    t.stage = nullptr;
    t.stage_cap = 0;
    t.stage_items = 0;
    XDG_HIP_CHECK(hipMalloc(&t.stage, transport_stage_bytes(nc, ni, transport_mask_words(t))));
    t.stage_cap = nc;
    t.stage_items = ni;
    ++t.generation;
  }
  if (transport_staged(c, t)) t.start_grid[mixed ? 1 : 0] = transport_start_grid(mixed, t.stack_need);
  return L;
}

void transport_launch(xdg_hip_ctx* c, TransportHost& t, const TransportLaunch& L, bool mixed, uint64_t first,
                      int64_t n, uint64_t* tally, int32_t* events, uint8_t* status, uint64_t* hash, hipStream_t s)
{
  TransportRun R{};
  R.first = first;
  R.n = n;
  R.next = t.d_next;
  R.tally_fx = (unsigned long long*)tally;
  R.events = events;
  R.status = status;
  R.hash = hash;
  R.chunk = t.chunk;
  R.lds_stack = t.lds_stack;
  R.stack_cap = t.stack_need;
  R.spill = t.stack_need > t.lds_stack ? t.spill : nullptr;
  R.spill_lanes = L.grid * L.block;
  R.lds_tally = t.dev.nvol <= TRANSPORT_LDS_TALLY_MAX ? 1 : 0;
  R.refill = t.refill;
  R.staged = transport_staged(c, t) && t.stage_cap > 0 ? c->cfg.tr_start : 0;
  if (R.staged) {
    // component-major doubles first, then the 4-byte arrays, then bytes;
    // slices of at most 2^transport_slice histories (the layout follows the
    // slice, inside the buffer sized for stage_cap >= slice)
    const int64_t cap = std::min<int64_t>(t.stage_cap, transport_slice_cap(c, t));
    R.stage.cap = cap;
    R.stage.ru = (double*)t.stage;
    R.stage.hh = (unsigned long long*)(R.stage.ru + 6 * cap);
    R.stage.items = (int2*)(R.stage.hh + cap);
    R.stage.item_cap = (int64_t)std::min(c->cfg.tr_items, t.stage_items) * cap;
    R.stage.mask_words = transport_mask_words(t);
    R.stage.masks = (unsigned long long*)(R.stage.items + R.stage.item_cap);
    R.stage.vol = (int32_t*)(R.stage.masks + (int64_t)R.stage.mask_words * cap);
    R.stage.dc = (uint32_t*)(R.stage.vol + cap);
    R.stage.key = R.stage.dc + cap;
    R.stage.idx = (int32_t*)(R.stage.key + cap);
    R.stage.k2 = (uint32_t*)(R.stage.idx + cap);
    R.stage.v2 = (int32_t*)(R.stage.k2 + cap);
    R.stage.found = R.stage.v2 + cap;
    R.stage.ev = R.stage.found + cap;
    R.stage.hist8 = R.stage.ev + cap;
    R.stage.hist = (uint32_t*)(R.stage.hist8 + 8 * cap);
    R.stage.st = (uint8_t*)(R.stage.hist + radix_sort_hist_words(cap));
    R.stage.over = R.stage.st + cap;
    R.stage.hn = R.stage.over + cap;
    R.stage.live = t.d_live;
    R.stage.claim = (unsigned long long*)(t.d_live + 2);  // 8-byte aligned, inside the 64-byte buffer
    R.rounds = c->cfg.tr_rounds;
    R.events_pt = c->cfg.tr_events;
    R.start_grid = t.start_grid[mixed ? 1 : 0];
  }
  launch_transport(t.dev, R, mixed, L, s);
}

}  // namespace

extern "C" {

int xdg_hip_transport_create(xdg_hip_ctx* c, const xdg_hip_transport_problem* p, int32_t* handle)
{
  return guarded(c, [&]() {
    require(p && handle, "null argument");
    require(p->nvol >= 1 && p->vol_ids && p->vol_trees, "need at least one volume");
    require(p->nsurf_ids >= 1 && p->sense_fwd && p->sense_rev, "need the surface sense tables");
    require(std::isfinite(p->mfp) && p->mfp > 0.0, "mfp must be finite and > 0");
    require(p->p_absorb >= 0.0 && p->p_absorb <= 1.0, "p_absorb must be in [0, 1]");
    require(p->max_events >= 1, "max_events must be >= 1");
    for (int k = 0; k < 3; ++k) require(std::isfinite(p->lo[k]) && std::isfinite(p->hi[k]), "start box not finite");
    int32_t max_id = p->ipc;
    for (int v = 0; v < p->nvol; ++v) {
      require(p->vol_ids[v] >= 0, "volume ids must be >= 0");
      surf_tree(c, p->vol_trees[v]);
      max_id = std::max(max_id, p->vol_ids[v]);
    }
    require(max_id < (1 << 24), "volume ids must be < 2^24");
    for (int i = 0; i < p->nsurf_ids; ++i) {
      max_id = std::max(max_id, std::max(p->sense_fwd[i], p->sense_rev[i]));
    }
    std::vector<int32_t> id2slot((size_t)max_id + 1, -1);
    for (int v = 0; v < p->nvol; ++v) {
      require(id2slot[p->vol_ids[v]] < 0, "duplicate volume id");
      id2slot[p->vol_ids[v]] = v;
    }
    auto* t = new TransportHost();
    try {
      auto up = [&](const int32_t* src, size_t n) {
        int32_t* d = (int32_t*)dev_alloc(n * sizeof(int32_t), t->bufs);
        XDG_HIP_CHECK(hipMemcpy(d, src, n * sizeof(int32_t), hipMemcpyHostToDevice));
        return d;
      };
      t->dev.nvol = p->nvol;
      t->dev.vol_tree = up(p->vol_trees, (size_t)p->nvol);
      t->dev.vol_id = up(p->vol_ids, (size_t)p->nvol);
      t->dev.id2slot = up(id2slot.data(), id2slot.size());
      t->dev.n_ids = (int32_t)id2slot.size();
      t->dev.ipc = p->ipc;
      t->dev.nsurf = p->nsurf_ids;
      t->dev.sense_fwd = up(p->sense_fwd, (size_t)p->nsurf_ids);
      t->dev.sense_rev = up(p->sense_rev, (size_t)p->nsurf_ids);
      t->dev.seed = p->seed;
      t->dev.stream = p->stream;
      for (int k = 0; k < 3; ++k) {
        t->dev.lo[k] = p->lo[k];
        t->dev.hi[k] = p->hi[k];
      }
      t->dev.mfp = p->mfp;
      t->dev.p_absorb = p->p_absorb;
      t->dev.max_events = p->max_events;
      t->dev.surf_bc = nullptr;
      if (p->surf_bc) {
        for (int i = 0; i < p->nsurf_ids; ++i)
          require(p->surf_bc[i] <= XDG_HIP_BC_REFLECTING, "unknown boundary condition code");
        uint8_t* d = (uint8_t*)dev_alloc((size_t)p->nsurf_ids, t->bufs);
        XDG_HIP_CHECK(hipMemcpy(d, p->surf_bc, (size_t)p->nsurf_ids, hipMemcpyHostToDevice));
        t->dev.surf_bc = d;
      }
      t->d_next = (unsigned long long*)dev_alloc(64, t->bufs);
      t->d_live = (uint32_t*)dev_alloc(64, t->bufs);
      t->dev.pre_box = (double*)dev_alloc((size_t)p->nvol * 6 * sizeof(double), t->bufs);
      t->dev.slot_cnt = (uint32_t*)dev_alloc(((size_t)p->nvol + 1) * sizeof(uint32_t), t->bufs);
      t->dev.slot_cur = (uint32_t*)dev_alloc((size_t)p->nvol * sizeof(uint32_t), t->bufs);
      t->host_trees.assign(p->vol_trees, p->vol_trees + p->nvol);
      t->block = c->cfg.block > 256 ? 256 : c->cfg.block;
      transport_refresh(c, *t);
    } catch (...) {
      dev_free_all(t->bufs);
      delete t;
      throw;
    }
    c->transports.push_back(t);
    *handle = (int32_t)c->transports.size();
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_destroy(xdg_hip_ctx* c, int32_t handle)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    XDG_HIP_CHECK(hipDeviceSynchronize());
    if (t.spill) XDG_HIP_CHECK(hipFree(t.spill));
    if (t.stage) XDG_HIP_CHECK(hipFree(t.stage));
    dev_free_all(t.bufs);
    delete &t;
    c->transports[handle - 1] = nullptr;
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_set_config(xdg_hip_ctx* c, int32_t handle, int block, int lds_stack, int chunk, int refill)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    require(block == 0 || (block >= 64 && block <= 256 && block % 64 == 0), "block must be 64, 128, 192 or 256");
    require(lds_stack >= 0 && lds_stack <= 64, "lds_stack must be in [0, 64]");
    require(chunk >= 0 && chunk <= 4096, "chunk must be in [0, 4096]");
    require(refill >= 0 && refill <= 64, "refill must be in [0, 64]");
    if (block) t.block = block;
    if (lds_stack) t.lds_stack = lds_stack;
    if (chunk) t.chunk = chunk;
    if (refill) t.refill = refill;
    transport_refresh(c, t);
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_get_config(xdg_hip_ctx* c, int32_t handle, int* block, int* lds_stack, int* chunk,
                                 int* refill, int* stack_need)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    transport_refresh(c, t);
    if (block) *block = t.block;
    if (lds_stack) *lds_stack = t.lds_stack;
    if (chunk) *chunk = t.chunk;
    if (refill) *refill = t.refill;
    if (stack_need) *stack_need = t.stack_need;
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_run(xdg_hip_ctx* c, int32_t handle, uint64_t first, int64_t n, int precision,
                          uint64_t* tally_fx, int32_t* events, uint8_t* status, uint64_t* hash, void* stream)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    require(n >= 0, "negative n");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(tally_fx != nullptr, "tally_fx is null");
    hipStream_t s = pick(c, stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    XDG_HIP_CHECK(hipStreamIsCapturing(s, &cs));
    const bool mixed = precision == XDG_HIP_MIXED;
    const TransportLaunch L = transport_prepare(c, t, mixed, cs == hipStreamCaptureStatusNone, n);
    transport_launch(c, t, L, mixed, first, n, tally_fx, events, status, hash, s);
    tally_guard(c, tally_fx, t.dev.nvol, XDG_HIP_STATUS_TALLY_RANGE, s);
    XDG_HIP_CHECK(hipGetLastError());
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_capture(xdg_hip_ctx* c, int32_t handle, uint64_t first, int64_t n, int precision,
                              uint64_t* tally_fx, int32_t* events, uint8_t* status, uint64_t* hash, void** graph)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    require(n >= 0, "negative n");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    require(tally_fx != nullptr && graph != nullptr, "null argument");
    const bool mixed = precision == XDG_HIP_MIXED;
    const TransportLaunch L = transport_prepare(c, t, mixed, true, n);
    XDG_HIP_CHECK(hipStreamSynchronize(c->stream));
    hipGraph_t g = nullptr;
    XDG_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      transport_launch(c, t, L, mixed, first, n, tally_fx, events, status, hash, c->stream);
      tally_guard(c, tally_fx, t.dev.nvol, XDG_HIP_STATUS_TALLY_RANGE, c->stream);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    XDG_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    XDG_HIP_CHECK(e);
    auto* gh = new GraphHandle{ex, c, c->generation, handle, t.generation};
    *graph = (void*)gh;
    return XDG_HIP_OK;
  });
}

int xdg_hip_graph_launch(xdg_hip_ctx* c, void* graph, void* stream)
{
  return guarded(c, [&]() {
    require(graph != nullptr, "null graph");
    const GraphHandle* g = (const GraphHandle*)graph;
    require(g->ctx == c, "graph was captured on another context");
    require(g->ctx_generation == c->generation,
            "stale graph: trees were built, released or attached since the capture; capture it again");
    require(g->transport >= 1 && (size_t)g->transport <= c->transports.size() && c->transports[g->transport - 1],
            "stale graph: its transport problem was destroyed");
    require(c->transports[g->transport - 1]->generation == g->transport_generation,
            "stale graph: the transport's stack spill area or start buffer moved since the capture; capture it again");
    XDG_HIP_CHECK(hipGraphLaunch(g->exec, pick(c, stream)));
    return XDG_HIP_OK;
  });
}

int xdg_hip_graph_destroy(xdg_hip_ctx* c, void* graph)
{
  return guarded(c, [&]() {
    if (!graph) return XDG_HIP_OK;
    GraphHandle* g = (GraphHandle*)graph;
    require(g->ctx == c, "graph was captured on another context");
    const hipError_t e = hipGraphExecDestroy(g->exec);
    delete g;
    XDG_HIP_CHECK(e);
    return XDG_HIP_OK;
  });
}

int xdg_hip_transport_autotune(xdg_hip_ctx* c, int32_t handle, uint64_t first, int64_t sample, int precision,
                               double* best_ms)
{
  return guarded(c, [&]() {
    TransportHost& t = transport_of(c, handle);
    require(sample >= 1, "sample must be >= 1");
    require(precision == XDG_HIP_PLAIN || precision == XDG_HIP_MIXED, "unknown precision");
    const bool mixed = precision == XDG_HIP_MIXED;
    transport_refresh(c, t);
    std::vector<DevBuf> tmp;
    uint64_t* d_tally = (uint64_t*)dev_alloc((size_t)t.dev.nvol * 8, tmp);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XDG_HIP_CHECK(hipEventCreate(&e0));
    XDG_HIP_CHECK(hipEventCreate(&e1));
    const int keep[4] = {t.block, t.lds_stack, t.chunk, t.refill};
    double best = -1.0;
    int bb = keep[0], bs = keep[1], bc = keep[2], br = keep[3];
    try {
      std::vector<int> stacks = {std::min(t.stack_need, 8), std::min(t.stack_need, 16), t.stack_need};
      std::sort(stacks.begin(), stacks.end());
      stacks.erase(std::unique(stacks.begin(), stacks.end()), stacks.end());
      for (int block : {64, 128, 256})
        for (int ls : stacks)
          for (int chunk : {1, 4, 16})
            for (int refill : {1, 32, 48}) {
            t.block = block;
            t.lds_stack = ls;
            t.chunk = chunk;
            t.refill = refill;
            TransportLaunch L;
            try {
              L = transport_prepare(c, t, mixed, true, sample);
            } catch (const HipError&) {
              continue;  // shape does not fit a CU
            }
            transport_launch(c, t, L, mixed, first, sample, d_tally, nullptr, nullptr, nullptr, c->stream);
            XDG_HIP_CHECK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < 2; ++r)
              transport_launch(c, t, L, mixed, first, sample, d_tally, nullptr, nullptr, nullptr, c->stream);
            XDG_HIP_CHECK(hipEventRecord(e1, c->stream));
            XDG_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            XDG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (best < 0.0 || ms * 0.5 < best) {
              best = ms * 0.5;
              bb = block;
              bs = ls;
              bc = chunk;
              br = refill;
            }
            }
    } catch (...) {
      t.block = keep[0];
      t.lds_stack = keep[1];
      t.chunk = keep[2];
      t.refill = keep[3];
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      dev_free_all(tmp);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    XDG_HIP_CHECK(hipStreamSynchronize(c->stream));
    dev_free_all(tmp);
    require(best >= 0.0, "no launch shape fits");
    t.block = bb;
    t.lds_stack = bs;
    t.chunk = bc;
    t.refill = br;
    transport_prepare(c, t, mixed, true, sample);
    if (best_ms) *best_ms = best;
    return XDG_HIP_OK;
  });
}

}  // extern "C"
