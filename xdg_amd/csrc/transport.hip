// transport.hip -- batched particle random walk in one persistent launch.
//
// Each lane runs whole particle histories: it claims `chunk` history
// indices at a time from a device counter (one atomic per wave per refill),
// samples the start, finds the volume, then loops events -- ray_fire to the
// volume's boundary, sample the flight, tally, collide or cross -- until the
// particle is absorbed, leaks or hits the event cap, and claims the next.
// Histories of very different lengths therefore share waves without a host
// round trip, and the launch is a fixed sequence of kernels that a caller
// can capture in a HIP graph.
//
// Semantics follow tools/particle_sim.h:22-154 (Particle::initialize,
// surf_dist, sample_collision_distance, advance, collide, cross_surface;
// transport_particles) with: Philox streams in place of drand48, starts
// uniform in a box with isotropic directions, absorption probability per
// collision, leakage into the implicit complement (ipc_graveyard), an event
// cap.  The random-number protocol is the one the host sampler uses, so a
// history is reproducible bit for bit on the CPU.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "surface_traverse.h"

using namespace xdgd;

namespace xdgh {
namespace {

constexpr int HIST = 8;  // exclusion history: last hit faces since the last collision
enum : uint8_t { TR_ALIVE = 0, TR_ABSORBED, TR_LEAKED, TR_LOST, TR_STUCK, TR_CAPPED, TR_ERROR, TR_VACUUM };
enum : uint8_t { BC_TRANSMISSION = 0, BC_VACUUM = 1, BC_REFLECTING = 2 };

__device__ __forceinline__ RayD make_ray_d3(const D3& o, const D3& d)
{
  RayD r;
  r.o = o;
  r.d = d;
  r.rb = cross(d, o);  // plucker.h:75
  r.idx = major_axis(d);
  r.oi = pick3(o.x, o.y, o.z, r.idx);
  r.di = pick3(d.x, d.y, d.z, r.idx);
  return r;
}

// Can the ray meet the tree at all?  Slab test against the tree's vertex
// bounds dilated by the box bump and a relative slack, so a ray that
// reaches any leaf box passes; a ray that fails it misses every triangle.
__device__ __forceinline__ bool may_hit_tree(const SurfTreeDev& T, const D3& o, const D3& d)
{
  double a = 0.0, b = INFTY;
  const double om[3] = {o.x, o.y, o.z}, dm[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double slack = 2.0 * T.bump + 0x1p-30 * (fabs((double)T.box_lo[k]) + fabs((double)T.box_hi[k]) + 1.0);
    const double lo = (double)T.box_lo[k] - slack, hi = (double)T.box_hi[k] + slack;
    if (dm[k] == 0.0) {
      if (om[k] < lo || om[k] > hi) return false;
      continue;
    }
    const double inv = 1.0 / dm[k];
    double t0 = (lo - om[k]) * inv, t1 = (hi - om[k]) * inv;
    if (t0 > t1) {
      const double t = t0;
      t0 = t1;
      t1 = t;
    }
    a = fmax(a, t0);
    b = fmin(b, t1);
  }
  return a <= b;
}

// may_hit_tree's dilated box of every volume slot, written at each launch
// from the live tree table (a refit moves the boxes); a slot that can hold
// no hit (the implicit complement, an unknown or empty tree) gets an empty
// box (lo > hi).
__global__ void k_transport_prefilter(TransportDev P, double* box)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.nvol) return;
  const int32_t t = P.vol_tree[v];
  bool ok = P.vol_id[v] != P.ipc && t >= 1 && t <= P.ntrees;
  const SurfTreeDev* T = ok ? &P.table[t] : nullptr;
  ok = ok && T->n_prims > 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double lo = 1.0, hi = -1.0;
    if (ok) {
      const double slack = 2.0 * T->bump + 0x1p-30 * (fabs((double)T->box_lo[k]) + fabs((double)T->box_hi[k]) + 1.0);
      lo = (double)T->box_lo[k] - slack;
      hi = (double)T->box_hi[k] + slack;
    }
    box[6 * v + k] = lo;
    box[6 * v + 3 + k] = hi;
  }
}

// may_hit_tree on a prefilter box (same arithmetic, reciprocals hoisted)
__device__ __forceinline__ bool may_hit_box(const double* b, const D3& o, const D3& d, const D3& inv)
{
  if (!(b[0] <= b[3])) return false;
  double a = 0.0, e = INFTY;
  const double om[3] = {o.x, o.y, o.z}, dm[3] = {d.x, d.y, d.z}, im[3] = {inv.x, inv.y, inv.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = b[k], hi = b[3 + k];
    if (dm[k] == 0.0) {
      if (om[k] < lo || om[k] > hi) return false;
      continue;
    }
    double t0 = (lo - om[k]) * im[k], t1 = (hi - om[k]) * im[k];
    if (t0 > t1) {
      const double t = t0;
      t0 = t1;
      t1 = t;
    }
    a = fmax(a, t0);
    e = fmin(e, t1);
  }
  return a <= e;
}

__device__ __forceinline__ int wave_min_i(int x)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const int y = __shfl_xor(x, m);
    x = y < x ? y : x;
  }
  return x;
}

template <int MODE, bool MIXED, class STK>
__device__ __forceinline__ void trace_s(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                        int k, STK& S, HitState& h)
{
  if (MIXED && mixed_safe(T, ray))
    traverse4_s<MODE, true, false, false>(T, ray, orientation, ex, k, S, h);
  else
    traverse4_s<MODE, false, false, false>(T, ray, orientation, ex, k, S, h);
}

// Particle::initialize (particle_sim.h:34-40) with a sampled start: a
// uniform point in the source box and an isotropic direction
__device__ __forceinline__ void sample_start(const TransportDev& P, uint64_t item, uint32_t& dc, D3& r, D3& u,
                                             uint8_t& st)
{
  dc = 0;
  double ua, ub, uc, unused;
  uniform2(P.seed, P.stream, item, dc++, ua, ub);
  uniform2(P.seed, P.stream, item, dc++, uc, unused);
  r = d3(P.lo[0] + (P.hi[0] - P.lo[0]) * ua, P.lo[1] + (P.hi[1] - P.lo[1]) * ub, P.lo[2] + (P.hi[2] - P.lo[2]) * uc);
  st = TR_ALIVE;
  if (!iso_dir(P.seed, P.stream, item, dc, u)) st = TR_ERROR;
}

// XDG::find_volume (xdg.cpp:113-128) with the particle's direction: the
// first volume in the caller's order whose point_in_volume is true; none
// means the implicit complement, where the particle has leaked.  Lanes that
// run it together step through the volumes together, so the tree
// descriptors are wave-uniform loads.
template <bool MIXED, class STK>
__device__ __forceinline__ int32_t start_volume(const TransportDev& P, const D3& r, const D3& u, STK& S, uint8_t& st)
{
  int32_t vol = P.ipc;
  if (st != TR_ALIVE) return vol;
  for (int v = 0; v < P.nvol; ++v) {
    if (P.vol_id[v] == P.ipc) continue;
    const int32_t t = P.vol_tree[v];
    if (t < 1 || t > P.ntrees) continue;
    const SurfTreeDev& T = P.table[t];
    if (T.n_prims <= 0 || !may_hit_tree(T, r, u)) continue;
    const RayD ray = make_ray_d3(r, u);
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    trace_s<PIV, MIXED>(T, ray, ORIENT_ANY, nullptr, 0, S, h);
    if (h.surf != ID_NONE && h.inside) {
      vol = P.vol_id[v];
      break;
    }
  }
  if (vol == P.ipc) st = TR_LEAKED;
  return vol;
}

// The start stage on its own, per slice of histories:
//  1. k_transport_sample: one lane per history samples the start and writes
//     it, with a coherence key, to HBM;
//  2. (sorted form) a stable radix sort of the keys orders the histories;
//  3. k_transport_cand, twice: each history lists its candidate volumes,
//     those whose prefilter box its ray reaches (the first pass counts them
//     per volume, the second scatters (history, volume) items grouped by
//     volume);
//  4. k_transport_piv: one lane per item runs point_in_volume; a wave's
//     items mostly share one volume's tree, and a history keeps the lowest
//     volume slot found to contain it (atomicMin);
//  5. k_transport_resolve: the start volume is that slot's volume, or the
//     implicit complement.
// The event kernel then claims histories in the sorted order.  A volume no
// candidate can contain the start (its prefilter box is conservative, as in
// start_volume), so the lowest containing candidate is the first containing
// volume in the caller's order: XDG::find_volume's answer.  A history whose
// items do not fit the item buffer runs start_volume itself in step 5.
// No result depends on the order or on the item buffer: every history is
// computed from its own index.
//
// Key (16 bits): the start's cell on an 8 x 8 x 8 grid over the source box
// (Morton order), the major axis and sign of the direction (6 faces of the
// direction cube) and the other two components over the major one, 4 bins
// each.
__device__ __forceinline__ uint32_t start_key(const TransportDev& P, const D3& r, const D3& u)
{
  const double rm[3] = {r.x, r.y, r.z};
  uint32_t cell[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double w = P.hi[k] - P.lo[k];
    const double f = w > 0.0 ? (rm[k] - P.lo[k]) / w * 8.0 : 0.0;
    cell[k] = f >= 7.0 ? 7u : (f > 0.0 ? (uint32_t)f : 0u);
  }
  uint32_t morton = 0;
#pragma unroll
  for (int b = 2; b >= 0; --b)
    morton = (morton << 3) | (((cell[0] >> b) & 1u) << 2) | (((cell[1] >> b) & 1u) << 1) | ((cell[2] >> b) & 1u);
  const double ax = fabs(u.x), ay = fabs(u.y), az = fabs(u.z);
  const int m = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  const double um = m == 0 ? u.x : (m == 1 ? u.y : u.z);
  const double ua = m == 0 ? u.y : u.x, ub = m == 2 ? u.y : u.z;
  const double inv = um != 0.0 ? 1.0 / fabs(um) : 0.0;
  auto bin4 = [](double x) {  // x in [-1, 1] -> 0..3
    const double f = (x + 1.0) * 2.0;
    return f >= 3.0 ? 3u : (f > 0.0 ? (uint32_t)f : 0u);
  };
  const uint32_t face = (uint32_t)(2 * m + (um < 0.0 ? 1 : 0));
  return (morton << 7) | (face << 4) | (bin4(ua * inv) << 2) | bin4(ub * inv);
}

__global__ __launch_bounds__(256) void k_transport_sample(TransportDev P, TransportRun R)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= R.n) return;
  uint32_t dc;
  D3 r, u;
  uint8_t st;
  sample_start(P, R.first + (uint64_t)i, dc, r, u, st);
  const TransportStage& G = R.stage;
  G.ru[i] = r.x;
  G.ru[G.cap + i] = r.y;
  G.ru[2 * G.cap + i] = r.z;
  G.ru[3 * G.cap + i] = u.x;
  G.ru[4 * G.cap + i] = u.y;
  G.ru[5 * G.cap + i] = u.z;
  G.dc[i] = dc;
  G.st[i] = st;
  G.key[i] = st == TR_ALIVE ? start_key(P, r, u) : 0u;
  G.idx[i] = (int32_t)i;
  G.found[i] = INT32_MAX;
  G.over[i] = 0;
}

__device__ __forceinline__ void load_start(const TransportStage& G, int64_t i, D3& r, D3& u)
{
  r = d3(G.ru[i], G.ru[G.cap + i], G.ru[2 * G.cap + i]);
  u = d3(G.ru[3 * G.cap + i], G.ru[4 * G.cap + i], G.ru[5 * G.cap + i]);
}

// Candidate volumes of each history, in chunks of 64 volume slots: the wave
// stages the chunk's prefilter boxes in LDS and each lane tests its ray
// against all of them (broadcast reads).  Per volume of the chunk that any
// lane's ray reaches, one lane of the wave counts (SCATTER false) or
// reserves (SCATTER true) the wave's items for it.  When the stage keeps
// candidate masks (few enough chunks), the counting pass saves each lane's
// masks and the scatter pass reads them instead of testing again.
//
// With at most CAND_LDS_SLOTS volumes the counts and reservations go
// through LDS first: the workgroup's waves add into per-slot LDS counters,
// one lane per slot adds the workgroup's total to the global counter (the
// scatter pass reserves the workgroup's range there), and the waves then
// take their places from the LDS cursor.  A volume that nearly every ray
// reaches (the box around a lattice) then sees one global atomic per
// workgroup instead of one per wave.
constexpr int CAND_BLOCK = 1024;
constexpr int CAND_LDS_SLOTS = 1024;

// one chunk's candidate mask of this lane: read from the kept masks
// (from_masks), or tested and then kept if the stage keeps masks
__device__ __forceinline__ uint64_t cand_mask(const TransportDev& P, const TransportStage& G, bool from_masks,
                                              bool keep, bool open, int v0, int cn, int64_t j, int lane,
                                              double* wbox, const D3& r, const D3& u, const D3& inv)
{
  uint64_t m = 0;
  if (from_masks) {
    if (open) m = G.masks[(int64_t)(v0 >> 6) * G.cap + j];
    return m;
  }
  __builtin_amdgcn_wave_barrier();
  if (lane < cn) {
#pragma unroll
    for (int k = 0; k < 6; ++k) wbox[6 * lane + k] = P.pre_box[6 * (int64_t)(v0 + lane) + k];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (open)
    for (int q = 0; q < cn; ++q)
      if (may_hit_box(wbox + 6 * q, r, u, inv)) m |= 1ull << q;
  if (keep && open) G.masks[(int64_t)(v0 >> 6) * G.cap + j] = m;
  __builtin_amdgcn_wave_barrier();
  return m;
}

// MODE 0: count into the global counters; 1: count into the LDS counters;
// 2: reserve from the global cursors and write items; 3: reserve from the
// LDS cursors and write items
template <int MODE>
__device__ __forceinline__ void cand_walk(const TransportDev& P, const TransportStage& G, bool from_masks,
                                          bool keep, bool open, int64_t i, int64_t j, int lane, double* wbox,
                                          uint32_t* lcnt, const D3& r, const D3& u, const D3& inv)
{
  for (int v0 = 0; v0 < P.nvol; v0 += 64) {
    const int cn = P.nvol - v0 < 64 ? P.nvol - v0 : 64;
    uint64_t m = cand_mask(P, G, from_masks, keep, open, v0, cn, j, lane, wbox, r, u, inv);
    while (true) {
      const int lq = m ? __ffsll((long long)m) - 1 : 64;
      const int q = wave_min_i(lq);
      if (q >= 64) break;
      const bool mine = lq == q;
      const unsigned long long bal = __ballot(mine);
      const int cnt = __popcll(bal);
      const int leader = __ffsll((long long)bal) - 1;
      if (MODE == 0) {
        if (lane == leader) atomicAdd(P.slot_cnt + v0 + q, (uint32_t)cnt);
      } else if (MODE == 1) {
        if (lane == leader) atomicAdd(lcnt + v0 + q, (uint32_t)cnt);
      } else {
        uint32_t base = 0;
        if (lane == leader)
          base = MODE == 2 ? P.slot_cnt[v0 + q] + atomicAdd(P.slot_cur + v0 + q, (uint32_t)cnt)
                           : atomicAdd(lcnt + v0 + q, (uint32_t)cnt);
        base = __shfl(base, leader);
        if (mine) {
          const uint64_t pos = (uint64_t)base + (uint64_t)__popcll(bal & ((1ull << lane) - 1ull));
          if (pos < (uint64_t)G.item_cap)
            G.items[pos] = make_int2((int32_t)i, v0 + q);
          else
            G.over[i] = 1;
        }
      }
      if (mine) m &= m - 1ull;
    }
  }
}

template <bool SCATTER>
__global__ __launch_bounds__(CAND_BLOCK) void k_transport_cand(TransportDev P, TransportRun R)
{
  __shared__ double sbox[CAND_BLOCK * 6];
  __shared__ uint32_t lcnt[CAND_LDS_SLOTS];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool valid = j < R.n;  // lanes past the end still take part in the wave's and workgroup's steps
  const int lane = threadIdx.x & 63;
  double* wbox = sbox + (size_t)(threadIdx.x >> 6) * (64 * 6);
  const TransportStage& G = R.stage;
  const int64_t i = valid ? (int64_t)G.idx[j] : 0;
  D3 r = d3(0, 0, 0), u = d3(0, 0, 1);
  bool open = false;
  if (valid && G.st[i] == TR_ALIVE) {
    load_start(G, i, r, u);
    open = true;
  }
  const D3 inv = d3(u.x != 0.0 ? 1.0 / u.x : 0.0, u.y != 0.0 ? 1.0 / u.y : 0.0, u.z != 0.0 ? 1.0 / u.z : 0.0);
  const bool any = __ballot(open) != 0;  // wave-uniform
  const bool keep = G.mask_words > 0;
  // block-uniform: LDS aggregation, which the scatter pass does from the
  // kept masks only (its count and write walks must see the same masks)
  const bool blk = P.nvol <= CAND_LDS_SLOTS && (!SCATTER || keep);
  // the scatter pass reads the masks the counting pass kept
  const bool from_masks = SCATTER && keep;
  if (!blk) {
    if (any) cand_walk<SCATTER ? 2 : 0>(P, G, from_masks, keep, open, i, j, lane, wbox, lcnt, r, u, inv);
    return;
  }
  for (int q = threadIdx.x; q < P.nvol; q += blockDim.x) lcnt[q] = 0u;
  __syncthreads();
  if (any) cand_walk<1>(P, G, from_masks, keep, open, i, j, lane, wbox, lcnt, r, u, inv);
  __syncthreads();
  if (!SCATTER) {
    for (int q = threadIdx.x; q < P.nvol; q += blockDim.x)
      if (lcnt[q]) atomicAdd(P.slot_cnt + q, lcnt[q]);
    return;
  }
  for (int q = threadIdx.x; q < P.nvol; q += blockDim.x)
    if (lcnt[q]) lcnt[q] = P.slot_cnt[q] + atomicAdd(P.slot_cur + q, lcnt[q]);
  __syncthreads();
  if (any) cand_walk<3>(P, G, true, keep, open, i, j, lane, wbox, lcnt, r, u, inv);
}

// point_in_volume of each (history, volume slot) item, grid-stride over the
// items (grouped by slot, so a wave's lanes mostly walk one tree).  An item
// whose history already holds a lower containing slot is skipped: it cannot
// lower the minimum, so the skip changes no result.
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport_piv(TransportDev P,
                                                                                              TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  HybridStack S;
  S.lbase = (int32_t*)smem + threadIdx.x;
  S.lstride = blockDim.x;
  S.gbase = nullptr;
  S.gstride = 0;
  S.lcap = R.stack_cap;
  S.cap = R.stack_cap;
  const TransportStage& G = R.stage;
  const uint64_t total = min((uint64_t)P.slot_cnt[P.nvol], (uint64_t)G.item_cap);
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < total;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const int2 it = G.items[k];
    if (G.over[it.x] || __hip_atomic_load(G.found + it.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < it.y)
      continue;
    D3 r, u;
    load_start(G, it.x, r, u);
    const SurfTreeDev& T = P.table[P.vol_tree[it.y]];
    const RayD ray = make_ray_d3(r, u);
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    S.reset();
    trace_s<PIV, MIXED>(T, ray, ORIENT_ANY, nullptr, 0, S, h);
    if (h.surf != ID_NONE && h.inside) atomicMin(G.found + it.x, it.y);
  }
}

// The same items with the traversal as a per-lane state machine (as in the
// persistent ray_fire): a lane whose point_in_volume has ended waits until
// R.refill lanes of its wave (or all remaining ones) have too, then those
// lanes record their results and claim the next items together, while the
// other lanes keep descending.  Items are claimed in list order from the
// counter *G.claim (zeroed before the launch).
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport_piv_pt(TransportDev P,
                                                                                                 TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  LdsStack S((int32_t*)smem + threadIdx.x, blockDim.x, R.stack_cap);
  const int lane = threadIdx.x & 63;
  const int32_t DONE = CHILD_EMPTY;
  const TransportStage& G = R.stage;
  const uint64_t total = min((uint64_t)P.slot_cnt[P.nvol], (uint64_t)G.item_cap);
  bool traced = false;
  int2 it = make_int2(0, 0);
  TreeView T{nullptr, nullptr, nullptr, nullptr};
  RayD ray;
  BoxRayF bf{0, 0, 0, 0, 0, 0};
  BoxRayD bd{0, 0, 0, 0, 0, 0};
  PlaneOfs po{0, 0, 0, 0, 0, 0};
  HitState h{INFTY, INT32_MAX, ID_NONE, false};
  float tf = 0.0f;
  int32_t node = DONE, leaf = DONE;
  while (true) {
    const bool done = node == DONE && leaf == DONE;
    const unsigned long long dmask = __ballot(done);
    const int nact = __popcll(__ballot(1));
    const bool go = done && __popcll(dmask) >= (R.refill < nact ? R.refill : nact);
    if (go && traced) {
      traced = false;
      if (h.surf != ID_NONE && h.inside) atomicMin(G.found + it.x, it.y);
    }
    const unsigned long long cmask = __ballot(go);
    if (cmask) {
      const int leader = __ffsll((long long)cmask) - 1;
      unsigned long long b = 0;
      if (lane == leader) b = atomicAdd(G.claim, (unsigned long long)__popcll(cmask));
      b = __shfl(b, leader);
      if (go) {
        const uint64_t k = b + (unsigned long long)__popcll(cmask & ((1ull << lane) - 1ull));
        if (k >= total) break;
        it = G.items[k];
        // an item whose history holds a lower containing slot cannot lower
        // the minimum: skipped (no result depends on the skip)
        if (G.over[it.x] || __hip_atomic_load(G.found + it.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < it.y)
          continue;
        const SurfTreeDev& TT = P.table[P.vol_tree[it.y]];
        D3 r, u;
        load_start(G, it.x, r, u);
        ray = make_ray_d3(r, u);
        h = HitState{INFTY, INT32_MAX, ID_NONE, false};
        T.nodes4_d = TT.nodes4_d;
        T.nodes4_f = TT.nodes4_f;
        T.tris = TT.tris;
        T.normals = TT.normals;
        traced = true;
        S.reset();
        if (MIXED && !mixed_safe(TT, ray)) {
          trace_s<PIV, false>(TT, ray, ORIENT_ANY, nullptr, 0, S, h);
          continue;
        }
        if (MIXED) {
          const float ox = (float)ray.o.x, oy = (float)ray.o.y, oz = (float)ray.o.z;
          bf.ix = safe_inv_f((float)ray.d.x);
          bf.iy = safe_inv_f((float)ray.d.y);
          bf.iz = safe_inv_f((float)ray.d.z);
          bf.nx = -ox * bf.ix;
          bf.ny = -oy * bf.iy;
          bf.nz = -oz * bf.iz;
        } else {
          bd.ix = safe_inv_d(ray.d.x);
          bd.iy = safe_inv_d(ray.d.y);
          bd.iz = safe_inv_d(ray.d.z);
          bd.nx = -ray.o.x * bd.ix;
          bd.ny = -ray.o.y * bd.iy;
          bd.nz = -ray.o.z * bd.iz;
        }
        po = MIXED ? plane_ofs_f(bf) : plane_ofs_d(bd);
        tf = f32_up(h.t);
        node = TT.n_prims > 0 ? TT.root4 : DONE;
      }
    }
    while (node >= 0) {
      float te[4];
      int32_t ch[4];
      node4_test<MIXED>(T, node, bf, bd, po, tf, h.t, te, ch);
      cswap(te[0], ch[0], te[1], ch[1]);
      cswap(te[2], ch[2], te[3], ch[3]);
      cswap(te[0], ch[0], te[2], ch[2]);
      cswap(te[1], ch[1], te[3], ch[3]);
      cswap(te[1], ch[1], te[2], ch[2]);
      if (ch[0] != CHILD_EMPTY) {
        if (ch[3] != CHILD_EMPTY) S.push(ch[3]);
        if (ch[2] != CHILD_EMPTY) S.push(ch[2]);
        if (ch[1] != CHILD_EMPTY) S.push(ch[1]);
        node = ch[0];
      } else {
        node = S.empty() ? DONE : S.pop();
      }
      if (node < 0 && node != DONE && leaf == DONE) {
        leaf = node;
        node = S.empty() ? DONE : S.pop();
      }
      if (__all(leaf != DONE)) break;
    }
    if (leaf == DONE && node != DONE) {
      leaf = node;
      node = S.empty() ? DONE : S.pop();
    }
    if (leaf != DONE) {
      const int32_t first = leaf_first(leaf);
      const int32_t cnt = leaf_count(leaf);
      leaf = DONE;
      for (int32_t j = 0; j < cnt; ++j) tri_closest_hit<PIV>(T, first + j, ray, ORIENT_ANY, nullptr, 0, h);
      tf = f32_up(h.t);
    }
  }
}

// Start volume of each history from its lowest containing slot; histories
// whose items overflowed the buffer run start_volume.
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport_resolve(TransportDev P,
                                                                                                  TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= R.n) return;
  const TransportStage& G = R.stage;
  uint8_t st = G.st[i];
  int32_t vol = P.ipc;
  if (G.over[i]) {
    HybridStack S;
    S.lbase = (int32_t*)smem + threadIdx.x;
    S.lstride = blockDim.x;
    S.gbase = nullptr;
    S.gstride = 0;
    S.lcap = R.stack_cap;
    S.cap = R.stack_cap;
    S.reset();
    D3 r, u;
    load_start(G, i, r, u);
    vol = start_volume<MIXED>(P, r, u, S, st);
  } else if (st == TR_ALIVE) {
    const int32_t f = G.found[i];
    vol = f < P.nvol ? P.vol_id[f] : P.ipc;
    if (vol == P.ipc) st = TR_LEAKED;
  }
  G.vol[i] = vol;
  G.st[i] = st;
  G.ev[i] = 0;
  G.hh[i] = HASH_SEED;
  G.hn[i] = 0;
  if (st != TR_ALIVE) {  // the history ends at its start
    if (R.events) R.events[i] = 0;
    if (R.status) R.status[i] = st;
    if (R.hash) R.hash[i] = HASH_SEED;
  }
}

// One history's state between events.
struct Hist {
  D3 r, u;
  int32_t vol, ev;
  uint32_t dc;
  int hn;
  uint8_t st;
  unsigned long long hh;
  uint64_t item;
};

// The rest of an event once its ray_fire hit h in volume slot `slot` is
// known: sample the flight, tally, then collide or cross
// (particle_sim.h:56-130).  normals: the traced tree's stored normals.
__device__ __forceinline__ void event_finish(const TransportDev& P, const TransportRun& R,
                                             unsigned long long* ltally, int32_t* hist, Hist& H, const HitState& h,
                                             const double* normals, int32_t slot)
{
  if (h.surf == ID_NONE) {
    H.st = TR_LOST;
    return;
  }
  if (h.t == 0.0) {
    H.st = TR_STUCK;
    return;
  }
  // the hit face joins the exclusion history (ray_tracer.cpp:324)
  if (H.hn == HIST) {
    for (int j = 1; j < HIST; ++j) hist[j - 1] = hist[j];
    --H.hn;
  }
  hist[H.hn++] = h.face;
  // sample_collision_distance (particle_sim.h:56-58)
  double ua, ub;
  uniform2(P.seed, P.stream, H.item, H.dc++, ua, ub);
  const double dcol = -log_series(1.0 - ua) * P.mfp;
  // advance (particle_sim.h:70-83)
  const double s = dcol < h.t ? dcol : h.t;
  H.r = d3(H.r.x + s * H.u.x, H.r.y + s * H.u.y, H.r.z + s * H.u.z);
  const unsigned long long q = (unsigned long long)__double2ll_rn(s * 4294967296.0);  // s * 2^32
  if (R.lds_tally)
    atomicAdd(ltally + slot, q);
  else
    atomicAdd(R.tally_fx + slot, q);
  H.hh = mix64(H.hh ^ (unsigned long long)(uint32_t)H.vol);
  H.hh = mix64(H.hh ^ (unsigned long long)__double_as_longlong(s));
  ++H.ev;
  if (dcol < h.t) {
    // collide (particle_sim.h:60-68), absorption first
    if (ub < P.p_absorb) {
      H.st = TR_ABSORBED;
    } else {
      if (!iso_dir(P.seed, P.stream, H.item, H.dc, H.u)) H.st = TR_ERROR;
      H.hn = 0;
    }
  } else {
    // cross_surface (particle_sim.h:85-130): the hit surface's boundary
    // condition (MeshManager::get_surface_property, default transmission)
    const int32_t sid = h.surf;
    const bool known = sid >= 0 && sid < P.nsurf;
    const uint8_t bc = (P.surf_bc && known) ? P.surf_bc[sid] : (uint8_t)BC_TRANSMISSION;
    if (bc == BC_VACUUM) {
      H.st = TR_VACUUM;
    } else if (bc == BC_REFLECTING) {
      // particle_sim.h:90-104; XDG::surface_normal (xdg.cpp:272-290) is
      // face_normal of the last history entry, the face just hit, whose
      // unflipped normal the tree stores in leaf order
      const double* pn = normals + 3 * (int64_t)h.leaf;
      const D3 nrm = d3(pn[0], pn[1], pn[2]);
      const double proj = nrm.x * H.u.x + nrm.y * H.u.y + nrm.z * H.u.z;
      const double mag = sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
      const double sc = 2.0 * proj / mag;
      H.u = normalize(d3(H.u.x - sc * nrm.x, H.u.y - sc * nrm.y, H.u.z - sc * nrm.z));
      hist[0] = h.face;  // history_ = {history_.back()}
      H.hn = 1;
    } else {
      // MeshManager::next_volume (mesh_manager_interface.cpp:202-214)
      int32_t nv = ID_NONE;
      if (known) {
        if (P.sense_fwd[sid] == H.vol)
          nv = P.sense_rev[sid];
        else if (P.sense_rev[sid] == H.vol)
          nv = P.sense_fwd[sid];
      }
      if (nv == ID_NONE) {
        H.st = TR_ERROR;
      } else {
        H.vol = nv;
        if (H.vol == P.ipc) H.st = TR_LEAKED;
      }
    }
  }
  if (H.st == TR_ALIVE && H.ev >= P.max_events) H.st = TR_CAPPED;
}

// volume slot and surface tree of the history's current volume (-1: none)
__device__ __forceinline__ int32_t event_slot(const TransportDev& P, const Hist& H, int32_t& tree)
{
  const int32_t slot = (H.vol >= 0 && H.vol < P.n_ids) ? P.id2slot[H.vol] : -1;
  tree = slot >= 0 ? P.vol_tree[slot] : -1;
  return (tree < 1 || tree > P.ntrees) ? -1 : slot;
}

// One event of a live history (particle_sim.h:136-154): ray_fire to the
// volume's boundary with the exclusion history (Particle::surf_dist,
// particle_sim.h:41-54), then event_finish.  hist is the lane's exclusion
// list (HIST entries).
template <bool MIXED, class STK>
__device__ __forceinline__ void event_step(const TransportDev& P, const TransportRun& R,
                                           unsigned long long* ltally, STK& S, int32_t* hist, Hist& H)
{
  int32_t t;
  const int32_t slot = event_slot(P, H, t);
  if (slot < 0) {
    H.st = TR_ERROR;
    return;
  }
  const SurfTreeDev& T = P.table[t];
  const RayD ray = make_ray_d3(H.r, H.u);
  HitState h{INFTY, INT32_MAX, ID_NONE, false};
  S.reset();
  trace_s<RAYFIRE, MIXED>(T, ray, ORIENT_EXITING, hist, H.hn, S, h);
  event_finish(P, R, ltally, hist, H, h, T.normals, slot);
}

__device__ __forceinline__ void record_end(const TransportRun& R, int64_t i, const Hist& H)
{
  if (R.events) R.events[i] = H.ev;
  if (R.status) R.status[i] = H.st;
  if (R.hash) R.hash[i] = H.hh;
}

// saved state of the staged form (start stage or event rounds)
__device__ __forceinline__ void load_hist(const TransportRun& R, int64_t i, Hist& H, int32_t* hist)
{
  const TransportStage& G = R.stage;
  load_start(G, i, H.r, H.u);
  H.vol = G.vol[i];
  H.dc = G.dc[i];
  H.st = G.st[i];
  H.ev = G.ev[i];
  H.hh = G.hh[i];
  H.hn = G.hn[i];
  for (int k = 0; k < H.hn; ++k) hist[k] = G.hist8[k * G.cap + i];
  H.item = R.first + (uint64_t)i;
}

__device__ __forceinline__ void store_hist(const TransportRun& R, int64_t i, const Hist& H, const int32_t* hist)
{
  const TransportStage& G = R.stage;
  G.ru[i] = H.r.x;
  G.ru[G.cap + i] = H.r.y;
  G.ru[2 * G.cap + i] = H.r.z;
  G.ru[3 * G.cap + i] = H.u.x;
  G.ru[4 * G.cap + i] = H.u.y;
  G.ru[5 * G.cap + i] = H.u.z;
  G.vol[i] = H.vol;
  G.dc[i] = H.dc;
  G.st[i] = H.st;
  G.ev[i] = H.ev;
  G.hh[i] = H.hh;
  G.hn[i] = (uint8_t)H.hn;
  for (int k = 0; k < H.hn; ++k) G.hist8[k * G.cap + i] = hist[k];
}

// LDS of the event kernels: per-workgroup tallies, then the stack share,
// then the exclusion histories
struct EventLds {
  unsigned long long* ltally;
  int32_t* hist;
  HybridStack S;
};

__device__ __forceinline__ EventLds event_lds(unsigned char* smem, const TransportDev& P, const TransportRun& R)
{
  EventLds E;
  E.ltally = (unsigned long long*)smem;
  const size_t tally_bytes = R.lds_tally ? (((size_t)P.nvol * 8 + 15) & ~(size_t)15) : 0;
  int32_t* lstack = (int32_t*)(smem + tally_bytes);
  E.hist = lstack + (size_t)R.lds_stack * blockDim.x + threadIdx.x * HIST;
  if (R.lds_tally) {
    for (int i = threadIdx.x; i < P.nvol; i += blockDim.x) E.ltally[i] = 0ull;
    __syncthreads();
  }
  const int64_t gl = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  E.S.lbase = lstack + threadIdx.x;
  E.S.lstride = blockDim.x;
  E.S.gbase = R.spill ? R.spill + gl : nullptr;
  E.S.gstride = R.spill_lanes;
  E.S.lcap = R.lds_stack;
  E.S.cap = R.spill ? R.stack_cap : R.lds_stack;
  E.S.reset();
  return E;
}

__device__ __forceinline__ void flush_tally(const TransportDev& P, const TransportRun& R, const EventLds& E)
{
  if (R.lds_tally) {
    __syncthreads();
    for (int i = threadIdx.x; i < P.nvol; i += blockDim.x)
      if (E.ltally[i]) atomicAdd(R.tally_fx + i, E.ltally[i]);
  }
}

// Whole histories keep more state live than one query: workgroups of at
// most 256 lanes let the kernel use more than 128 VGPRs instead of spilling
// to scratch; three waves per SIMD (168 VGPRs) is the target.
// STAGED: the histories are the staged form's list R.stage.idx[0 .. n),
// n = *R.stage.live after event rounds and R.n otherwise, resumed from
// their saved state (one that has ended already wrote its record);
// otherwise histories [0, R.n) are sampled and started here.
template <bool MIXED, bool STAGED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport(TransportDev P, TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  EventLds E = event_lds(smem, P, R);
  const int lane = threadIdx.x & 63;
  const uint64_t n = (STAGED && R.rounds > 0) ? (uint64_t)*R.stage.live : (uint64_t)R.n;
  uint64_t cur = 0, end = 0;  // claimed range of list positions
  bool alive = false, have = false;
  int64_t hix = 0;
  Hist H;
  H.st = TR_ALIVE;
  while (true) {
    // a lane whose history ended waits until at least R.refill lanes of the
    // wave (or all remaining ones) are idle: starting a history (sampling,
    // find_volume over every volume) is long, and the wave runs it for all
    // lanes that start together
    const bool idle = !alive;
    if (idle && have) {
      record_end(R, hix, H);
      have = false;
    }
    const int nact = __popcll(__ballot(1));
    const int nidle = __popcll(__ballot(idle));
    const bool go = nidle >= (R.refill < nact ? R.refill : nact);
    if (idle) {
      if (!go) continue;
      const bool need = cur >= end;
      const unsigned long long mask = __ballot(need);
      if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long b = 0;
        if (lane == leader) b = atomicAdd(R.next, (unsigned long long)__popcll(mask) * (unsigned long long)R.chunk);
        b = __shfl(b, leader);
        if (need) {
          cur = b + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)) * (unsigned long long)R.chunk;
          end = cur + (uint64_t)R.chunk;
          if (end > n) end = n;
        }
      }
      if (cur >= n) break;
      have = true;
      if (STAGED) {
        hix = (int64_t)R.stage.idx[cur];
        load_hist(R, hix, H, E.hist);
        have = H.st == TR_ALIVE;
      } else {
        hix = (int64_t)cur;
        H.item = R.first + (uint64_t)hix;
        H.ev = 0;
        H.hh = HASH_SEED;
        H.hn = 0;
        sample_start(P, H.item, H.dc, H.r, H.u, H.st);
        E.S.reset();
        H.vol = start_volume<MIXED>(P, H.r, H.u, E.S, H.st);
      }
      ++cur;
      alive = H.st == TR_ALIVE;
      continue;
    }
    event_step<MIXED>(P, R, E.ltally, E.S, E.hist, H);
    alive = H.st == TR_ALIVE;
  }
  flush_tally(P, R, E);
}

// The staged form's event kernel with the traversal as a per-lane state
// machine, as in the persistent ray_fire: a lane whose ray_fire has ended
// waits until R.refill lanes of its wave (or all remaining ones) have too,
// then those lanes finish their events together (flight, tally, collide or
// cross), claim new histories where theirs ended, and start their next
// rays; meanwhile the other lanes keep descending.  A wave therefore never
// waits for its longest traversal before its other lanes go on.  The
// histories are R.stage.idx[0 .. n), resumed from their saved state.
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport_pt(TransportDev P,
                                                                                             TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  EventLds E = event_lds(smem, P, R);
  const int lane = threadIdx.x & 63;
  const int32_t DONE = CHILD_EMPTY;
  const uint64_t n = R.rounds > 0 ? (uint64_t)*R.stage.live : (uint64_t)R.n;
  bool have = false;    // the lane holds a live history
  bool traced = false;  // its current event's ray is traced (or being traced)
  int64_t hix = 0;
  Hist H;
  H.st = TR_ALIVE;
  TreeView T{nullptr, nullptr, nullptr, nullptr};
  RayD ray;
  BoxRayF bf{0, 0, 0, 0, 0, 0};
  BoxRayD bd{0, 0, 0, 0, 0, 0};
  PlaneOfs po{0, 0, 0, 0, 0, 0};
  HitState h{INFTY, INT32_MAX, ID_NONE, false};
  float tf = 0.0f;
  int32_t node = DONE, leaf = DONE, slot = -1;
  while (true) {
    const bool done = node == DONE && leaf == DONE;
    const unsigned long long dmask = __ballot(done);
    const int nact = __popcll(__ballot(1));
    const bool go = done && __popcll(dmask) >= (R.refill < nact ? R.refill : nact);
    if (go && traced) {
      traced = false;
      event_finish(P, R, E.ltally, E.hist, H, h, T.normals, slot);
      if (H.st != TR_ALIVE) {
        record_end(R, hix, H);
        have = false;
      }
    }
    // lanes without a history claim the next ones of the list
    const bool need = go && !have;
    const unsigned long long cmask = __ballot(need);
    if (cmask) {
      const int leader = __ffsll((long long)cmask) - 1;
      unsigned long long b = 0;
      if (lane == leader) b = atomicAdd(R.next, (unsigned long long)__popcll(cmask));
      b = __shfl(b, leader);
      if (need) {
        const uint64_t pos = b + (unsigned long long)__popcll(cmask & ((1ull << lane) - 1ull));
        if (pos >= n) break;
        hix = (int64_t)R.stage.idx[pos];
        load_hist(R, hix, H, E.hist);
        have = H.st == TR_ALIVE;  // one that ended at its start has its record already
      }
    }
    // lanes with a live history and no ray start the next event's ray
    if (go && have) {
      int32_t t;
      slot = event_slot(P, H, t);
      if (slot < 0) {
        H.st = TR_ERROR;
        record_end(R, hix, H);
        have = false;
        continue;
      }
      const SurfTreeDev& TT = P.table[t];
      ray = make_ray_d3(H.r, H.u);
      h = HitState{INFTY, INT32_MAX, ID_NONE, false};
      T.nodes4_d = TT.nodes4_d;
      T.nodes4_f = TT.nodes4_f;
      T.tris = TT.tris;
      T.normals = TT.normals;
      traced = true;
      E.S.reset();
      if (MIXED && !mixed_safe(TT, ray)) {
        // an origin beyond the tree's FP32 limit: FP64 boxes, traced at once
        trace_s<RAYFIRE, false>(TT, ray, ORIENT_EXITING, E.hist, H.hn, E.S, h);
        continue;
      }
      if (MIXED) {
        const float ox = (float)ray.o.x, oy = (float)ray.o.y, oz = (float)ray.o.z;
        bf.ix = safe_inv_f((float)ray.d.x);
        bf.iy = safe_inv_f((float)ray.d.y);
        bf.iz = safe_inv_f((float)ray.d.z);
        bf.nx = -ox * bf.ix;
        bf.ny = -oy * bf.iy;
        bf.nz = -oz * bf.iz;
      } else {
        bd.ix = safe_inv_d(ray.d.x);
        bd.iy = safe_inv_d(ray.d.y);
        bd.iz = safe_inv_d(ray.d.z);
        bd.nx = -ray.o.x * bd.ix;
        bd.ny = -ray.o.y * bd.iy;
        bd.nz = -ray.o.z * bd.iz;
      }
      po = MIXED ? plane_ofs_f(bf) : plane_ofs_d(bd);
      tf = f32_up(h.t);
      node = TT.n_prims > 0 ? TT.root4 : DONE;
    }
    while (node >= 0) {
      float te[4];
      int32_t ch[4];
      node4_test<MIXED>(T, node, bf, bd, po, tf, h.t, te, ch);
      cswap(te[0], ch[0], te[1], ch[1]);
      cswap(te[2], ch[2], te[3], ch[3]);
      cswap(te[0], ch[0], te[2], ch[2]);
      cswap(te[1], ch[1], te[3], ch[3]);
      cswap(te[1], ch[1], te[2], ch[2]);
      if (ch[0] != CHILD_EMPTY) {
        if (ch[3] != CHILD_EMPTY) E.S.push(ch[3]);
        if (ch[2] != CHILD_EMPTY) E.S.push(ch[2]);
        if (ch[1] != CHILD_EMPTY) E.S.push(ch[1]);
        node = ch[0];
      } else {
        node = E.S.empty() ? DONE : E.S.pop();
      }
      if (node < 0 && node != DONE && leaf == DONE) {
        leaf = node;
        node = E.S.empty() ? DONE : E.S.pop();
      }
      if (__all(leaf != DONE)) break;
    }
    if (leaf == DONE && node != DONE) {
      leaf = node;
      node = E.S.empty() ? DONE : E.S.pop();
    }
    if (leaf != DONE) {
      const int32_t first = leaf_first(leaf);
      const int32_t cnt = leaf_count(leaf);
      leaf = DONE;
      for (int32_t j = 0; j < cnt; ++j)
        tri_closest_hit<RAYFIRE>(T, first + j, ray, ORIENT_EXITING, E.hist, H.hn, h);
      tf = f32_up(h.t);
    }
  }
  flush_tally(P, R, E);
}

// One event round of the staged form: each live history of the sorted list
// R.stage.idx[0 .. *R.stage.live) runs one event (grid-stride, so a wave's
// lanes take neighbours in the sorted order: the same volume, nearby starts
// and directions).  The state goes back to HBM; a history that ends writes
// its record.
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport_round(TransportDev P,
                                                                                                TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  EventLds E = event_lds(smem, P, R);
  const uint64_t n = *R.stage.live;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t i = R.stage.idx[j];
    Hist H;
    load_hist(R, i, H, E.hist);
    event_step<MIXED>(P, R, E.ltally, E.S, E.hist, H);
    store_hist(R, i, H, E.hist);
    if (H.st != TR_ALIVE) record_end(R, i, H);
  }
  flush_tally(P, R, E);
}

// Order of the next round: live histories first, by volume slot, then by
// cell and direction (start_key's bits); ended histories get the largest
// key (dead_key), so after the sort the live ones lead.
__device__ __forceinline__ uint32_t dead_key(int key_bits)
{
  return key_bits >= 32 ? 0xffffffffu : ((1u << key_bits) - 1u);
}

__global__ __launch_bounds__(256) void k_transport_keys(TransportDev P, TransportRun R, int slot_bits, int key_bits)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const TransportStage& G = R.stage;
  if (i < R.n) {
    uint32_t key = dead_key(key_bits);
    if (G.st[i] == TR_ALIVE) {
      const int32_t v = G.vol[i];
      const int32_t slot = (v >= 0 && v < P.n_ids) ? P.id2slot[v] : -1;
      D3 r, u;
      load_start(G, i, r, u);
      // up to 9 spatial bits: the cell on a 4 x 4 x 4 grid (the top 6
      // Morton bits of start_key) and the direction's face
      const uint32_t k16 = start_key(P, r, u);
      const uint32_t k9 = ((k16 >> 10) << 3) | ((k16 >> 4) & 7u);
      const int sp = key_bits - slot_bits;
      key = ((uint32_t)(slot < 0 ? 0 : slot) << sp) | (k9 >> (9 - sp));
    }
    G.key[i] = key;
    G.idx[i] = (int32_t)i;
  }
}

// live count = position of the first dead key in the sorted keys
__global__ void k_transport_live(TransportRun R, int key_bits)
{
  if (threadIdx.x != 0) return;
  const uint32_t dead = dead_key(key_bits);
  int64_t lo = 0, hi = R.n;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (R.stage.key[mid] < dead)
      lo = mid + 1;
    else
      hi = mid;
  }
  *R.stage.live = (uint32_t)lo;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

size_t transport_lds(int block, int lds_stack, int nvol)
{
  const size_t tally = nvol <= TRANSPORT_LDS_TALLY_MAX ? (((size_t)nvol * 8 + 15) & ~(size_t)15) : 0;
  return tally + (size_t)block * 4 * (size_t)(lds_stack + HIST);
}

}  // namespace

TransportLaunch transport_shape(bool mixed, int block, int lds_stack, int nvol)
{
  TransportLaunch L;
  L.block = block;
  L.lds_bytes = transport_lds(block, lds_stack, nvol);
  int dev = 0, cu = 0, nb = 0;
  XDG_HIP_CHECK(hipGetDevice(&dev));
  XDG_HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  // the staged event kernel has the same resources as the plain one
  if (mixed)
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport<true, false>, block, L.lds_bytes));
  else
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport<false, false>, block, L.lds_bytes));
  if (nb <= 0) throw HipError{"transport launch shape does not fit one CU (block " + std::to_string(block) + ")"};
  L.grid = (int64_t)nb * cu;
  int np = 0;
  if (mixed)
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&np, k_transport_pt<true>, block, L.lds_bytes));
  else
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&np, k_transport_pt<false>, block, L.lds_bytes));
  L.grid_pt = (int64_t)std::max(np, 1) * cu;
  L.grid = std::max(L.grid, L.grid_pt);  // the stack spill area covers either grid
  return L;
}

size_t transport_start_lds(int block, int stack_cap) { return (size_t)block * 4 * (size_t)stack_cap; }

int transport_start_block(int stack_cap)
{
  for (int b = 256; b >= 64; b -= 64)
    if (transport_start_lds(b, stack_cap) <= TRANSPORT_START_LDS_MAX) return b;
  return 0;
}

int64_t transport_start_grid(bool mixed, int stack_cap)
{
  const int sb = transport_start_block(stack_cap);
  if (sb == 0) return 0;
  int dev = 0, cu = 0, nb = 0;
  XDG_HIP_CHECK(hipGetDevice(&dev));
  XDG_HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t sh = transport_start_lds(sb, stack_cap);
  int nb2 = 0;
  if (mixed) {
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_piv<true>, sb, sh));
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, k_transport_piv_pt<true>, sb, sh));
  } else {
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_piv<false>, sb, sh));
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, k_transport_piv_pt<false>, sb, sh));
  }
  return (int64_t)std::max(std::min(nb, nb2), 1) * cu;
}

void launch_transport(const TransportDev& P, const TransportRun& R, bool mixed, const TransportLaunch& L,
                      hipStream_t s)
{
  if (P.nvol > 0)
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((P.nvol + 255) / 256)), dim3(256), 0, s, R.tally_fx,
                       (int64_t)P.nvol, 0ull);
  if (R.n <= 0) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, R.next, (int64_t)1, 0ull);
    return;
  }
  // the staged form runs the histories in slices of at most stage.cap: the
  // start kernel for the slice, then the event kernel over it
  const int64_t slice = R.staged ? R.stage.cap : R.n;
  const int sb = R.staged ? transport_start_block(R.stack_cap) : 0;
  if (R.staged && (slice <= 0 || sb == 0)) throw HipError{"staged transport: no start buffer or stack too deep"};
  for (int64_t off = 0; off < R.n; off += slice) {
    TransportRun Q = R;
    Q.first = R.first + (uint64_t)off;
    Q.n = std::min(slice, R.n - off);
    if (R.events) Q.events = R.events + off;
    if (R.status) Q.status = R.status + off;
    if (R.hash) Q.hash = R.hash + off;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, Q.next, (int64_t)1, 0ull);
    if (R.staged) {
      if (off == 0)
        hipLaunchKernelGGL(k_transport_prefilter, dim3((unsigned)((P.nvol + 255) / 256)), dim3(256), 0, s, P,
                           P.pre_box);
      const dim3 gn((unsigned)((Q.n + 255) / 256));
      hipLaunchKernelGGL(k_transport_sample, gn, dim3(256), 0, s, P, Q);
      if (R.staged == 2)
        radix_sort_pairs(R.stage.key, R.stage.idx, R.stage.k2, R.stage.v2, Q.n, 16, R.stage.hist, s);
      const dim3 gv((unsigned)((P.nvol + 256) / 256));
      hipLaunchKernelGGL(k_fill_u32, gv, dim3(256), 0, s, P.slot_cnt, (int64_t)P.nvol + 1, 0u);
      hipLaunchKernelGGL(k_fill_u32, gv, dim3(256), 0, s, P.slot_cur, (int64_t)P.nvol, 0u);
      const dim3 gc((unsigned)((Q.n + CAND_BLOCK - 1) / CAND_BLOCK));
      hipLaunchKernelGGL(k_transport_cand<false>, gc, dim3(CAND_BLOCK), 0, s, P, Q);
      launch_scan_excl(P.slot_cnt, P.nvol, P.slot_cnt + P.nvol, s);
      hipLaunchKernelGGL(k_transport_cand<true>, gc, dim3(CAND_BLOCK), 0, s, P, Q);
      const size_t sh = transport_start_lds(sb, R.stack_cap);
      const unsigned gp = (unsigned)std::max<int64_t>(1, R.start_grid);
      if (R.events_pt) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, R.stage.claim, (int64_t)1, 0ull);
        if (mixed)
          hipLaunchKernelGGL(k_transport_piv_pt<true>, dim3(gp), dim3(sb), sh, s, P, Q);
        else
          hipLaunchKernelGGL(k_transport_piv_pt<false>, dim3(gp), dim3(sb), sh, s, P, Q);
      } else if (mixed) {
        hipLaunchKernelGGL(k_transport_piv<true>, dim3(gp), dim3(sb), sh, s, P, Q);
      } else {
        hipLaunchKernelGGL(k_transport_piv<false>, dim3(gp), dim3(sb), sh, s, P, Q);
      }
      const dim3 gs((unsigned)((Q.n + sb - 1) / sb));
      if (mixed)
        hipLaunchKernelGGL(k_transport_resolve<true>, gs, dim3(sb), sh, s, P, Q);
      else
        hipLaunchKernelGGL(k_transport_resolve<false>, gs, dim3(sb), sh, s, P, Q);
    }
    const int64_t need = (Q.n + (int64_t)L.block * Q.chunk - 1) / ((int64_t)L.block * Q.chunk);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min(L.grid, need));
    if (R.staged) {
      // R.rounds sorted one-event rounds, then the event kernel takes the
      // live histories, each until it ends
      int slot_bits = 1;
      while (((int64_t)1 << slot_bits) <= (int64_t)P.nvol) ++slot_bits;
      const int key_bits = std::min(32, slot_bits + 9);
      const dim3 gk((unsigned)((Q.n + 255) / 256));
      const unsigned gr = (unsigned)std::max<int64_t>(1, L.grid);
      for (int rd = 0; R.rounds > 0; ++rd) {
        hipLaunchKernelGGL(k_transport_keys, gk, dim3(256), 0, s, P, Q, slot_bits, key_bits);
        radix_sort_pairs(R.stage.key, R.stage.idx, R.stage.k2, R.stage.v2, Q.n, key_bits, R.stage.hist, s);
        hipLaunchKernelGGL(k_transport_live, dim3(1), dim3(64), 0, s, Q, key_bits);
        if (rd >= R.rounds) break;
        if (mixed)
          hipLaunchKernelGGL(k_transport_round<true>, dim3(gr), dim3(L.block), L.lds_bytes, s, P, Q);
        else
          hipLaunchKernelGGL(k_transport_round<false>, dim3(gr), dim3(L.block), L.lds_bytes, s, P, Q);
      }
      if (R.events_pt) {
        const unsigned gp = (unsigned)std::max<int64_t>(1, std::min(L.grid_pt, need));
        if (mixed)
          hipLaunchKernelGGL(k_transport_pt<true>, dim3(gp), dim3(L.block), L.lds_bytes, s, P, Q);
        else
          hipLaunchKernelGGL(k_transport_pt<false>, dim3(gp), dim3(L.block), L.lds_bytes, s, P, Q);
      } else if (mixed) {
        hipLaunchKernelGGL((k_transport<true, true>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
      } else {
        hipLaunchKernelGGL((k_transport<false, true>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
      }
    } else {
      if (mixed)
        hipLaunchKernelGGL((k_transport<true, false>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
      else
        hipLaunchKernelGGL((k_transport<false, false>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
    }
  }
}

}  // namespace xdgh
