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
// reserves (SCATTER true) the wave's items for it.
template <bool SCATTER>
__global__ __launch_bounds__(256) void k_transport_cand(TransportDev P, TransportRun R)
{
  __shared__ double sbox[256 * 6];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool valid = j < R.n;  // lanes past the end still take part in the wave's steps
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
  if (__ballot(open) == 0) return;  // wave-uniform
  for (int v0 = 0; v0 < P.nvol; v0 += 64) {
    const int cn = P.nvol - v0 < 64 ? P.nvol - v0 : 64;
    __builtin_amdgcn_wave_barrier();
    if (lane < cn) {
#pragma unroll
      for (int k = 0; k < 6; ++k) wbox[6 * lane + k] = P.pre_box[6 * (int64_t)(v0 + lane) + k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint64_t m = 0;
    if (open)
      for (int q = 0; q < cn; ++q)
        if (may_hit_box(wbox + 6 * q, r, u, inv)) m |= 1ull << q;
    while (true) {
      const int lq = m ? __ffsll((long long)m) - 1 : 64;
      const int q = wave_min_i(lq);
      if (q >= 64) break;
      const bool mine = lq == q;
      const unsigned long long bal = __ballot(mine);
      const int cnt = __popcll(bal);
      const int leader = __ffsll((long long)bal) - 1;
      if (!SCATTER) {
        if (lane == leader) atomicAdd(P.slot_cnt + v0 + q, (uint32_t)cnt);
      } else {
        uint32_t base = 0;
        if (lane == leader) base = P.slot_cnt[v0 + q] + atomicAdd(P.slot_cur + v0 + q, (uint32_t)cnt);
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
    __builtin_amdgcn_wave_barrier();
  }
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
}

// Whole histories keep more state live than one query: workgroups of at
// most 256 lanes let the kernel use more than 128 VGPRs instead of spilling
// to scratch; three waves per SIMD (168 VGPRs) is the target.
// STAGED: the start state comes from k_transport_start instead of being
// sampled here.
template <bool MIXED, bool STAGED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_transport(TransportDev P, TransportRun R)
{
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned long long* ltally = (unsigned long long*)smem;
  const size_t tally_bytes = R.lds_tally ? (((size_t)P.nvol * 8 + 15) & ~(size_t)15) : 0;
  int32_t* lstack = (int32_t*)(smem + tally_bytes);
  int32_t* lhist = lstack + (size_t)R.lds_stack * blockDim.x;
  if (R.lds_tally) {
    for (int i = threadIdx.x; i < P.nvol; i += blockDim.x) ltally[i] = 0ull;
    __syncthreads();
  }
  const int64_t gl = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  HybridStack S;
  S.lbase = lstack + threadIdx.x;
  S.lstride = blockDim.x;
  S.gbase = R.spill ? R.spill + gl : nullptr;
  S.gstride = R.spill_lanes;
  S.lcap = R.lds_stack;
  S.cap = R.spill ? R.stack_cap : R.lds_stack;
  S.reset();
  int32_t* hist = lhist + threadIdx.x * HIST;
  const int lane = threadIdx.x & 63;

  uint64_t cur = 0, end = 0;  // claimed history range (relative indices)
  bool alive = false, have = false;
  int64_t hix = 0;
  uint64_t item = 0;
  uint32_t dc = 0;
  D3 r = d3(0, 0, 0), u = d3(0, 0, 0);
  int32_t vol = ID_NONE, ev = 0;
  int hn = 0;
  uint8_t st = TR_ALIVE;
  unsigned long long hh = HASH_SEED;
  while (true) {
    // a lane whose history ended waits until at least R.refill lanes of the
    // wave (or all remaining ones) are idle: starting a history (sampling,
    // find_volume over every volume) is long, and the wave runs it for all
    // lanes that start together
    const bool idle = !alive;
    if (idle && have) {
      if (R.events) R.events[hix] = ev;
      if (R.status) R.status[hix] = st;
      if (R.hash) R.hash[hix] = hh;
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
          if (end > (uint64_t)R.n) end = (uint64_t)R.n;
        }
      }
      if (cur >= (uint64_t)R.n) break;
      hix = STAGED ? (int64_t)R.stage.idx[cur] : (int64_t)cur;
      ++cur;
      have = true;
      item = R.first + (uint64_t)hix;
      ev = 0;
      hh = HASH_SEED;
      hn = 0;
      if (STAGED) {
        const TransportStage& G = R.stage;
        r = d3(G.ru[hix], G.ru[G.cap + hix], G.ru[2 * G.cap + hix]);
        u = d3(G.ru[3 * G.cap + hix], G.ru[4 * G.cap + hix], G.ru[5 * G.cap + hix]);
        vol = G.vol[hix];
        dc = G.dc[hix];
        st = G.st[hix];
      } else {
        sample_start(P, item, dc, r, u, st);
        vol = start_volume<MIXED>(P, r, u, S, st);
      }
      alive = st == TR_ALIVE;
      continue;
    }
    // one event
    const int32_t slot = (vol >= 0 && vol < P.n_ids) ? P.id2slot[vol] : -1;
    const int32_t t = slot >= 0 ? P.vol_tree[slot] : -1;
    if (t < 1 || t > P.ntrees) {
      st = TR_ERROR;
      alive = false;
      continue;
    }
    const SurfTreeDev& T = P.table[t];
    // Particle::surf_dist (particle_sim.h:41-54)
    const RayD ray = make_ray_d3(r, u);
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    trace_s<RAYFIRE, MIXED>(T, ray, ORIENT_EXITING, hist, hn, S, h);
    if (h.surf == ID_NONE) {
      st = TR_LOST;
    } else if (h.t == 0.0) {
      st = TR_STUCK;
    } else {
      // the hit face joins the exclusion history (ray_tracer.cpp:324)
      if (hn == HIST) {
        for (int j = 1; j < HIST; ++j) hist[j - 1] = hist[j];
        --hn;
      }
      hist[hn++] = h.face;
      // sample_collision_distance (particle_sim.h:56-58)
      double ua, ub;
      uniform2(P.seed, P.stream, item, dc++, ua, ub);
      const double dcol = -log_series(1.0 - ua) * P.mfp;
      // advance (particle_sim.h:70-83)
      const double s = dcol < h.t ? dcol : h.t;
      r = d3(r.x + s * u.x, r.y + s * u.y, r.z + s * u.z);
      const unsigned long long q = (unsigned long long)__double2ll_rn(s * 4294967296.0);  // s * 2^32
      if (R.lds_tally)
        atomicAdd(ltally + slot, q);
      else
        atomicAdd(R.tally_fx + slot, q);
      hh = mix64(hh ^ (unsigned long long)(uint32_t)vol);
      hh = mix64(hh ^ (unsigned long long)__double_as_longlong(s));
      ++ev;
      if (dcol < h.t) {
        // collide (particle_sim.h:60-68), absorption first
        if (ub < P.p_absorb) {
          st = TR_ABSORBED;
        } else {
          if (!iso_dir(P.seed, P.stream, item, dc, u)) st = TR_ERROR;
          hn = 0;
        }
      } else {
        // cross_surface (particle_sim.h:85-130): the hit surface's boundary
        // condition (MeshManager::get_surface_property, default transmission)
        const int32_t sid = h.surf;
        const bool known = sid >= 0 && sid < P.nsurf;
        const uint8_t bc = (P.surf_bc && known) ? P.surf_bc[sid] : (uint8_t)BC_TRANSMISSION;
        if (bc == BC_VACUUM) {
          st = TR_VACUUM;
        } else if (bc == BC_REFLECTING) {
          // particle_sim.h:90-104; XDG::surface_normal (xdg.cpp:272-290) is
          // face_normal of the last history entry, the face just hit, whose
          // unflipped normal the tree stores in leaf order
          const double* pn = T.normals + 3 * (int64_t)h.leaf;
          const D3 nrm = d3(pn[0], pn[1], pn[2]);
          const double proj = nrm.x * u.x + nrm.y * u.y + nrm.z * u.z;
          const double mag = sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
          const double sc = 2.0 * proj / mag;
          u = normalize(d3(u.x - sc * nrm.x, u.y - sc * nrm.y, u.z - sc * nrm.z));
          hist[0] = h.face;  // history_ = {history_.back()}
          hn = 1;
        } else {
          // MeshManager::next_volume (mesh_manager_interface.cpp:202-214)
          int32_t nv = ID_NONE;
          if (known) {
            if (P.sense_fwd[sid] == vol)
              nv = P.sense_rev[sid];
            else if (P.sense_rev[sid] == vol)
              nv = P.sense_fwd[sid];
          }
          if (nv == ID_NONE) {
            st = TR_ERROR;
          } else {
            vol = nv;
            if (vol == P.ipc) st = TR_LEAKED;
          }
        }
      }
      if (st == TR_ALIVE && ev >= P.max_events) st = TR_CAPPED;
    }
    alive = st == TR_ALIVE;
  }
  if (R.lds_tally) {
    __syncthreads();
    for (int i = threadIdx.x; i < P.nvol; i += blockDim.x)
      if (ltally[i]) atomicAdd(R.tally_fx + i, ltally[i]);
  }
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
  if (mixed)
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_piv<true>, sb, sh));
  else
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_piv<false>, sb, sh));
  return (int64_t)std::max(nb, 1) * cu;
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
      hipLaunchKernelGGL(k_transport_cand<false>, gn, dim3(256), 0, s, P, Q);
      launch_scan_excl(P.slot_cnt, P.nvol, P.slot_cnt + P.nvol, s);
      hipLaunchKernelGGL(k_transport_cand<true>, gn, dim3(256), 0, s, P, Q);
      const size_t sh = transport_start_lds(sb, R.stack_cap);
      const unsigned gp = (unsigned)std::max<int64_t>(1, R.start_grid);
      if (mixed)
        hipLaunchKernelGGL(k_transport_piv<true>, dim3(gp), dim3(sb), sh, s, P, Q);
      else
        hipLaunchKernelGGL(k_transport_piv<false>, dim3(gp), dim3(sb), sh, s, P, Q);
      const dim3 gs((unsigned)((Q.n + sb - 1) / sb));
      if (mixed)
        hipLaunchKernelGGL(k_transport_resolve<true>, gs, dim3(sb), sh, s, P, Q);
      else
        hipLaunchKernelGGL(k_transport_resolve<false>, gs, dim3(sb), sh, s, P, Q);
    }
    const int64_t need = (Q.n + (int64_t)L.block * Q.chunk - 1) / ((int64_t)L.block * Q.chunk);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min(L.grid, need));
    if (R.staged) {
      if (mixed)
        hipLaunchKernelGGL((k_transport<true, true>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
      else
        hipLaunchKernelGGL((k_transport<false, true>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
    } else {
      if (mixed)
        hipLaunchKernelGGL((k_transport<true, false>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
      else
        hipLaunchKernelGGL((k_transport<false, false>), dim3(g), dim3(L.block), L.lds_bytes, s, P, Q);
    }
  }
}

}  // namespace xdgh
