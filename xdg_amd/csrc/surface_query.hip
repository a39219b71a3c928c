// surface_query.hip -- batched surface-tree queries on one MI355X.
//
// Reference semantics (one ray per call there, one ray per lane here):
//   ray_fire         EmbreeRayTracer::ray_fire  src/embree/ray_tracer.cpp:290-326
//                    + TriangleIntersectionFunc src/triangle_intersect.cpp:52-109
//   point_in_volume  ray_tracer.cpp:259-287
//   find_volume      XDG::find_volume src/xdg.cpp:113-128
//   occluded         ray_tracer.cpp:347-370 + TriangleOcclusionFunc
//                    triangle_intersect.cpp:139-161
//   closest          ray_tracer.cpp:328-345 + TriangleClosestFunc
//                    triangle_intersect.cpp:111-137
// The answer of every query is the exact FP64 minimum over all primitives
// passing the reference's culls (the BVH only prunes conservatively), with
// one deterministic tie rule: among bitwise-equal distances the smallest
// face id wins (the reference keeps the last candidate Embree visits).
//
// Traversal: one ray per lane, wave64, over the BVH4 (WIDE, default) or the
// binary LBVH; near child first, others pushed on a per-lane stack in LDS
// laid out [slot][lane] (conflict-free).  PLAIN tests FP64 boxes; MIXED
// tests the outward-rounded FP32 boxes with an FP32 ray and keeps FP64 only
// for the Pluecker test, the distance and the normal/parity decision.  A
// mixed ray whose origin lies beyond the tree's conservative-rounding limit
// is traced with FP64 boxes (for ray_fire: queued to a second pass so that
// the mixed kernel carries no FP64 box state).
#include <algorithm>

#include "internal.h"
#include "surface_traverse.h"
#include "xdg_hip.h"

namespace xdgh {

using namespace xdgd;

namespace {

constexpr int MAXK = XDG_HIP_MAX_EXCLUDE;  // exclusion entries per ray (read in place)

// One ray of ray_fire.  DEFER (mixed only): a ray beyond the FP32-rounding
// limit is queued for the FP64 pass instead of traced here.
template <bool MIXED, bool DEFER, bool WIDE, bool SPEC = false>
__device__ __forceinline__ void ray_fire_one(const RayFireArgs& a, int64_t i, int32_t* stk, int cap)
{
  int32_t tree = a.tree_per_ray ? a.tree_per_ray[i] : a.tree;
  double* tout = a.t_out;
  if (tree < 1 || tree > a.ntrees) {
    tout[i] = INFTY;
    a.surf_out[i] = ID_NONE;
    if (a.face_out) a.face_out[i] = ID_NONE;
    return;
  }
  const SurfTreeDev& T = a.table[tree];
  RayD ray = make_ray(a.org + 3 * i, a.dir + 3 * i);
  if (DEFER && !mixed_safe(T, ray)) {
    const unsigned int j = atomicAdd(a.fb_count, 1u);
    a.fb_list[j] = (int32_t)i;
    return;
  }
  const int32_t* ex = a.excl ? a.excl + (int64_t)a.k * i : nullptr;
  const int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  HitState h;
  h.t = a.tmax ? a.tmax[i] : INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  if (DEFER)
    trace_f32<RAYFIRE, false, WIDE, SPEC>(T, ray, a.orientation, ex, k, stk, blockDim.x, cap, h);
  else
    trace<RAYFIRE, MIXED, false, WIDE, SPEC>(T, ray, a.orientation, ex, k, stk, blockDim.x, cap, h);
  if (h.surf == ID_NONE) {
    tout[i] = INFTY;
    a.surf_out[i] = ID_NONE;
    if (a.face_out) a.face_out[i] = ID_NONE;
  } else {
    tout[i] = h.t;
    a.surf_out[i] = h.surf;
    if (a.face_out) a.face_out[i] = h.face;
  }
}

template <bool MIXED, bool DEFER, bool WIDE, bool SPEC = false>
__global__ void k_ray_fire(RayFireArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  ray_fire_one<MIXED, DEFER, WIDE, SPEC>(a, a.order ? (int64_t)a.order[i] : i, lds + threadIdx.x, cap);
}

// FP64 pass over the rays the mixed kernel queued (grid-stride; the count is
// read on the device, so no host round trip separates the two passes).
template <bool WIDE>
__global__ void k_ray_fire_deferred(RayFireArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  const unsigned int cnt = *a.fb_count;
  for (unsigned int j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x)
    ray_fire_one<false, false, WIDE>(a, a.fb_list[j], lds + threadIdx.x, cap);
}

__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }
__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

__device__ __forceinline__ void write_hit(const RayFireArgs& a, int64_t i, const HitState& h)
{
  if (h.surf == ID_NONE) {
    a.t_out[i] = INFTY;
    a.surf_out[i] = ID_NONE;
    if (a.face_out) a.face_out[i] = ID_NONE;
  } else {
    a.t_out[i] = h.t;
    a.surf_out[i] = h.surf;
    if (a.face_out) a.face_out[i] = h.face;
  }
}

// Persistent ray_fire over the BVH4.  A fixed grid of resident waves works
// through the batch: a lane whose ray is finished writes its result and
// takes the next unclaimed ray (one 64-bit atomic per wave per refill), so a
// wave does not idle behind its slowest lane.  A lane that reaches a leaf
// while others still descend parks it and keeps descending until every
// lane holds one (speculative traversal, Aila & Laine 2009).  MIXED queues far-origin rays
// for the FP64 pass (k_ray_fire_deferred).  Every lane leaves the loop once
// the counter passes n, so the grid always drains.
template <bool MIXED>
__device__ __forceinline__ void ray_fire_pt_body(const RayFireArgs& a, int cap, unsigned long long* next_ray,
                                                 int refill_min)
{
  extern __shared__ int32_t lds[];
  LdsStack S(lds + threadIdx.x, blockDim.x, cap);
  const int lane = threadIdx.x & 63;
  const int32_t DONE = CHILD_EMPTY;
  int64_t ri = -1;
  bool have = false;  // ri's result is pending
  TreeView T{nullptr, nullptr, nullptr, nullptr};
  RayD ray;
  BoxRayF bf{0, 0, 0, 0, 0, 0};
  BoxRayD bd{0, 0, 0, 0, 0, 0};
  HitState h{INFTY, INT32_MAX, ID_NONE, false};
  const int32_t* ex = nullptr;
  int k = 0;
  float tf = 0.0f;
  PlaneOfs po{0, 0, 0, 0, 0, 0};
  int32_t node = DONE;
  int32_t leaf = DONE;  // a leaf parked while the wave keeps descending
  while (true) {
    const bool done = node == DONE && leaf == DONE;
    if (done && have) {
      write_hit(a, ri, h);
      have = false;
    }
    // Refill only once enough lanes are idle (or every remaining lane is):
    // the refill path is long, and running it for one lane at a time costs
    // the whole wave more than the idle lanes waste.
    const unsigned long long dmask = __ballot(done);
    const int nact = __popcll(__ballot(1));
    const bool go = __popcll(dmask) >= (refill_min < nact ? refill_min : nact);
    const bool need = go && done;
    const unsigned long long mask = go ? dmask : 0ull;
    unsigned long long base = 0;
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long b = 0;
      if (lane == leader) b = atomicAdd(next_ray, (unsigned long long)__popcll(mask));
      base = __shfl(b, leader);
    }
    if (need) {
      const int64_t slot = (int64_t)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
      if (slot >= a.n) break;
      ri = a.order ? (int64_t)a.order[slot] : slot;
      const int32_t tree = a.tree_per_ray ? a.tree_per_ray[ri] : a.tree;
      if (tree < 1 || tree > a.ntrees) {
        HitState miss{INFTY, INT32_MAX, ID_NONE, false};
        write_hit(a, ri, miss);
        continue;
      }
      const SurfTreeDev& TT = a.table[tree];
      ray = make_ray(a.org + 3 * ri, a.dir + 3 * ri);
      if (MIXED && !mixed_safe(TT, ray)) {
        const unsigned int j = atomicAdd(a.fb_count, 1u);
        a.fb_list[j] = (int32_t)ri;
        continue;
      }
      T.nodes4_d = TT.nodes4_d;
      T.nodes4_f = TT.nodes4_f;
      T.tris = TT.tris;
      T.normals = TT.normals;
      ex = a.excl ? a.excl + (int64_t)a.k * ri : nullptr;
      k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
      h.t = a.tmax ? a.tmax[ri] : INFTY;
      h.face = INT32_MAX;
      h.surf = ID_NONE;
      h.inside = false;
      tf = f32_up(h.t);
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
      S.reset();
      have = true;
      node = TT.n_prims > 0 ? TT.root4 : DONE;
      if (node == DONE) continue;
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
      for (int32_t j = 0; j < cnt; ++j) tri_closest_hit<RAYFIRE>(T, first + j, ray, a.orientation, ex, k, h);
      tf = f32_up(h.t);
    }
  }
}

__global__ void k_ray_fire_pt_mixed(RayFireArgs a, int cap, unsigned long long* next_ray, int refill_min)
{
  ray_fire_pt_body<true>(a, cap, next_ray, refill_min);
}
__global__ void k_ray_fire_pt_plain(RayFireArgs a, int cap, unsigned long long* next_ray, int refill_min)
{
  ray_fire_pt_body<false>(a, cap, next_ray, refill_min);
}

// Instrumented ray_fire: identical traversal, plus per-ray work counters
// summed into counters[0..7] = rays, V, F, E, C, N, hits, mixed-fallbacks.
template <bool MIXED, bool WIDE>
__global__ void k_ray_fire_stats(RayFireArgs a, int cap, unsigned long long* counters, int spec)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.order) i = a.order[i];
  int32_t tree = a.tree_per_ray ? a.tree_per_ray[i] : a.tree;
  if (tree < 1 || tree > a.ntrees) return;
  const SurfTreeDev& T = a.table[tree];
  RayD ray = make_ray(a.org + 3 * i, a.dir + 3 * i);
  const int32_t* ex = a.excl ? a.excl + (int64_t)a.k * i : nullptr;
  const int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  HitState h;
  h.t = a.tmax ? a.tmax[i] : INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  TraceStats st{0, 0, 0, 0, 0, 0, 0};
  if (WIDE && spec)
    trace<RAYFIRE, MIXED, true, WIDE, true>(T, ray, a.orientation, ex, k, lds + threadIdx.x, blockDim.x, cap, h, &st);
  else
    trace<RAYFIRE, MIXED, true, WIDE>(T, ray, a.orientation, ex, k, lds + threadIdx.x, blockDim.x, cap, h, &st);
  atomicAdd(counters + 0, 1ULL);
  atomicAdd(counters + 1, (unsigned long long)st.V);
  atomicAdd(counters + 2, (unsigned long long)st.F);
  atomicAdd(counters + 3, (unsigned long long)st.E);
  atomicAdd(counters + 4, (unsigned long long)st.C);
  atomicAdd(counters + 5, (unsigned long long)st.N);
  if (h.surf != ID_NONE) atomicAdd(counters + 6, 1ULL);
  if (MIXED && !mixed_safe(T, ray)) atomicAdd(counters + 7, 1ULL);
  if (st.WV) atomicAdd(counters + 8, (unsigned long long)st.WV);
  if (st.WF) atomicAdd(counters + 9, (unsigned long long)st.WF);
  // longest ray of the wave (node visits, triangle tests): the iterations a
  // wave needs if lanes never wait for each other between leaves
  const unsigned int mv = wave_max_u24(st.V), mf = wave_max_u24(st.F);
  if (first_active_lane()) {
    atomicAdd(counters + 10, (unsigned long long)mv);
    atomicAdd(counters + 11, (unsigned long long)mf);
  }
}

template <bool MIXED, bool WIDE>
__global__ void k_piv(PivArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.tree < 1 || a.tree > a.ntrees) {
    a.inside[i] = 0;
    return;
  }
  const SurfTreeDev& T = a.table[a.tree];
  const double* d = a.dir ? a.dir + 3 * i : a.dflt;
  RayD ray = make_ray(a.pts + 3 * i, d);
  const int32_t* ex = a.excl ? a.excl + (int64_t)a.k * i : nullptr;
  const int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  HitState h;
  h.t = INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  trace<PIV, MIXED, false, WIDE>(T, ray, ORIENT_ANY, ex, k, lds + threadIdx.x, blockDim.x, cap, h);
  a.inside[i] = (h.surf != ID_NONE && h.inside) ? 1 : 0;
}

template <bool MIXED, bool WIDE>
__global__ void k_find_volume(FindVolumeArgs a, int cap, int spec)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double* d = a.dir ? a.dir + 3 * i : a.dflt;
  RayD ray = make_ray(a.pts + 3 * i, d);
  int32_t found = a.ipc;
  for (int v = 0; v < a.nvol; ++v) {
    const int32_t vid = a.vol_ids[v];
    const int32_t tree = a.vol_trees[v];
    if (vid == a.ipc || tree < 1 || tree > a.ntrees) continue;
    HitState h;
    h.t = INFTY;
    h.face = INT32_MAX;
    h.surf = ID_NONE;
    h.inside = false;
    if (WIDE && spec)
      trace<PIV, MIXED, false, WIDE, true>(a.table[tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x, blockDim.x,
                                           cap, h);
    else
      trace<PIV, MIXED, false, WIDE>(a.table[tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x, blockDim.x, cap,
                                     h);
    if (h.surf != ID_NONE && h.inside) {
      found = vid;
      break;
    }
  }
  a.vol_out[i] = found;
}

__device__ __forceinline__ void find_volume_point(const FindVolumeArgs& a, int64_t i, int32_t* stk, int cap)
{
  const double* d = a.dir ? a.dir + 3 * i : a.dflt;
  RayD ray = make_ray(a.pts + 3 * i, d);
  int32_t found = a.ipc;
  for (int v = 0; v < a.nvol; ++v) {
    const int32_t vid = a.vol_ids[v];
    const int32_t tree = a.vol_trees[v];
    if (vid == a.ipc || tree < 1 || tree > a.ntrees) continue;
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    trace<PIV, false, false, true>(a.table[tree], ray, ORIENT_ANY, nullptr, 0, stk, blockDim.x, cap, h);
    if (h.surf != ID_NONE && h.inside) {
      found = vid;
      break;
    }
  }
  a.vol_out[i] = found;
}

// FP64 pass over the points the mixed persistent kernel queued
__global__ void k_find_volume_deferred(FindVolumeArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  const unsigned int cnt = *a.fb_count;
  for (unsigned int j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x)
    find_volume_point(a, a.fb_list[j], lds + threadIdx.x, cap);
}

// Persistent find_volume over the BVH4: a lane holds one point and runs its
// point_in_volume traversals volume after volume (a new traversal on the
// same ray costs only a reset), as a per-lane state machine like the
// persistent ray_fire: lanes whose traversal has ended record the answer or
// start the next volume, and claim new points, together once refill_min of
// the wave are ready, while the other lanes keep descending (with
// speculative leaf parking).
template <bool MIXED>
__device__ __forceinline__ void find_volume_pt_body(const FindVolumeArgs& a, int cap, int refill_min)
{
  extern __shared__ int32_t lds[];
  LdsStack S(lds + threadIdx.x, blockDim.x, cap);
  const int lane = threadIdx.x & 63;
  const int32_t DONE = CHILD_EMPTY;
  int64_t pi = 0;
  bool hasp = false;   // the lane holds a point whose volume is not known yet
  bool traced = false; // a point_in_volume traversal of volume v is under way or ended
  int v = 0;
  TreeView T{nullptr, nullptr, nullptr, nullptr};
  RayD ray;
  BoxRayF bf{0, 0, 0, 0, 0, 0};
  BoxRayD bd{0, 0, 0, 0, 0, 0};
  PlaneOfs po{0, 0, 0, 0, 0, 0};
  HitState h{INFTY, INT32_MAX, ID_NONE, false};
  float tf = INFINITY;
  int32_t node = DONE, leaf = DONE;
  while (true) {
    // lanes whose traversal has ended wait until refill_min of the wave (or
    // all remaining ones) have too; then they act together: record or
    // advance to the next volume, claim new points, start traversals
    const bool done = node == DONE && leaf == DONE;
    const unsigned long long dmask = __ballot(done);
    const int nact = __popcll(__ballot(1));
    const bool go = done && __popcll(dmask) >= (refill_min < nact ? refill_min : nact);
    if (go && traced) {
      traced = false;
      if (h.surf != ID_NONE && h.inside) {
        a.vol_out[pi] = a.vol_ids[v];
        hasp = false;
      } else {
        ++v;
      }
    }
    const bool need = go && !hasp;
    const unsigned long long cmask = __ballot(need);
    if (cmask) {
      const int leader = __ffsll((long long)cmask) - 1;
      unsigned long long b = 0;
      if (lane == leader) b = atomicAdd(a.next, (unsigned long long)__popcll(cmask));
      b = __shfl(b, leader);
      if (need) {
        pi = (int64_t)(b + (unsigned long long)__popcll(cmask & ((1ull << lane) - 1ull)));
        if (pi >= a.n) break;
        const double* d = a.dir ? a.dir + 3 * pi : a.dflt;
        ray = make_ray(a.pts + 3 * pi, d);
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
        v = 0;
        hasp = true;
      }
    }
    if (go && hasp) {  // the point's next volume
      while (v < a.nvol && (a.vol_ids[v] == a.ipc || a.vol_trees[v] < 1 || a.vol_trees[v] > a.ntrees)) ++v;
      if (v >= a.nvol) {
        a.vol_out[pi] = a.ipc;
        hasp = false;
      } else {
        const SurfTreeDev& TT = a.table[a.vol_trees[v]];
        if (MIXED && !mixed_safe(TT, ray)) {  // the whole point goes to the FP64 pass
          const unsigned int j = atomicAdd(a.fb_count, 1u);
          a.fb_list[j] = (int32_t)pi;
          hasp = false;
        } else {
          T.nodes4_d = TT.nodes4_d;
          T.nodes4_f = TT.nodes4_f;
          T.tris = TT.tris;
          T.normals = TT.normals;
          h = HitState{INFTY, INT32_MAX, ID_NONE, false};
          tf = INFINITY;
          S.reset();
          node = TT.n_prims > 0 ? TT.root4 : DONE;
          traced = true;
        }
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

__global__ void k_find_volume_pt_mixed(FindVolumeArgs a, int cap, int refill_min)
{
  find_volume_pt_body<true>(a, cap, refill_min);
}
__global__ void k_find_volume_pt_plain(FindVolumeArgs a, int cap, int refill_min)
{
  find_volume_pt_body<false>(a, cap, refill_min);
}

template <bool MIXED, bool WIDE>
__global__ void k_occluded(OccludedArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  bool hit = false;
  if (a.tree >= 1 && a.tree <= a.ntrees) {
    RayD ray = make_ray(a.org + 3 * i, a.dir + 3 * i);
    HitState h;
    h.t = INFTY;
    h.face = INT32_MAX;
    h.surf = ID_NONE;
    h.inside = false;
    hit = trace<OCCL, MIXED, false, WIDE>(a.table[a.tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x,
                                          blockDim.x, cap, h);
  }
  a.hit[i] = hit ? 1 : 0;
  if (a.dist) a.dist[i] = hit ? -INFTY : INFTY;
}

// squared distance from p to an FP64 box
__device__ __forceinline__ double box_d2(const double* b, const D3& p)
{
  double ex = fmax(fmax(b[0] - p.x, p.x - b[3]), 0.0);
  double ey = fmax(fmax(b[1] - p.y, p.y - b[4]), 0.0);
  double ez = fmax(fmax(b[2] - p.z, p.z - b[5]), 0.0);
  return ex * ex + ey * ey + ez * ez;
}

__global__ void k_closest(ClosestArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double best = INFTY;
  int32_t bface = INT32_MAX;
  if (a.tree >= 1 && a.tree <= a.ntrees && a.table[a.tree].n_prims > 0) {
    const SurfTreeDev& T = a.table[a.tree];
    const D3 p = d3(a.pts[3 * i], a.pts[3 * i + 1], a.pts[3 * i + 2]);
    int32_t* stk = lds + threadIdx.x;
    const int stride = blockDim.x;
    int sp = 0;
    int32_t node = T.root;
    while (true) {
      if (node >= 0) {
        const NodeD nd = T.nodes_d[node];
        double d0 = box_d2(nd.b0, p), d1 = box_d2(nd.b1, p);
        double lim = best * best;
        bool h0 = d0 <= lim, h1 = d1 <= lim;
        int32_t c0 = nd.c0, c1 = nd.c1;
        if (h0 && h1) {
          if (d1 < d0) {
            int32_t tmp = c0;
            c0 = c1;
            c1 = tmp;
          }
          if (sp < cap) stk[sp++ * stride] = c1;
          node = c0;
        } else if (h0) {
          node = c0;
        } else if (h1) {
          node = c1;
        } else {
          if (sp == 0) break;
          node = stk[--sp * stride];
        }
        continue;
      }
      const int32_t first = leaf_first(node);
      const int32_t cnt = leaf_count(node);
      for (int32_t j = 0; j < cnt; ++j) {
        const TriRec& r = T.tris[first + j];
        const D3 v0 = d3(r.v[0], r.v[1], r.v[2]);
        const D3 v1 = d3(r.v[3], r.v[4], r.v[5]);
        const D3 v2 = d3(r.v[6], r.v[7], r.v[8]);
        const D3 q = closest_on_triangle(v0, v1, v2, p);
        const D3 e = sub(q, p);
        const double dist = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
        if (dist < best || (dist == best && r.face < bface)) {
          best = dist;
          bface = r.face;
        }
      }
      if (sp == 0) break;
      node = stk[--sp * stride];
    }
  }
  if (bface == INT32_MAX) {
    a.dist[i] = INFTY;
    a.face[i] = ID_NONE;
  } else {
    a.dist[i] = best;
    a.face[i] = bface;
  }
}

inline dim3 grid1(int64_t n, int blk) { return dim3((unsigned)((n + blk - 1) / blk)); }
inline size_t lds_bytes(const LaunchCfg& c) { return sizeof(int32_t) * (size_t)c.stack * c.block; }

#define XDG_DISPATCH2(KERNEL, MIXED, WIDE, ...)                                              \
  do {                                                                                       \
    if (MIXED) {                                                                             \
      if (WIDE)                                                                              \
        hipLaunchKernelGGL((KERNEL<true, true>), __VA_ARGS__);                               \
      else                                                                                   \
        hipLaunchKernelGGL((KERNEL<true, false>), __VA_ARGS__);                              \
    } else {                                                                                 \
      if (WIDE)                                                                              \
        hipLaunchKernelGGL((KERNEL<false, true>), __VA_ARGS__);                              \
      else                                                                                   \
        hipLaunchKernelGGL((KERNEL<false, false>), __VA_ARGS__);                             \
    }                                                                                        \
  } while (0)

}  // namespace

void launch_ray_fire(const RayFireArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  const dim3 g = grid1(a.n, c.block), b(c.block);
  const size_t sh = lds_bytes(c);
  const bool persist = (c.persistent & (mixed ? 2 : 1)) != 0;
  if (persist && c.wide && a.next_ray && (!mixed || (a.fb_list && a.fb_count))) {
    const int64_t need = (a.n + c.block - 1) / c.block;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, a.next_ray);
    if (mixed) {
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, a.fb_count);
      const int64_t gp = std::min(need, resident_grid(k_ray_fire_pt_mixed, c.block, sh));
      hipLaunchKernelGGL(k_ray_fire_pt_mixed, dim3((unsigned)gp), b, sh, s, a, c.stack, a.next_ray, c.refill);
      const unsigned gd = (unsigned)std::min<int64_t>(need, 1024);
      hipLaunchKernelGGL((k_ray_fire_deferred<true>), dim3(gd), b, sh, s, a, c.stack);
    } else {
      const int64_t gp = std::min(need, resident_grid(k_ray_fire_pt_plain, c.block, sh));
      hipLaunchKernelGGL(k_ray_fire_pt_plain, dim3((unsigned)gp), b, sh, s, a, c.stack, a.next_ray, c.refill);
    }
    return;
  }
  if (mixed && a.fb_list && a.fb_count) {
    const unsigned gd = (unsigned)std::min<int64_t>((a.n + c.block - 1) / c.block, 1024);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, a.fb_count);
    if (c.wide) {
      if (c.spec)
        hipLaunchKernelGGL((k_ray_fire<true, true, true, true>), g, b, sh, s, a, c.stack);
      else
        hipLaunchKernelGGL((k_ray_fire<true, true, true>), g, b, sh, s, a, c.stack);
      hipLaunchKernelGGL((k_ray_fire_deferred<true>), dim3(gd), b, sh, s, a, c.stack);
    } else {
      hipLaunchKernelGGL((k_ray_fire<true, true, false>), g, b, sh, s, a, c.stack);
      hipLaunchKernelGGL((k_ray_fire_deferred<false>), dim3(gd), b, sh, s, a, c.stack);
    }
  } else if (mixed) {
    if (c.wide)
      hipLaunchKernelGGL((k_ray_fire<true, false, true>), g, b, sh, s, a, c.stack);
    else
      hipLaunchKernelGGL((k_ray_fire<true, false, false>), g, b, sh, s, a, c.stack);
  } else {
    if (c.wide && c.spec)
      hipLaunchKernelGGL((k_ray_fire<false, false, true, true>), g, b, sh, s, a, c.stack);
    else if (c.wide)
      hipLaunchKernelGGL((k_ray_fire<false, false, true>), g, b, sh, s, a, c.stack);
    else
      hipLaunchKernelGGL((k_ray_fire<false, false, false>), g, b, sh, s, a, c.stack);
  }
}

void launch_ray_fire_stats(const RayFireArgs& a, bool mixed, const LaunchCfg& c, unsigned long long* counters,
                           hipStream_t s)
{
  if (a.n <= 0) return;
  XDG_DISPATCH2(k_ray_fire_stats, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack,
                counters, c.spec);
}

void launch_point_in_volume(const PivArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  XDG_DISPATCH2(k_piv, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_find_volume(const FindVolumeArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  // bit 2 of the persistent mask: the resident grid whose lanes move on to
  // their next volume or point while others still traverse (default), or
  // the per-point kernel that keeps a wave's lanes on the same tree volume
  // after volume but waits for its longest traversal at each.  The refill
  // threshold that measured fastest differs by precision (nested spheres,
  // 10 M points: mixed 7.7 ms at 56, plain 10.4 ms at 40-44).
  const bool persist = (c.persistent & 4) != 0;
  const int refill = c.fv_refill ? c.fv_refill : (mixed ? 56 : 40);
  if (persist && c.wide && a.next && (!mixed || (a.fb_list && a.fb_count))) {
    const size_t sh = lds_bytes(c);
    const int64_t need = (a.n + c.block - 1) / c.block;
    const dim3 b(c.block);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, a.next);
    if (mixed) {
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, a.fb_count);
      const int64_t gp = std::min(need, resident_grid(k_find_volume_pt_mixed, c.block, sh));
      hipLaunchKernelGGL(k_find_volume_pt_mixed, dim3((unsigned)gp), b, sh, s, a, c.stack, refill);
      const unsigned gd = (unsigned)std::min<int64_t>(need, 1024);
      hipLaunchKernelGGL(k_find_volume_deferred, dim3(gd), b, sh, s, a, c.stack);
    } else {
      const int64_t gp = std::min(need, resident_grid(k_find_volume_pt_plain, c.block, sh));
      hipLaunchKernelGGL(k_find_volume_pt_plain, dim3((unsigned)gp), b, sh, s, a, c.stack, refill);
    }
    return;
  }
  // speculative leaf parking measured faster for the plain kernel only
  // (nested spheres, 10 M points: plain 14.98 -> 14.16 ms, mixed 9.67 -> 9.77 ms)
  XDG_DISPATCH2(k_find_volume, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack,
                mixed ? 0 : c.spec);
}

// check_location_for_overlap (tools/overlap.cpp:9-61): point_in_volume of
// each location, with its own direction and no exclusions, against every
// volume of the list in order (no early exit).  The first max_k volumes
// that contain it are written, and their count.
template <bool MIXED, bool WIDE>
__global__ void k_point_in_volumes(VolumesArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const RayD ray = make_ray(a.pts + 3 * i, a.dir + 3 * i);
  int32_t* ids = a.ids_out + (int64_t)a.max_k * i;
  int cnt = 0;
  for (int v = 0; v < a.nvol; ++v) {
    const int32_t tree = a.vol_trees[v];
    if (tree < 1 || tree > a.ntrees) continue;
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    trace<PIV, MIXED, false, WIDE>(a.table[tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x, blockDim.x, cap,
                                   h);
    if (h.surf != ID_NONE && h.inside) {
      if (cnt < a.max_k) ids[cnt] = a.vol_ids[v];
      ++cnt;
    }
  }
  a.count_out[i] = cnt;
}

// check_along_edge (tools/overlap.cpp:229-265): XDG::ray_fire(vol, origin,
// direction, edge length) (EXITING, no exclusions) against each volume of the
// list in order except the ray's two skipped volumes (the parents of the
// edge's surface); the first volume with a hit ends the search.
template <bool MIXED, bool WIDE>
__global__ void k_edge_rays(VolumesArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const RayD ray = make_ray(a.pts + 3 * i, a.dir + 3 * i);
  const int32_t s0 = a.skip[2 * i], s1 = a.skip[2 * i + 1];
  const double lim = a.len[i];
  double t = INFTY;
  int32_t surf = ID_NONE, vol = ID_NONE;
  for (int v = 0; v < a.nvol; ++v) {
    const int32_t vid = a.vol_ids[v];
    const int32_t tree = a.vol_trees[v];
    if (vid == s0 || vid == s1 || tree < 1 || tree > a.ntrees) continue;
    HitState h{lim, INT32_MAX, ID_NONE, false};
    trace<RAYFIRE, MIXED, false, WIDE>(a.table[tree], ray, ORIENT_EXITING, nullptr, 0, lds + threadIdx.x, blockDim.x,
                                       cap, h);
    if (h.surf != ID_NONE) {
      t = h.t;
      surf = h.surf;
      vol = vid;
      break;
    }
  }
  a.t_out[i] = t;
  a.surf_out[i] = surf;
  a.vol_out[i] = vol;
}

void launch_point_in_volumes(const VolumesArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  XDG_DISPATCH2(k_point_in_volumes, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_edge_rays(const VolumesArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  XDG_DISPATCH2(k_edge_rays, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_occluded(const OccludedArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  XDG_DISPATCH2(k_occluded, mixed, c.wide, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_closest(const ClosestArgs& a, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_closest, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

// ---- primitive evaluation (parity checks of the device geometry kernels) --

namespace {

__global__ void k_eval(int kind, int64_t n, const double* __restrict__ in, double* __restrict__ out)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == 0) {  // plucker: v0 v1 v2 o d tmax tmin use ori -> hit, t
    const double* c = in + 19 * i;
    const D3 v0 = d3(c[0], c[1], c[2]), v1 = d3(c[3], c[4], c[5]), v2 = d3(c[6], c[7], c[8]);
    const D3 o = d3(c[9], c[10], c[11]), d = d3(c[12], c[13], c[14]);
    const D3 rb = cross(d, o);
    const int idx = major_axis(d);
    const double oi = pick3(o.x, o.y, o.z, idx), di = pick3(d.x, d.y, d.z, idx);
    double t = 0.0;
    bool hit;
    if (c[17] != 0.0)
      hit = plucker_oriented(v0, v1, v2, d, rb, idx, oi, di, (double)(int)c[18], c[15], c[16], t);
    else
      hit = plucker_any(v0, v1, v2, d, rb, idx, oi, di, c[15], c[16], t);
    out[2 * i] = hit ? 1.0 : 0.0;
    out[2 * i + 1] = hit ? t : 0.0;
  } else if (kind == 1) {  // face normal: 9 -> 3
    const double* c = in + 9 * i;
    const D3 nrm = face_normal(d3(c[0], c[1], c[2]), d3(c[3], c[4], c[5]), d3(c[6], c[7], c[8]));
    out[3 * i] = nrm.x;
    out[3 * i + 1] = nrm.y;
    out[3 * i + 2] = nrm.z;
  } else if (kind == 2) {  // closest point: 12 -> 3
    const double* c = in + 12 * i;
    const D3 q = closest_on_triangle(d3(c[0], c[1], c[2]), d3(c[3], c[4], c[5]), d3(c[6], c[7], c[8]),
                                     d3(c[9], c[10], c[11]));
    out[3 * i] = q.x;
    out[3 * i + 1] = q.y;
    out[3 * i + 2] = q.z;
  } else if (kind == 3) {  // tet containment: p v0 v1 v2 v3 (15) -> 1
    const double* c = in + 15 * i;
    out[i] = tet_contains(d3(c[0], c[1], c[2]), d3(c[3], c[4], c[5]), d3(c[6], c[7], c[8]),
                          d3(c[9], c[10], c[11]), d3(c[12], c[13], c[14]))
                 ? 1.0
                 : 0.0;
  } else if (kind == 4) {  // sqrt (1 -> 1)
    out[i] = sqrt(in[i]);
  } else if (kind == 5) {  // division a / b (2 -> 1)
    out[i] = in[2 * i] / in[2 * i + 1];
  }
}

}  // namespace

void launch_eval(int kind, int64_t n, const double* in, double* out, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_eval, grid1(n, 256), dim3(256), 0, s, kind, n, in, out);
}

// ---- build helpers --------------------------------------------------------

namespace {

__global__ void k_tri_boxes(const double* __restrict__ v9, int64_t n, double bump, double* box6)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* v = v9 + 9 * i;
  for (int k = 0; k < 3; ++k) {
    double lo = fmin(fmin(v[k], v[3 + k]), v[6 + k]);
    double hi = fmax(fmax(v[k], v[3 + k]), v[6 + k]);
    box6[6 * i + k] = lo - bump;
    box6[6 * i + 3 + k] = hi + bump;
  }
}

__global__ void k_write_tris(const double* __restrict__ v9, const int32_t* __restrict__ face,
                             const int32_t* __restrict__ surf, const uint8_t* __restrict__ flip,
                             const int32_t* __restrict__ order, int64_t n, TriRec* out)
{
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  int64_t p = order[s];
  TriRec r;
  for (int k = 0; k < 9; ++k) r.v[k] = v9[9 * p + k];
  r.face = face[p];
  const unsigned fb = tri_first_bits(d3(r.v[0], r.v[1], r.v[2]), d3(r.v[3], r.v[4], r.v[5]), d3(r.v[6], r.v[7], r.v[8]));
  r.surf2 = (surf[p] << TRI_SURF_SHIFT) | (int32_t)(fb << 1) | (flip[p] ? 1 : 0);
  out[s] = r;
}

// face_normal of every triangle record (mesh_manager_interface.cpp:216-220)
__global__ void k_tri_normals(const TriRec* __restrict__ tris, int64_t n, double* out)
{
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  const TriRec& r = tris[s];
  const D3 nrm = face_normal(d3(r.v[0], r.v[1], r.v[2]), d3(r.v[3], r.v[4], r.v[5]), d3(r.v[6], r.v[7], r.v[8]));
  out[3 * s] = nrm.x;
  out[3 * s + 1] = nrm.y;
  out[3 * s + 2] = nrm.z;
}

}  // namespace

void launch_tri_normals(const TriRec* tris, int64_t n, double* out, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_tri_normals, grid1(n, 256), dim3(256), 0, s, tris, n, out);
}

void launch_tri_boxes(const double* verts9, int64_t n, double bump, double* box6, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_tri_boxes, grid1(n, 256), dim3(256), 0, s, verts9, n, bump, box6);
}

void launch_write_tris(const double* verts9, const int32_t* face, const int32_t* surf, const uint8_t* flip,
                       const int32_t* order, int64_t n, TriRec* out, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_write_tris, grid1(n, 256), dim3(256), 0, s, verts9, face, surf, flip, order, n, out);
}

}  // namespace xdgh
