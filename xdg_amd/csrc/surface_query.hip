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
// Traversal: one ray per lane, wave64; BVH2 nodes carry both child boxes;
// near child first, far child pushed on a per-lane stack in LDS laid out
// [slot][lane] (conflict-free).  PLAIN tests FP64 boxes; MIXED tests the
// outward-rounded FP32 boxes with an FP32 ray and keeps FP64 only for the
// Pluecker test, the distance and the normal/parity decision.  MIXED falls
// back to the FP64 boxes for a ray whose origin lies beyond the tree's
// conservative-rounding limit.
#include "internal.h"
#include "surface_traverse.h"

namespace xdgh {

using namespace xdgd;

namespace {

constexpr int MAXK = 8;

template <bool MIXED>
__global__ void k_ray_fire(RayFireArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
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
  int32_t ex[MAXK];
  int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  for (int j = 0; j < k; ++j) ex[j] = a.excl[(int64_t)a.k * i + j];
  HitState h;
  h.t = a.tmax ? a.tmax[i] : INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  trace<RAYFIRE, MIXED>(T, ray, a.orientation, ex, k, lds + threadIdx.x, blockDim.x, cap, h);
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

// Instrumented ray_fire: identical traversal, plus per-ray work counters
// summed into counters[0..7] = rays, V, F, E, C, N, hits, mixed-fallbacks.
template <bool MIXED>
__global__ void k_ray_fire_stats(RayFireArgs a, int cap, unsigned long long* counters)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int32_t tree = a.tree_per_ray ? a.tree_per_ray[i] : a.tree;
  if (tree < 1 || tree > a.ntrees) return;
  const SurfTreeDev& T = a.table[tree];
  RayD ray = make_ray(a.org + 3 * i, a.dir + 3 * i);
  int32_t ex[MAXK];
  int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  for (int j = 0; j < k; ++j) ex[j] = a.excl[(int64_t)a.k * i + j];
  HitState h;
  h.t = a.tmax ? a.tmax[i] : INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  TraceStats st{0, 0, 0, 0, 0};
  trace<RAYFIRE, MIXED, true>(T, ray, a.orientation, ex, k, lds + threadIdx.x, blockDim.x, cap, h, &st);
  atomicAdd(counters + 0, 1ULL);
  atomicAdd(counters + 1, (unsigned long long)st.V);
  atomicAdd(counters + 2, (unsigned long long)st.F);
  atomicAdd(counters + 3, (unsigned long long)st.E);
  atomicAdd(counters + 4, (unsigned long long)st.C);
  atomicAdd(counters + 5, (unsigned long long)st.N);
  if (h.surf != ID_NONE) atomicAdd(counters + 6, 1ULL);
  if (MIXED && !mixed_safe(T, ray)) atomicAdd(counters + 7, 1ULL);
}

template <bool MIXED>
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
  int32_t ex[MAXK];
  int k = a.excl ? (a.k < MAXK ? a.k : MAXK) : 0;
  for (int j = 0; j < k; ++j) ex[j] = a.excl[(int64_t)a.k * i + j];
  HitState h;
  h.t = INFTY;
  h.face = INT32_MAX;
  h.surf = ID_NONE;
  h.inside = false;
  trace<PIV, MIXED>(T, ray, ORIENT_ANY, ex, k, lds + threadIdx.x, blockDim.x, cap, h);
  a.inside[i] = (h.surf != ID_NONE && h.inside) ? 1 : 0;
}

template <bool MIXED>
__global__ void k_find_volume(FindVolumeArgs a, int cap)
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
    trace<PIV, MIXED>(a.table[tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x, blockDim.x, cap, h);
    if (h.surf != ID_NONE && h.inside) {
      found = vid;
      break;
    }
  }
  a.vol_out[i] = found;
}

template <bool MIXED>
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
    hit = trace<OCCL, MIXED>(a.table[a.tree], ray, ORIENT_ANY, nullptr, 0, lds + threadIdx.x, blockDim.x, cap, h);
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

}  // namespace

void launch_ray_fire(const RayFireArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_ray_fire<true>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
  else
    hipLaunchKernelGGL(k_ray_fire<false>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_ray_fire_stats(const RayFireArgs& a, bool mixed, const LaunchCfg& c, unsigned long long* counters,
                           hipStream_t s)
{
  if (a.n <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_ray_fire_stats<true>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack,
                       counters);
  else
    hipLaunchKernelGGL(k_ray_fire_stats<false>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack,
                       counters);
}

void launch_point_in_volume(const PivArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_piv<true>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
  else
    hipLaunchKernelGGL(k_piv<false>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_find_volume(const FindVolumeArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_find_volume<true>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
  else
    hipLaunchKernelGGL(k_find_volume<false>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

void launch_occluded(const OccludedArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_occluded<true>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
  else
    hipLaunchKernelGGL(k_occluded<false>, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
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
  r.surf2 = (surf[p] << 1) | (flip[p] ? 1 : 0);
  out[s] = r;
}

}  // namespace

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
