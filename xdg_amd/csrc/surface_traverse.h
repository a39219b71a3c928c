// surface_traverse.h -- BVH traversal core for surface-tree queries.
//
// Shared by the surface query kernels and the tet walker (which fires rays
// against the implicit complement, src/xdg.cpp:164-185).  See
// surface_query.hip for the semantics and the tie rule.
#pragma once

#include "xdg_device.h"

namespace xdgd {

enum Mode { RAYFIRE = 0, PIV = 1, OCCL = 2 };

struct RayD {
  D3 o, d, rb;
  int idx;
  double oi, di;  // origin / direction on the major axis
};

struct BoxRayD {
  double ix, iy, iz, nx, ny, nz;  // inv dir, -o * inv
};
struct BoxRayF {
  float ix, iy, iz, nx, ny, nz;
};

__device__ __forceinline__ double safe_inv_d(double d)
{
  const double eps = 1e-300;
  return 1.0 / (fabs(d) > eps ? d : copysign(eps, d));
}
__device__ __forceinline__ float safe_inv_f(float d)
{
  const float eps = 1e-30f;
  return 1.0f / (fabsf(d) > eps ? d : copysignf(eps, d));
}

__device__ __forceinline__ float f32_up(double x)
{
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}
__device__ __forceinline__ float f32_down_d(double x)
{
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}

// slab test of one box; returns entry distance in tn
__device__ __forceinline__ bool slab_d(const double* b, const BoxRayD& r, double tmax, double& tn)
{
  double x0 = fma(b[0], r.ix, r.nx), x1 = fma(b[3], r.ix, r.nx);
  double y0 = fma(b[1], r.iy, r.ny), y1 = fma(b[4], r.iy, r.ny);
  double z0 = fma(b[2], r.iz, r.nz), z1 = fma(b[5], r.iz, r.nz);
  double a = fmax(fmax(fmin(x0, x1), fmin(y0, y1)), fmax(fmin(z0, z1), 0.0));
  double c = fmin(fmin(fmax(x0, x1), fmax(y0, y1)), fmin(fmax(z0, z1), tmax));
  tn = a;
  return a <= c;
}

__device__ __forceinline__ bool slab_f(const float* b, const BoxRayF& r, float tmax, float& tn)
{
  float x0 = fmaf(b[0], r.ix, r.nx), x1 = fmaf(b[3], r.ix, r.nx);
  float y0 = fmaf(b[1], r.iy, r.ny), y1 = fmaf(b[4], r.iy, r.ny);
  float z0 = fmaf(b[2], r.iz, r.nz), z1 = fmaf(b[5], r.iz, r.nz);
  float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
  float c = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
  tn = a;
  return a <= c;
}

// Per-lane traversal stack in LDS, interleaved across the workgroup (entry i
// of lane l at base[i * stride + l], so a wave's pushes hit 64 consecutive
// words).  The top is kept as a pointer: a push or pop is one add instead of
// a 32-bit multiply.  Pushes beyond the capacity are dropped; callers size
// the capacity to the tree's proven stack bound, so none is dropped.
struct LdsStack {
  int32_t* top;
  int32_t* base;
  int32_t* end;
  int stride;
  __device__ __forceinline__ LdsStack(int32_t* b, int s, int cap) : top(b), base(b), end(b + cap * s), stride(s) {}
  __device__ __forceinline__ void reset() { top = base; }
  __device__ __forceinline__ bool empty() const { return top == base; }
  __device__ __forceinline__ void push(int32_t v)
  {
    if (top < end) {
      *top = v;
      top += stride;
    }
  }
  __device__ __forceinline__ int32_t pop()
  {
    top -= stride;
    return *top;
  }
};

struct HitState {
  double t;
  int32_t face;
  int32_t surf;
  bool inside;  // PIV: dot(dir, sense-flipped normal) > 0 of the committed hit
  int32_t leaf;  // tree position (leaf order) of the committed triangle
};

// Work counters of the instrumented traversal (roofline accounting): V node
// visits, F triangle tests, E Pluecker edge tests, C tests reaching the
// distance computation, N normal evaluations; WV / WF count the node and
// triangle iterations each wave executes (credited to its lowest active
// lane), so V / (64 WV) is the SIMD efficiency of the node loop.
struct TraceStats {
  unsigned int V, F, E, C, N, WV, WF;
};

// max of v (< 2^24) over the active lanes of the wave, in every lane
// (bitwise from the top: only ballots, so inactive lanes never contribute)
__device__ __forceinline__ unsigned int wave_max_u24(unsigned int v)
{
  const int lane = threadIdx.x & 63;
  unsigned long long cand = __ballot(1);
  unsigned int m = 0;
  for (int b = 23; b >= 0; --b) {
    const unsigned long long has = __ballot(((v >> b) & 1u) != 0u && ((cand >> lane) & 1ull) != 0ull);
    if (has) {
      m |= 1u << b;
      cand &= has;
    }
  }
  return m;
}

// 1 in the lowest active lane of the wave, else 0
__device__ __forceinline__ unsigned int first_active_lane()
{
  const unsigned long long m = __ballot(1);
  return (unsigned int)((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1);
}

// Counts the edge tests plucker_any performs (same control flow).
__device__ __forceinline__ void plucker_count(const D3& v0, const D3& v1, const D3& v2, const D3& d, const D3& rb,
                                              TraceStats& st)
{
  const double c0 = edge_test(v0, v1, d, rb);
  const double c1 = edge_test(v1, v2, d, rb);
  st.E += 2;
  if ((0.0 < c0 && 0.0 > c1) || (0.0 > c0 && 0.0 < c1)) return;
  const double c2 = edge_test(v2, v0, d, rb);
  st.E += 1;
  if ((0.0 < c1 && 0.0 > c2) || (0.0 > c1 && 0.0 < c2) || (0.0 < c0 && 0.0 > c2) || (0.0 > c0 && 0.0 < c2))
    return;
  if (0.0 == c0 && 0.0 == c1 && 0.0 == c2) return;
  st.C += 1;
}

// Leaf triangle for RAYFIRE / PIV: TriangleIntersectionFunc with the tie rule.
// The normal is read from the build-time array (same bits as recomputing
// face_normal here), so the leaf carries no sqrt/divide.
template <int MODE, bool STATS = false, class TT = SurfTreeDev>
__device__ __forceinline__ void tri_closest_hit(const TT& T, int32_t p, const RayD& ray, int orientation,
                                                const int32_t* ex, int k, HitState& h, TraceStats* st = nullptr)
{
  const auto r = as_global(T.tris + p);
  const D3 v0 = d3(r->v[0], r->v[1], r->v[2]);
  const D3 v1 = d3(r->v[3], r->v[4], r->v[5]);
  const D3 v2 = d3(r->v[6], r->v[7], r->v[8]);
  double t;
  if (STATS) {
    st->F += 1;
    plucker_count(v0, v1, v2, ray.d, ray.rb, *st);
  }
  const int32_t s2 = r->surf2;
  if (!plucker_any_f(v0, v1, v2, (unsigned)(s2 >> 1) & 7u, ray.d, ray.rb, ray.idx, ray.oi, ray.di, h.t, 0.0, t))
    return;
  const int32_t face = r->face;
  if (t == h.t && face >= h.face) return;  // tie rule
  if (STATS) st->N += 1;
  const auto pn = as_global(T.normals + 3 * (int64_t)p);
  D3 n = d3(pn[0], pn[1], pn[2]);
  if (s2 & 1) n = neg(n);
  if (MODE == RAYFIRE && orientation != ORIENT_ANY) {
    const double dp = dot(ray.d, n);
    if (orientation == ORIENT_EXITING && dp < 0.0) return;
    if (orientation == ORIENT_ENTERING && dp >= 0.0) return;
  }
  for (int j = 0; j < k; ++j)
    if (ex[j] == face) return;
  h.t = t;
  h.face = face;
  h.leaf = p;
  h.surf = s2 >> TRI_SURF_SHIFT;
  if (MODE == PIV) h.inside = dot(ray.d, n) > 0.0;
}

template <int MODE>
__device__ __forceinline__ bool tri_any_hit(const TriRec* __restrict__ tris, int32_t p, const RayD& ray)
{
  const auto r = as_global(tris + p);
  const D3 v0 = d3(r->v[0], r->v[1], r->v[2]);
  const D3 v1 = d3(r->v[3], r->v[4], r->v[5]);
  const D3 v2 = d3(r->v[6], r->v[7], r->v[8]);
  double t;
  return plucker_any_f(v0, v1, v2, (unsigned)(r->surf2 >> 1) & 7u, ray.d, ray.rb, ray.idx, ray.oi, ray.di, INFTY,
                       0.0, t);
}

// Generic traversal.  Returns true for OCCL when any hit is found.
template <int MODE, bool MIXED, bool STATS = false>
__device__ __forceinline__ bool traverse(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                         int k, int32_t* stk, int stride, int depth_cap, HitState& h,
                                         TraceStats* st = nullptr)
{
  if (T.n_prims <= 0) return false;
  BoxRayD bd;
  BoxRayF bf;
  if (MIXED) {
    float ox = (float)ray.o.x, oy = (float)ray.o.y, oz = (float)ray.o.z;
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
  // The stack holds child codes, [slot][lane] in LDS (4 bytes per entry:
  // storing entry distances as well halves the occupancy LDS allows).
  LdsStack S(stk, stride, depth_cap);
  float tf = MODE == OCCL ? INFINITY : f32_up(h.t);
  int32_t node = T.root;
  while (true) {
    if (node >= 0) {
      int32_t c0, c1;
      bool h0, h1;
      float e0, e1;
      if (STATS) st->V += 1;
      if (MIXED) {
        const NodeF nd = gload(T.nodes_f + node);
        h0 = slab_f(nd.b0, bf, tf, e0);
        h1 = slab_f(nd.b1, bf, tf, e1);
        c0 = nd.c0;
        c1 = nd.c1;
      } else {
        const NodeD nd = gload(T.nodes_d + node);
        double t0, t1;
        double tm = MODE == OCCL ? INFTY : h.t;
        h0 = slab_d(nd.b0, bd, tm, t0);
        h1 = slab_d(nd.b1, bd, tm, t1);
        e0 = f32_down_d(t0);
        e1 = f32_down_d(t1);
        c0 = nd.c0;
        c1 = nd.c1;
      }
      if (h0 && h1) {
        if (e1 < e0) {
          int32_t tmp = c0;
          c0 = c1;
          c1 = tmp;
          e1 = e0;
        }
        S.push(c1);
        node = c0;
        continue;
      } else if (h0) {
        node = c0;
        continue;
      } else if (h1) {
        node = c1;
        continue;
      }
    } else {
      const int32_t first = leaf_first(node);
      const int32_t cnt = leaf_count(node);
      for (int32_t j = 0; j < cnt; ++j) {
        if (MODE == OCCL) {
          if (tri_any_hit<MODE>(T.tris, first + j, ray)) return true;
        } else {
          tri_closest_hit<MODE, STATS>(T, first + j, ray, orientation, ex, k, h, st);
        }
      }
      if (MODE != OCCL) tf = f32_up(h.t);
    }
    if (S.empty()) break;
    node = S.pop();
  }
  return false;
}

// ---------------------------------------------------------------------------
// BVH4 traversal: one node fetch tests four child boxes; hits are ordered by
// entry distance (5-comparator network), the nearest is visited next and the
// others pushed farthest-first.
// ---------------------------------------------------------------------------

constexpr int32_t CHILD_EMPTY = INT32_MIN;  // never a valid leaf or node code

__device__ __forceinline__ void cswap(float& ta, int32_t& ca, float& tb, int32_t& cb)
{
  if (tb < ta) {
    const float t = ta;
    ta = tb;
    tb = t;
    const int32_t c = ca;
    ca = cb;
    cb = c;
  }
}

// The pointers a traversal needs, per lane (persistent kernels, where
// neighbouring lanes may trace different trees).
struct TreeView {
  const Node4D* nodes4_d;
  const Node4F* nodes4_f;
  const TriRec* tris;
  const double* normals;
};

// Byte offsets of a ray's near and far planes inside a BVH4 node: with the
// direction's sign per axis known, the entry distance of a slab is the near
// plane's and the exit distance the far plane's, so the slab test needs no
// min/max pair per axis.  An empty slot's inverted box (lo = +inf,
// hi = -inf) then gives entry +inf and can never be hit.
struct PlaneOfs {
  uint32_t nx, fx, ny, fy, nz, fz;
};
template <bool MIXED>
__device__ __forceinline__ PlaneOfs plane_ofs(bool negx, bool negy, bool negz)
{
  constexpr uint32_t P = MIXED ? 16u : 32u;  // bytes per plane (4 children)
  PlaneOfs o;
  o.nx = negx ? P : 0u;
  o.fx = P - o.nx;
  o.ny = 2 * P + (negy ? P : 0u);
  o.fy = 5 * P - o.ny;
  o.nz = 4 * P + (negz ? P : 0u);
  o.fz = 9 * P - o.nz;
  return o;
}
__device__ __forceinline__ PlaneOfs plane_ofs_f(const BoxRayF& b)
{
  return plane_ofs<true>(signbit(b.ix), signbit(b.iy), signbit(b.iz));
}
__device__ __forceinline__ PlaneOfs plane_ofs_d(const BoxRayD& b)
{
  return plane_ofs<false>(signbit(b.ix), signbit(b.iy), signbit(b.iz));
}

template <class V, class P>
__device__ __forceinline__ V gload_at(const P* base, uint32_t byte_ofs)
{
  return gload(reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_ofs));
}

template <bool MIXED, class TT = SurfTreeDev>
__device__ __forceinline__ void node4_test(const TT& T, int32_t node, const BoxRayF& bf, const BoxRayD& bd,
                                           const PlaneOfs& po, float tf, double tm, float* te, int32_t* ch)
{
  if (MIXED) {
    const Node4F* nd = T.nodes4_f + node;
    const float4 nx = gload_at<float4>(nd, po.nx), fx = gload_at<float4>(nd, po.fx);
    const float4 ny = gload_at<float4>(nd, po.ny), fy = gload_at<float4>(nd, po.fy);
    const float4 nz = gload_at<float4>(nd, po.nz), fz = gload_at<float4>(nd, po.fz);
    const int4 cc = gload_at<int4>(nd, 96u);
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    const int32_t cca[4] = {cc.x, cc.y, cc.z, cc.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float a = fmaxf(fmaxf(fmaf(nxa[c], bf.ix, bf.nx), fmaf(nya[c], bf.iy, bf.ny)),
                            fmaxf(fmaf(nza[c], bf.iz, bf.nz), 0.0f));
      const float b = fminf(fminf(fmaf(fxa[c], bf.ix, bf.nx), fmaf(fya[c], bf.iy, bf.ny)),
                            fminf(fmaf(fza[c], bf.iz, bf.nz), tf));
      const bool hit = a <= b;
      te[c] = hit ? a : INFINITY;
      ch[c] = hit ? cca[c] : CHILD_EMPTY;
    }
  } else {
    const Node4D* nd = T.nodes4_d + node;
    const double4 nx = gload_at<double4>(nd, po.nx), fx = gload_at<double4>(nd, po.fx);
    const double4 ny = gload_at<double4>(nd, po.ny), fy = gload_at<double4>(nd, po.fy);
    const double4 nz = gload_at<double4>(nd, po.nz), fz = gload_at<double4>(nd, po.fz);
    const int4 cc = gload_at<int4>(nd, 192u);
    const double nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const double nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const double nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    const int32_t cca[4] = {cc.x, cc.y, cc.z, cc.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double a = fmax(fmax(fma(nxa[c], bd.ix, bd.nx), fma(nya[c], bd.iy, bd.ny)),
                            fmax(fma(nza[c], bd.iz, bd.nz), 0.0));
      const double b = fmin(fmin(fma(fxa[c], bd.ix, bd.nx), fma(fya[c], bd.iy, bd.ny)),
                            fmin(fma(fza[c], bd.iz, bd.nz), tm));
      const bool hit = a <= b;
      te[c] = hit ? f32_down_d(a) : INFINITY;
      ch[c] = hit ? cca[c] : CHILD_EMPTY;
    }
  }
}

template <int MODE, bool MIXED, bool STATS, bool SPEC, class STK>
__device__ __forceinline__ bool traverse4_s(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                            int k, STK& S, HitState& h, TraceStats* st = nullptr)
{
  if (T.n_prims <= 0) return false;
  BoxRayD bd;
  BoxRayF bf;
  if (MIXED) {
    float ox = (float)ray.o.x, oy = (float)ray.o.y, oz = (float)ray.o.z;
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
  const PlaneOfs po = MIXED ? plane_ofs_f(bf) : plane_ofs_d(bd);
  float tf = MODE == OCCL ? INFINITY : f32_up(h.t);
  int32_t node = T.root4;
  S.reset();
  // while-while: lanes descend internal nodes together until every lane of
  // the wave holds a leaf (or is done), then the leaves are tested together;
  // node and leaf work never share a divergent iteration.  SPEC: a lane that
  // reaches a leaf while others still descend parks it and keeps descending
  // (with the distance bound it has), so fewer lanes idle in the node loop.
  const int32_t DONE = CHILD_EMPTY;
  int32_t leaf = DONE;
  while (node != DONE) {
    while (node >= 0) {
      if (STATS) {
        st->V += 1;
        st->WV += first_active_lane();
      }
      float te[4];
      int32_t ch[4];
      node4_test<MIXED>(T, node, bf, bd, po, tf, MODE == OCCL ? INFTY : h.t, te, ch);
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
      if (SPEC) {
        if (node < 0 && node != DONE && leaf == DONE) {
          leaf = node;
          node = S.empty() ? DONE : S.pop();
        }
        if (__all(leaf != DONE)) break;
      }
    }
    if (leaf == DONE) {
      if (node == DONE) break;
      leaf = node;  // a leaf: node < 0 here
      node = S.empty() ? DONE : S.pop();
    }
    const int32_t first = leaf_first(leaf);
    const int32_t cnt = leaf_count(leaf);
    leaf = DONE;
    for (int32_t j = 0; j < cnt; ++j) {
      if (STATS) st->WF += first_active_lane();
      if (MODE == OCCL) {
        if (tri_any_hit<MODE>(T.tris, first + j, ray)) return true;
      } else {
        tri_closest_hit<MODE, STATS>(T, first + j, ray, orientation, ex, k, h, st);
      }
    }
    if (MODE != OCCL) tf = f32_up(h.t);
  }
  return false;
}

template <int MODE, bool MIXED, bool STATS = false, bool SPEC = false>
__device__ __forceinline__ bool traverse4(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                          int k, int32_t* stk, int stride, int depth_cap, HitState& h,
                                          TraceStats* st = nullptr)
{
  LdsStack S(stk, stride, depth_cap);
  return traverse4_s<MODE, MIXED, STATS, SPEC>(T, ray, orientation, ex, k, S, h, st);
}

// Stack with its first entries in LDS and the rest in a per-lane global
// region (entry i >= lds_cap of lane l at glb[(i - lds_cap) * gstride + l]).
// The LDS share bounds occupancy; the global share keeps deep trees exact.
struct HybridStack {
  int32_t* lt;      // next free LDS slot
  int32_t* lbase;
  int32_t* gt;      // next free global slot
  int32_t* gbase;
  int lstride;
  int64_t gstride;
  int n, lcap, cap;
  __device__ __forceinline__ void reset()
  {
    lt = lbase;
    gt = gbase;
    n = 0;
  }
  __device__ __forceinline__ bool empty() const { return n == 0; }
  __device__ __forceinline__ void push(int32_t v)
  {
    if (n >= cap) return;
    if (n < lcap) {
      *lt = v;
      lt += lstride;
    } else {
      *gt = v;
      gt += gstride;
    }
    ++n;
  }
  __device__ __forceinline__ int32_t pop()
  {
    --n;
    if (n < lcap) {
      lt -= lstride;
      return *lt;
    }
    gt -= gstride;
    return *gt;
  }
};

__device__ __forceinline__ RayD make_ray(const double* o, const double* d)
{
  RayD r;
  r.o = d3(o[0], o[1], o[2]);
  r.d = d3(d[0], d[1], d[2]);
  r.rb = cross(r.d, r.o);  // plucker.h:75
  r.idx = major_axis(r.d);
  r.oi = o[r.idx];
  r.di = d[r.idx];
  return r;
}

__device__ __forceinline__ bool mixed_safe(const SurfTreeDev& T, const RayD& r)
{
  double m = fmax(fmax(fabs(r.o.x), fabs(r.o.y)), fabs(r.o.z));
  return m <= (double)T.o_limit;
}

// MIXED traces with FP32 boxes when the origin is within the tree's limit
// and falls back to FP64 boxes otherwise; WIDE selects the BVH4.
template <int MODE, bool MIXED, bool STATS = false, bool WIDE = false, bool SPEC = false>
__device__ __forceinline__ bool trace(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                      int k, int32_t* stk, int stride, int cap, HitState& h,
                                      TraceStats* st = nullptr)
{
  if (WIDE) {
    if (MIXED && mixed_safe(T, ray))
      return traverse4<MODE, true, STATS, SPEC>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
    return traverse4<MODE, false, STATS, SPEC>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
  }
  if (MIXED) {
    if (mixed_safe(T, ray)) return traverse<MODE, true, STATS>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
  }
  return traverse<MODE, false, STATS>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
}

// Mixed-only traversal for rays already known to be within the FP32 limit.
template <int MODE, bool STATS = false, bool WIDE = false, bool SPEC = false>
__device__ __forceinline__ bool trace_f32(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                          int k, int32_t* stk, int stride, int cap, HitState& h,
                                          TraceStats* st = nullptr)
{
  if (WIDE) return traverse4<MODE, true, STATS, SPEC>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
  return traverse<MODE, true, STATS>(T, ray, orientation, ex, k, stk, stride, cap, h, st);
}

}  // namespace xdgd
