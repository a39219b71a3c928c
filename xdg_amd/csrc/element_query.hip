// element_query.hip -- tetrahedral element location and the element walk.
//
// Reference semantics:
//   find_element   EmbreeRayTracer::find_element src/embree/ray_tracer.cpp:226-257
//                  + TetrahedronOcclusionFunc src/tetrahedron_contain.cpp:94-114
//                  (first containing tet in Embree order; here: the lowest
//                  element index whose containment test passes)
//   next_element   MeshManager::next_element src/mesh_manager_interface.cpp:150-200
//   walk_elements  src/mesh_manager_interface.cpp:107-136
//   segments       XDG::segments src/xdg.cpp:142-198 (including the
//                  implicit-complement re-entry branch :164-185)
//
// One segment per lane.  The walker reads one 128-byte ElemRec per step
// (vertices + adjacency) and evaluates the four oriented Pluecker face tests
// exactly as next_element does.  Per-element track lengths are accumulated
// as fixed-point integers (len * 2^40, round-to-nearest-even) with 64-bit
// atomics, so the tally is independent of the order in which lanes add.
#include "internal.h"
#include "surface_traverse.h"

#include <type_traits>

namespace xdgh {

using namespace xdgd;

namespace {

constexpr int MAX_WALK_STEPS = 1 << 16;
constexpr int MAX_SEG_LOOPS = 64;

__device__ __forceinline__ float f32_dn(double x)
{
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}

__device__ __forceinline__ bool in_box_f(const float* b, float plo_x, float plo_y, float plo_z, float phi_x,
                                         float phi_y, float phi_z)
{
  // conservative: p in [lo, hi] (FP64) implies round_up(p) >= lo_f and
  // round_down(p) <= hi_f for outward-rounded lo_f <= lo, hi_f >= hi
  return phi_x >= b[0] && phi_y >= b[1] && phi_z >= b[2] && plo_x <= b[3] && plo_y <= b[4] && plo_z <= b[5];
}

__device__ __forceinline__ bool elem_contains(const ElemRec& e, const D3& p)
{
  return tet_contains(p, d3(e.v[0], e.v[1], e.v[2]), d3(e.v[3], e.v[4], e.v[5]), d3(e.v[6], e.v[7], e.v[8]),
                      d3(e.v[9], e.v[10], e.v[11]));
}

// lowest containing element index, -1 if none
__device__ int32_t find_element_dev(const ElemTreeDev& T, const D3& p, int32_t* stk, int stride, int cap)
{
  if (T.n_prims <= 0) return ID_NONE;
  const float lx = f32_dn(p.x), ly = f32_dn(p.y), lz = f32_dn(p.z);
  const float hx = f32_up(p.x), hy = f32_up(p.y), hz = f32_up(p.z);
  int32_t best = INT32_MAX;
  int32_t node = T.root;
  int sp = 0;
  while (true) {
    if (node >= 0) {
      const NodeF nd = T.nodes_f[node];
      const bool h0 = in_box_f(nd.b0, lx, ly, lz, hx, hy, hz);
      const bool h1 = in_box_f(nd.b1, lx, ly, lz, hx, hy, hz);
      if (h0 && h1) {
        if (sp < cap) stk[sp++ * stride] = nd.c1;
        node = nd.c0;
      } else if (h0) {
        node = nd.c0;
      } else if (h1) {
        node = nd.c1;
      } else {
        if (sp == 0) break;
        node = stk[--sp * stride];
      }
      continue;
    }
    const int32_t first = leaf_first(node);
    const int32_t cnt = leaf_count(node);
    for (int32_t j = 0; j < cnt; ++j) {
      const int32_t e = T.order[first + j];
      if (e >= best) continue;
      if (elem_contains(T.elems[e], p)) best = e;
    }
    if (sp == 0) break;
    node = stk[--sp * stride];
  }
  return best == INT32_MAX ? ID_NONE : best;
}

// Tet face orders (local vertex triples).  FO_MOAB: MOAB canonical
// (include/xdg/moab/direct_access.h:181-183); FO_MOCK: MeshMock
// (tests/mesh_mock.h:108-115); FO_TABLE: any other order, read from the tree.
enum FaceOrder { FO_MOAB = 0, FO_MOCK = 1, FO_TABLE = 2 };

template <int FO>
__device__ __forceinline__ int fo_at(const ElemTreeDev& T, int f, int j)
{
  constexpr int moab[12] = {0, 1, 3, 1, 2, 3, 2, 0, 3, 0, 2, 1};
  constexpr int mock[12] = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};
  if (FO == FO_MOAB) return moab[3 * f + j];
  if (FO == FO_MOCK) return mock[3 * f + j];
  return T.face_order[3 * f + j];
}

template <int FO>
__device__ __forceinline__ D3 face_vertex(const ElemTreeDev& T, const ElemRec& E, const D3* v, int f, int j)
{
  if (FO == FO_TABLE) {
    const int k = T.face_order[3 * f + j];
    return d3(E.v[3 * k], E.v[3 * k + 1], E.v[3 * k + 2]);
  }
  return v[fo_at<FO>(T, f, j)];
}

// Edge value of a face edge (i -> j) from the value of the vertex pair
// (min, max): plucker_edge_test orients every edge by the lexicographic
// order of its endpoints (plucker.h:35-62), so the reversed edge gives
// exactly the negated value, except that the reference keeps a value
// snapped to zero at +0.0 where the negation here gives -0.0.  No result
// depends on that sign: the face tests compare the values with 0 (equal
// for both zeros), and a zero coordinate reaches the distance only through
// products and sums whose zero results end in the (0 < t) ? t : 0 clamp.
template <int I, int J>
__device__ __forceinline__ double pair_edge(const double* e6)
{
  constexpr int lo = I < J ? I : J, hi = I < J ? J : I;
  constexpr int k = lo == 0 ? hi - 1 : (lo == 1 ? hi + 1 : 5);  // (0,1)(0,2)(0,3)(1,2)(1,3)(2,3)
  return I < J ? e6[k] : -e6[k];
}

// Pluecker coordinates of face F (vertices A, B, C of the face order) from
// the six canonical edge values, and whether the oriented test
// (plucker.h:64-161 with orientation 1) gets past its sign and coplanar
// rejections for that face.
template <int FO, int F>
__device__ __forceinline__ bool face_coords(const double* e6, double& c0, double& c1, double& c2)
{
  constexpr int moab[12] = {0, 1, 3, 1, 2, 3, 2, 0, 3, 0, 2, 1};
  constexpr int mock[12] = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};
  constexpr int A = FO == FO_MOAB ? moab[3 * F] : mock[3 * F];
  constexpr int B = FO == FO_MOAB ? moab[3 * F + 1] : mock[3 * F + 1];
  constexpr int C = FO == FO_MOAB ? moab[3 * F + 2] : mock[3 * F + 2];
  c0 = pair_edge<A, B>(e6);
  c1 = pair_edge<B, C>(e6);
  c2 = pair_edge<C, A>(e6);
  return !(1.0 * c0 > 0) && !(1.0 * c1 > 0) && !(1.0 * c2 > 0) && !(0.0 == c0 && 0.0 == c1 && 0.0 == c2);
}

template <int FO, int F, int J>
__device__ __forceinline__ double face_axis_vertex(const double* vi)
{
  constexpr int moab[12] = {0, 1, 3, 1, 2, 3, 2, 0, 3, 0, 2, 1};
  constexpr int mock[12] = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};
  return vi[FO == FO_MOAB ? moab[3 * F + J] : mock[3 * F + J]];
}

// x[f] for f in 0..3 by two levels of selects on the bits of f (a chain of
// f == k tests is turned into a switch over scratch memory by the compiler)
__device__ __forceinline__ double sel4(int f, double x0, double x1, double x2, double x3)
{
  const bool lo = (f & 1) != 0;
  const double a = lo ? x1 : x0;
  const double b = lo ? x3 : x2;
  return (f & 2) != 0 ? b : a;
}

template <int FO, int J>
__device__ __forceinline__ double pick_face_vertex(const double* vi, int f)
{
  return sel4(f, face_axis_vertex<FO, 0, J>(vi), face_axis_vertex<FO, 1, J>(vi), face_axis_vertex<FO, 2, J>(vi),
              face_axis_vertex<FO, 3, J>(vi));
}

// One element record in registers: its vertices, neighbours and edge
// orders (unused for FO_TABLE, whose step reads the record itself).
struct ElemLoad {
  D3 v[4];
  int32_t adj[4];
  uint32_t ef;
  int32_t elem;
};

template <int FO>
__device__ __forceinline__ void elem_load(const ElemTreeDev& T, int32_t elem, ElemLoad& L)
{
  L.elem = elem;
  if (FO == FO_TABLE) return;
  const auto E = as_global(T.elems + elem);
#pragma unroll
  for (int j = 0; j < 4; ++j) L.v[j] = d3(E->v[3 * j], E->v[3 * j + 1], E->v[3 * j + 2]);
#pragma unroll
  for (int j = 0; j < 4; ++j) L.adj[j] = E->adj[j];
  L.ef = E->edge_first;
}

template <int FO>
__device__ __forceinline__ void next_element_dev(const ElemTreeDev& T, int32_t elem, const D3& r, const D3& u,
                                                 int idx, double u_i, int32_t& next, double& dist);

// MeshManager::next_element (mesh_manager_interface.cpp:150-200) on a
// loaded record.  With a fixed face order the six edges of the tet are
// evaluated once each (every face edge is one of them, up to its sign),
// instead of up to twelve times with per-face early exits that make
// neighbouring lanes diverge.
template <int FO>
__device__ __forceinline__ void elem_step(const ElemTreeDev& T, const ElemLoad& L, const D3& r, const D3& u, int idx,
                                          double u_i, int32_t& next, double& dist)
{
  if (FO == FO_TABLE) {
    next_element_dev<FO_TABLE>(T, L.elem, r, u, idx, u_i, next, dist);
    return;
  }
  const D3* v = L.v;
  const D3 rb = cross(u, r);
  const double r_i = pick3(r.x, r.y, r.z, idx);
  const uint32_t ef = L.ef;
  const double e6[6] = {edge_test_f(v[0], v[1], ef & 1u, u, rb),  edge_test_f(v[0], v[2], ef & 2u, u, rb),
                        edge_test_f(v[0], v[3], ef & 4u, u, rb),  edge_test_f(v[1], v[2], ef & 8u, u, rb),
                        edge_test_f(v[1], v[3], ef & 16u, u, rb), edge_test_f(v[2], v[3], ef & 32u, u, rb)};
  double c0[4], c1[4], c2[4];
  unsigned pass = 0;
  pass |= face_coords<FO, 0>(e6, c0[0], c1[0], c2[0]) ? 1u : 0u;
  pass |= face_coords<FO, 1>(e6, c0[1], c1[1], c2[1]) ? 2u : 0u;
  pass |= face_coords<FO, 2>(e6, c0[2], c1[2], c2[2]) ? 4u : 0u;
  pass |= face_coords<FO, 3>(e6, c0[3], c1[3], c2[3]) ? 8u : 0u;
  double vi[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) vi[j] = pick3(v[j].x, v[j].y, v[j].z, idx);
  // A ray in general position passes the sign test on its exit face only,
  // so the distance (two divisions) is formed once per face that passes,
  // in face order; a failing face scores INFTY, which never replaces the
  // running minimum, so the first strict minimum is the reference's.
  int idx_out = ID_NONE;
  double min_dist = INFTY;
  while (pass) {
    const int f = __builtin_ctz(pass);
    pass &= pass - 1;
    const double a0 = sel4(f, c0[0], c0[1], c0[2], c0[3]);
    const double a1 = sel4(f, c1[0], c1[1], c1[2], c1[3]);
    const double a2 = sel4(f, c2[0], c2[1], c2[2], c2[3]);
    const double inverse_sum = 1.0 / (a0 + a1 + a2);
    const double w0 = a0 * inverse_sum;
    const double w1 = a1 * inverse_sum;
    const double w2 = a2 * inverse_sum;
    const double ix = w0 * pick_face_vertex<FO, 2>(vi, f) + w1 * pick_face_vertex<FO, 0>(vi, f) +
                      w2 * pick_face_vertex<FO, 1>(vi, f);
    double t = (ix - r_i) / u_i;
    if (w2 < 0.0 || w0 < 0.0 || (w2 + w0) > 1.0) t = -1.0;
    if (!(t < 0.0 || t > INFTY)) {
      const double df = (0.0 < t) ? t : 0.0;
      if (df < min_dist) {
        min_dist = df;
        idx_out = f;
      }
    }
  }
  // neighbour by select from the loaded record: no dependent load
  const int lo = idx_out & 1;
  const int32_t na = lo ? L.adj[1] : L.adj[0];
  const int32_t nb = lo ? L.adj[3] : L.adj[2];
  next = idx_out == ID_NONE ? ID_NONE : ((idx_out & 2) ? nb : na);
  dist = min_dist;
}

template <int FO>
__device__ __forceinline__ void next_element_dev(const ElemTreeDev& T, int32_t elem, const D3& r, const D3& u,
                                                 int idx, double u_i, int32_t& next, double& dist)
{
  if (FO != FO_TABLE) {
    ElemLoad L;
    elem_load<FO>(T, elem, L);
    elem_step<FO>(T, L, r, u, idx, u_i, next, dist);
    return;
  }
  const ElemRec& E = T.elems[elem];
  D3 v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = d3(E.v[3 * j], E.v[3 * j + 1], E.v[3 * j + 2]);
  const D3 rb = cross(u, r);
  const double r_i = pick3(r.x, r.y, r.z, idx);
  int idx_out = ID_NONE;
  double min_dist = INFTY;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const D3 a = face_vertex<FO>(T, E, v, f, 0);
    const D3 b = face_vertex<FO>(T, E, v, f, 1);
    const D3 c = face_vertex<FO>(T, E, v, f, 2);
    double t;
    double df = INFTY;
    if (plucker_oriented(a, b, c, u, rb, idx, r_i, u_i, 1.0, INFTY, 0.0, t)) df = (0.0 < t) ? t : 0.0;
    if (df < min_dist) {
      min_dist = df;
      idx_out = f;
    }
  }
  next = idx_out == ID_NONE ? ID_NONE : E.adj[idx_out];
  dist = min_dist;
}

__global__ void k_find_element(FindElementArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int32_t out = ID_NONE;
  if (a.tree >= 1 && a.tree <= a.nelem_trees) {
    const D3 p = d3(a.pts[3 * i], a.pts[3 * i + 1], a.pts[3 * i + 2]);
    out = find_element_dev(a.etable[a.tree], p, lds + threadIdx.x, blockDim.x, cap);
  }
  a.elem_out[i] = out;
}

__global__ void k_next_element(NextElementArgs a)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int32_t nx = ID_NONE;
  double d = INFTY;
  if (a.tree >= 1 && a.tree <= a.nelem_trees) {
    const ElemTreeDev& T = a.etable[a.tree];
    const int32_t e = a.elem[i];
    if (e >= 0 && e < T.n_prims) {
      const D3 r = d3(a.r[3 * i], a.r[3 * i + 1], a.r[3 * i + 2]);
      const D3 u = d3(a.u[3 * i], a.u[3 * i + 1], a.u[3 * i + 2]);
      const int idx = major_axis(u);
      next_element_dev<FO_TABLE>(T, e, r, u, idx, pick3(u.x, u.y, u.z, idx), nx, d);
    }
  }
  a.next_out[i] = nx;
  a.dist_out[i] = d;
}

struct Sink {
  int32_t count;
  unsigned long long hash;
  int64_t cap;
  int32_t* se;
  double* sl;
  unsigned long long* tally;
};

// What a walk writes besides counts and hashes.  OUT_TALLY: per-element
// tallies, and the tally atomic is unconditional, so the compiler knows it
// follows the walk's prefetch of the next record and waits for that record
// without waiting for the atomic.  OUT_SEQ: element sequences, plus
// tallies when s.tally is non-null.
enum WalkOut { OUT_NONE = 0, OUT_TALLY = 1, OUT_SEQ = 2 };

template <int OUT>
__device__ __forceinline__ void sink_push(Sink& s, int32_t e, double len)
{
  if (OUT == OUT_SEQ && s.count < s.cap) {
    s.se[s.count] = e;
    s.sl[s.count] = len;
  }
  s.count++;
  s.hash = mix64(s.hash ^ (unsigned long long)(uint32_t)e);
  s.hash = mix64(s.hash ^ (unsigned long long)__double_as_longlong(len));
  if (OUT == OUT_TALLY || (OUT == OUT_SEQ && s.tally)) {
    long long q = __double2ll_rn(len * 1099511627776.0);  // len * 2^40, exact scaling
    atomicAdd(s.tally + e, (unsigned long long)q);
  }
}

// MeshManager::walk_elements(elem, start, u, distance)
// (mesh_manager_interface.cpp:107-136) from element cur, fused with the
// sink; returns the std::accumulate sum of the lengths (xdg.cpp:190-191).
// `last` is the last element pushed; `steps` counts steps against
// MAX_WALK_STEPS over the whole segment.
template <int FO, int OUT>
__device__ __forceinline__ double walk_dev(const ElemTreeDev& T, int32_t cur, const D3& r, const D3& u, int idx,
                                           double u_i, double dist_w, Sink& s, int32_t& last, int& steps)
{
  D3 rr = r;
  double sum = 0.0;
  ElemLoad L;
  elem_load<FO>(T, cur, L);  // cur is an element here
  // Wait for the record here, so that at the loop head the only loads in
  // flight are the ones issued by the previous step, which the tally atomic
  // follows: the head then waits for the loads alone (one operation left
  // outstanding) instead of for everything.
  __builtin_amdgcn_s_waitcnt(0);
  while (dist_w > 0) {
    int32_t next;
    double ex;
    elem_step<FO>(T, L, rr, u, idx, u_i, next, ex);
    const double len = (dist_w < ex) ? dist_w : ex;
    dist_w -= len;
    const int32_t here = cur;
    last = cur;
    sum = sum + len;
    rr.x = rr.x + len * u.x;
    rr.y = rr.y + len * u.y;
    rr.z = rr.z + len * u.z;
    cur = next;
    const bool stop = cur == ID_NONE || ++steps >= MAX_WALK_STEPS;
    // the next record is requested before this step's tally and sequence
    // writes, so waiting for it does not wait for them; a lane that stops
    // requests its current record again (unused, always in range) so that
    // the loads are not under a branch
    elem_load<FO>(T, (!stop && dist_w > 0) ? cur : here, L);
    sink_push<OUT>(s, here, len);
    if (stop) break;
  }
  return sum;
}

template <int FO, int OUT>
__global__ void k_segments(SegmentsArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.order) i = a.order[i];
  Sink s;
  s.count = 0;
  s.hash = HASH_SEED;
  s.tally = a.tally_fx;
  if (OUT == OUT_SEQ) {
    s.se = a.seq_elem + a.seq_off[i];
    s.sl = a.seq_len + a.seq_off[i];
    s.cap = a.seq_off[i + 1] - a.seq_off[i];
  } else {
    s.se = nullptr;
    s.sl = nullptr;
    s.cap = 0;
  }
  if (a.tree < 1 || a.tree > a.nelem_trees) {
    a.count[i] = 0;
    a.hash[i] = s.hash;
    return;
  }
  // a copy read before any store, so it stays in scalar registers; the loop
  // below writes memory, after which fields read through a reference would
  // be fetched again (and waited for) on every step
  const ElemTreeDev T = a.etable[a.tree];
  int32_t* stk = lds + threadIdx.x;
  const int stride = blockDim.x;
  // xdg.cpp:150-153
  D3 r = d3(a.p0[3 * i], a.p0[3 * i + 1], a.p0[3 * i + 2]);
  D3 u = sub(d3(a.p1[3 * i], a.p1[3 * i + 1], a.p1[3 * i + 2]), r);
  double distance = sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
  u.x = u.x / distance;
  u.y = u.y / distance;
  u.z = u.z / distance;
  const int idx = major_axis(u);
  const double u_i = pick3(u.x, u.y, u.z, idx);
  int32_t last = ID_NONE;
  int steps = 0, loops = 0;
  bool capped = false;
  while (distance > 0) {
    if (++loops > MAX_SEG_LOOPS || steps >= MAX_WALK_STEPS) {
      capped = true;
      break;
    }
    int32_t cur = find_element_dev(T, r, stk, stride, cap);
    if (s.count > 0 && cur == last) cur = ID_NONE;
    if (cur == ID_NONE) {
      // xdg.cpp:164-185: fire against the implicit complement
      if (T.ipc_tree < 1 || T.ipc_tree > a.nsurf_trees) break;
      RayD ray;
      ray.o = d3(r.x + TINY_BIT * u.x, r.y + TINY_BIT * u.y, r.z + TINY_BIT * u.z);
      ray.d = u;
      ray.rb = cross(ray.d, ray.o);
      ray.idx = idx;
      ray.oi = pick3(ray.o.x, ray.o.y, ray.o.z, idx);
      ray.di = u_i;
      HitState h;
      h.t = INFTY;
      h.face = INT32_MAX;
      h.surf = ID_NONE;
      h.inside = false;
      trace<RAYFIRE, false>(a.stable[T.ipc_tree], ray, ORIENT_EXITING, nullptr, 0, stk, stride, cap, h);
      if (h.surf == ID_NONE || h.t > distance) break;
      const double hit_dist = h.t + TINY_BIT;
      r.x = r.x + hit_dist * u.x;
      r.y = r.y + hit_dist * u.y;
      r.z = r.z + hit_dist * u.z;
      distance -= hit_dist;
      const int32_t adjacent = T.bface_elem ? T.bface_elem[h.face] : ID_NONE;
      if (adjacent == ID_NONE) break;
      cur = adjacent;
    }
    // walk_elements (mesh_manager_interface.cpp:107-136) from cur
    const double sum = walk_dev<FO, OUT>(T, cur, r, u, idx, u_i, distance, s, last, steps);
    r.x = r.x + sum * u.x;
    r.y = r.y + sum * u.y;
    r.z = r.z + sum * u.z;
    distance -= sum;
  }
  a.count[i] = s.count;
  a.hash[i] = s.hash;
  if (capped && a.status) atomicOr(a.status, 1u);
}

// XDG::segments(volume, start, end) (xdg.cpp:200-227) over the volume's
// element tree, fused with the same sink as k_segments.  A start outside
// the mesh enters it through the volume's surface (ENTERING ray_fire, no
// exclusions); the walk then runs MeshManager::walk_elements(elem, start,
// end) (mesh_manager_interface.cpp:139-147), which takes its direction and
// length again from the moved start.
template <int FO, int OUT>
__global__ void k_segments_vol(SegmentsArgs a, int cap)
{
  extern __shared__ int32_t lds[];
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.order) i = a.order[i];
  Sink s;
  s.count = 0;
  s.hash = HASH_SEED;
  s.tally = a.tally_fx;
  if (OUT == OUT_SEQ) {
    s.se = a.seq_elem + a.seq_off[i];
    s.sl = a.seq_len + a.seq_off[i];
    s.cap = a.seq_off[i + 1] - a.seq_off[i];
  } else {
    s.se = nullptr;
    s.sl = nullptr;
    s.cap = 0;
  }
  if (a.tree < 1 || a.tree > a.nelem_trees) {
    a.count[i] = 0;
    a.hash[i] = s.hash;
    return;
  }
  const ElemTreeDev T = a.etable[a.tree];  // before any store: scalar registers
  int32_t* stk = lds + threadIdx.x;
  const int stride = blockDim.x;
  const D3 p0 = d3(a.p0[3 * i], a.p0[3 * i + 1], a.p0[3 * i + 2]);
  const D3 p1 = d3(a.p1[3 * i], a.p1[3 * i + 1], a.p1[3 * i + 2]);
  // xdg.cpp:205-207
  D3 start = p0;
  const D3 u = normalize(sub(p1, p0));
  int32_t e = find_element_dev(T, p0, stk, stride, cap);
  bool ok = true;
  uint32_t flags = 0;
  if (e == ID_NONE) {
    // xdg.cpp:211-222: the first entering hit on the volume's surface
    ok = a.surf_tree >= 1 && a.surf_tree <= a.nsurf_trees;
    if (ok) {
      RayD ray;
      ray.o = p0;
      ray.d = u;
      ray.rb = cross(u, p0);
      ray.idx = major_axis(u);
      ray.oi = pick3(p0.x, p0.y, p0.z, ray.idx);
      ray.di = pick3(u.x, u.y, u.z, ray.idx);
      HitState h{INFTY, INT32_MAX, ID_NONE, false};
      trace<RAYFIRE, false>(a.stable[a.surf_tree], ray, ORIENT_ENTERING, nullptr, 0, stk, stride, cap, h);
      ok = h.surf != ID_NONE;
      if (ok) {
        const double into = h.t + TINY_BIT;
        e = find_element_dev(T, d3(p0.x + u.x * into, p0.y + u.y * into, p0.z + u.z * into), stk, stride, cap);
        if (e == ID_NONE) {
          ok = false;
          flags |= 2u;  // the reference warns: surface hit, no element behind it
        } else {
          start = d3(start.x + u.x * h.t, start.y + u.y * h.t, start.z + u.z * h.t);
        }
      }
    }
  }
  if (ok) {
    // MeshManager::walk_elements(elem, start, end)
    const D3 w0 = sub(p1, start);
    const double distance = sqrt(w0.x * w0.x + w0.y * w0.y + w0.z * w0.z);
    const D3 w = normalize(w0);
    const int idx = major_axis(w);
    int32_t last = ID_NONE;
    int steps = 0;
    walk_dev<FO, OUT>(T, e, start, w, idx, pick3(w.x, w.y, w.z, idx), distance, s, last, steps);
    if (steps >= MAX_WALK_STEPS) flags |= 1u;
  }
  a.count[i] = s.count;
  a.hash[i] = s.hash;
  if (flags && a.status) atomicOr(a.status, flags);
}

// tools/walk_elements.h:53-117 (walk_elements): one particle per lane.
// Start: uniform samples in [lo, hi] until find_element finds an element;
// isotropic direction, normalized.  Each event: next_element, an
// exponential collision distance; a collision scatters isotropically, else
// the particle moves to the next element; a particle that leaves the mesh
// fires at the implicit complement (EXITING, the event's hit faces
// excluded) and re-enters at the element behind the hit.  Philox draws
// replace drand48 / rand_dir (the oracle replays the same streams).
enum : uint8_t { MW_ALIVE = 0, MW_LOST = 3, MW_LEFT = 2, MW_CAPPED = 5, MW_ERROR = 6 };
constexpr int MW_PRIMS = 8;        // re-entry exclusions kept per event
constexpr int MW_STARTS = 1000;    // start samples tried per particle
constexpr int MW_REENTRIES = 64;   // re-entry rays per event

// One particle's state between events.
struct MwState {
  D3 r, u;
  int32_t e;
  double dist;
  int32_t nev;
  unsigned long long hh;
  uint32_t dc;
  uint8_t st;
  uint64_t item;
};

// Particle::initialize of the tool: sample a location within the model
// (walk_elements.h:64-71), an isotropic direction, normalized.
__device__ __forceinline__ void mw_start(const MeshWalkArgs& a, const ElemTreeDev& T, int64_t i, MwState& S,
                                         int32_t* stk, int cap)
{
  S.item = a.first + (uint64_t)i;
  S.dc = 0;
  S.st = MW_ALIVE;
  S.e = ID_NONE;
  S.r = d3(0.0, 0.0, 0.0);
  S.u = d3(0.0, 0.0, 0.0);
  S.dist = 0.0;
  S.nev = 0;
  S.hh = HASH_SEED;
  for (int att = 0; att < MW_STARTS && S.e == ID_NONE; ++att) {
    double ua, ub, uc, unused;
    uniform2(a.seed, a.stream, S.item, S.dc++, ua, ub);
    uniform2(a.seed, a.stream, S.item, S.dc++, uc, unused);
    S.r = d3(a.lo[0] + (a.hi[0] - a.lo[0]) * ua, a.lo[1] + (a.hi[1] - a.lo[1]) * ub,
             a.lo[2] + (a.hi[2] - a.lo[2]) * uc);
    S.e = find_element_dev(T, S.r, stk, blockDim.x, cap);
  }
  if (S.e == ID_NONE) {
    S.st = MW_ERROR;
    return;
  }
  if (!iso_dir(a.seed, a.stream, S.item, S.dc, S.u))
    S.st = MW_ERROR;
  else
    S.u = normalize(S.u);
}

// The part of an event before any re-entry (walk_elements.h:73-90):
// next_element, the collision draw, the move, the tally.  L holds the
// record of S.e, requested by the previous step (or before the loop); the
// collision distance is drawn while it is in flight, and the record of the
// element the particle is in afterwards is requested before the tally
// atomic, so the next step waits for that record alone, not for the
// atomic as well.
template <int FO>
__device__ __forceinline__ void mw_step(const MeshWalkArgs& a, const ElemTreeDev& T, MwState& S, ElemLoad& L)
{
  const int idx = major_axis(S.u);
  double ua, ub;
  uniform2(a.seed, a.stream, S.item, S.dc++, ua, ub);
  const double coll = -log_series(1.0 - ua) * a.mfp;
  int32_t next;
  double ex;
  elem_step<FO>(T, L, S.r, S.u, idx, pick3(S.u.x, S.u.y, S.u.z, idx), next, ex);
  const int32_t here = S.e;
  const double s = coll < ex ? coll : ex;
  S.r = d3(S.r.x + S.u.x * s, S.r.y + S.u.y * s, S.r.z + S.u.z * s);
  S.dist += s;
  if (coll < ex) {
    if (!iso_dir(a.seed, a.stream, S.item, S.dc, S.u)) S.st = MW_ERROR;
  } else {
    S.e = next;
  }
  // a particle that left the mesh requests its last record again (unused,
  // always in range), so the load is not under a branch
  elem_load<FO>(T, S.e != ID_NONE ? S.e : here, L);
  S.hh = mix64(S.hh ^ (unsigned long long)(uint32_t)here);
  S.hh = mix64(S.hh ^ (unsigned long long)__double_as_longlong(s));
  if (a.tally_fx) atomicAdd(a.tally_fx + here, (unsigned long long)__double2ll_rn(s * 1099511627776.0));
}

// Re-entry through the implicit complement (walk_elements.h:92-102) of a
// particle that left the mesh; prims holds the event's hit faces.
__device__ __forceinline__ void mw_reenter(const MeshWalkArgs& a, const ElemTreeDev& T, MwState& S, int32_t* stk,
                                           int32_t* prims, int cap)
{
  int tries = 0, np = 0;
  while (S.st == MW_ALIVE && S.e == ID_NONE) {
    if (++tries > MW_REENTRIES) {
      S.st = MW_LOST;
      break;
    }
    RayD ray;
    ray.o = S.r;
    ray.d = S.u;
    ray.rb = cross(S.u, S.r);
    ray.idx = major_axis(S.u);
    ray.oi = pick3(S.r.x, S.r.y, S.r.z, ray.idx);
    ray.di = pick3(S.u.x, S.u.y, S.u.z, ray.idx);
    HitState h{INFTY, INT32_MAX, ID_NONE, false};
    trace<RAYFIRE, false>(a.stable[T.ipc_tree], ray, ORIENT_EXITING, prims, np, stk, blockDim.x, cap, h);
    if (h.surf == ID_NONE) break;
    // the hit face joins the exclusions (ray_tracer.cpp:324)
    if (np == MW_PRIMS) {
      for (int j = 1; j < MW_PRIMS; ++j) prims[j - 1] = prims[j];
      --np;
    }
    prims[np++] = h.face;
    S.r = d3(S.r.x + S.u.x * h.t, S.r.y + S.u.y * h.t, S.r.z + S.u.z * h.t);
    S.dist += h.t;
    S.e = find_element_dev(
        T, d3(S.r.x + S.u.x * TINY_BIT, S.r.y + S.u.y * TINY_BIT, S.r.z + S.u.z * TINY_BIT), stk, blockDim.x, cap);
  }
}

// The end of an event (walk_elements.h:103-104): count it; the event cap;
// a particle outside the mesh after its re-entry attempts has left.
__device__ __forceinline__ void mw_close(const MeshWalkArgs& a, MwState& S)
{
  ++S.nev;
  if (S.st == MW_ALIVE && S.e != ID_NONE && S.nev >= a.max_events) S.st = MW_CAPPED;
  if (S.st == MW_ALIVE && S.e == ID_NONE) S.st = MW_LEFT;
}

__device__ __forceinline__ void mw_finish(const MeshWalkArgs& a, int64_t i, const MwState& S)
{
  if (a.distance) a.distance[i] = S.dist;
  if (a.events) a.events[i] = S.nev;
  if (a.status) a.status[i] = S.st;
  if (a.hash) a.hash[i] = S.hh;
}

__device__ __forceinline__ void mw_store(MwSave& m, const MwState& S)
{
  m.r[0] = S.r.x;
  m.r[1] = S.r.y;
  m.r[2] = S.r.z;
  m.u[0] = S.u.x;
  m.u[1] = S.u.y;
  m.u[2] = S.u.z;
  m.dist = S.dist;
  m.hh = S.hh;
  m.e = S.e;
  m.nev = S.nev;
  m.dc = S.dc;
}

__device__ __forceinline__ void mw_load(const MwSave& m, uint64_t item, MwState& S)
{
  S.r = d3(m.r[0], m.r[1], m.r[2]);
  S.u = d3(m.u[0], m.u[1], m.u[2]);
  S.dist = m.dist;
  S.hh = m.hh;
  S.e = m.e;
  S.nev = m.nev;
  S.dc = m.dc;
  S.st = MW_ALIVE;
  S.item = item;
}

// Walk kernel of the wavefront: particles from in_list (every particle,
// started here, when in_list is null) run events until they end.  A
// particle that leaves the mesh is parked in out_list for the batched
// re-entry kernel; with out_list null it re-enters inline.  Tracing the
// re-entry rays of a whole list together keeps that long, rare traversal
// from running with one active lane while the rest of the wave waits.
// At most 256 lanes per workgroup, so the inline variant (which carries
// the re-entry traversal too) may hold more than 128 VGPRs instead of
// spilling.
template <int FO, bool INLINE>
__global__ __launch_bounds__(256) void k_mw_walk(MeshWalkArgs a, int cap, MwSave* save, const int32_t* in_list,
                                                 const unsigned int* in_count, int32_t* out_list,
                                                 unsigned int* out_count)
{
  extern __shared__ int32_t lds[];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n_in = in_count ? (int64_t)*in_count : a.n;
  if (j >= n_in) return;
  const int64_t i = in_list ? (int64_t)in_list[j] : j;
  int32_t* stk = lds + threadIdx.x;
  int32_t* prims = lds + (size_t)cap * blockDim.x + MW_PRIMS * threadIdx.x;
  const ElemTreeDev T = a.etable[a.tree];
  const bool has_ipc = T.ipc_tree >= 1 && T.ipc_tree <= a.nsurf_trees;
  MwState S;
  if (in_list)
    mw_load(save[i], a.first + (uint64_t)i, S);
  else
    mw_start(a, T, i, S, stk, cap);
  ElemLoad L;
  if (S.st == MW_ALIVE) elem_load<FO>(T, S.e, L);
  while (S.st == MW_ALIVE) {
    mw_step<FO>(a, T, S, L);
    if (S.st == MW_ALIVE && S.e == ID_NONE && has_ipc) {
      if (!INLINE) {
        mw_store(save[i], S);
        out_list[atomicAdd(out_count, 1u)] = (int32_t)i;
        return;
      }
      mw_reenter(a, T, S, stk, prims, cap);
      if (S.e != ID_NONE) elem_load<FO>(T, S.e, L);
    }
    mw_close(a, S);
  }
  mw_finish(a, i, S);
}

// The parking walk on a resident grid: a lane whose particle ended or was
// parked waits until `refill` lanes of its wave (or all remaining ones) are
// idle; those lanes then claim the next particles of the list together
// (one atomic on *claim) and start them (sampling and find_element, or the
// saved state), while the other lanes keep stepping.  A wave therefore no
// longer runs until its longest particle ends, and the long start path runs
// for a batch of lanes at a time.  Events are those of k_mw_walk.
// The ballots of an iteration are turned into each lane's claim position
// inside the wave-uniform claim block, and the position is pinned to a
// vector register there.  The compiler may run the stepping lanes in an
// inner loop that the claiming lanes leave early; a wave-uniform (scalar)
// mask read after that loop would hold a later iteration's value, whereas a
// lane's vector registers keep its own.
// Parked particles are listed by claim position: out_list[pos] holds the
// particle of position pos if it parked, else -1, and *out_count is set to
// the number of positions.  A shared append counter would take one atomic
// with return per parking iteration, and those serialise on one address.
template <int FO>
__global__ __launch_bounds__(256) void k_mw_walk_pt(MeshWalkArgs a, int cap, MwSave* save, const int32_t* in_list,
                                                    const unsigned int* in_count, int32_t* out_list,
                                                    unsigned int* out_count, unsigned long long* claim, int refill)
{
  extern __shared__ int32_t lds[];
  const int64_t n_in = in_count ? (int64_t)*in_count : a.n;
  int32_t* stk = lds + threadIdx.x;
  const ElemTreeDev T = a.etable[a.tree];
  const bool has_ipc = T.ipc_tree >= 1 && T.ipc_tree <= a.nsurf_trees;
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = (unsigned int)n_in;
  bool have = false;
  int64_t i = 0, cpos = 0;
  MwState S;
  S.st = MW_ALIVE;
  ElemLoad L;
  while (true) {
    const bool idle = !have;
    const unsigned long long imask = __ballot(idle);
    const int nact = __popcll(__ballot(1));
    const bool go = __popcll(imask) >= (refill < nact ? refill : nact);
    const unsigned long long mask = go ? imask : 0ull;
    int64_t pos = -1;  // this lane's claimed position, -1 when it claims none
    if (mask) {        // wave-uniform: every lane of the wave takes part
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long b = 0;
      if (lane == leader) b = atomicAdd(claim, (unsigned long long)__popcll(mask));
      const unsigned long long base = __shfl(b, leader);
      if (idle) pos = (int64_t)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
    }
    asm volatile("" : "+v"(pos));
    if (pos >= n_in) break;
    if (pos >= 0) {
      cpos = pos;
      i = in_list ? (int64_t)in_list[pos] : pos;
      if (in_list)
        mw_load(save[i], a.first + (uint64_t)i, S);
      else
        mw_start(a, T, i, S, stk, cap);
      if (S.st == MW_ALIVE) {
        elem_load<FO>(T, S.e, L);
        have = true;
      } else {
        mw_finish(a, i, S);
        out_list[cpos] = -1;
      }
    }
    if (have) {
      mw_step<FO>(a, T, S, L);
      if (S.st == MW_ALIVE && S.e == ID_NONE && has_ipc) {
        mw_store(save[i], S);
        out_list[cpos] = (int32_t)i;
        have = false;
      } else {
        mw_close(a, S);
        if (S.st != MW_ALIVE) {
          mw_finish(a, i, S);
          out_list[cpos] = -1;
          have = false;
        }
      }
    }
  }
}

// Re-entry kernel of the wavefront: the parked particles' re-entry rays;
// particles back in the mesh go on in out_list.  Negative entries of
// in_list are positions of the resident walk whose particle did not park.
__global__ void k_mw_reenter(MeshWalkArgs a, int cap, MwSave* save, const int32_t* in_list,
                             const unsigned int* in_count, int32_t* out_list, unsigned int* out_count)
{
  extern __shared__ int32_t lds[];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= (int64_t)*in_count) return;
  const int64_t i = in_list[j];
  if (i < 0) return;
  int32_t* stk = lds + threadIdx.x;
  int32_t* prims = lds + (size_t)cap * blockDim.x + MW_PRIMS * threadIdx.x;
  const ElemTreeDev T = a.etable[a.tree];
  MwState S;
  mw_load(save[i], a.first + (uint64_t)i, S);
  mw_reenter(a, T, S, stk, prims, cap);
  mw_close(a, S);
  if (S.st == MW_ALIVE) {
    mw_store(save[i], S);
    out_list[atomicAdd(out_count, 1u)] = (int32_t)i;
  } else {
    mw_finish(a, i, S);
  }
}

__global__ void k_zero_counters(unsigned int* p, int n)
{
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

__global__ void k_elem_boxes(const ElemRec* __restrict__ elems, int64_t n, double* box6)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ElemRec& e = elems[i];
  // BoundingBox::from_points + dilation (bbox.h:85-102): chord * 1e-6
  double lo[3] = {INFTY, INFTY, INFTY}, hi[3] = {-INFTY, -INFTY, -INFTY};
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], e.v[3 * j + k]);
      hi[k] = fmax(hi[k], e.v[3 * j + k]);
    }
  const double wx = hi[0] - lo[0], wy = hi[1] - lo[1], wz = hi[2] - lo[2];
  const double bump = sqrt(wx * wx + wy * wy + wz * wz) * 1e-6;
  for (int k = 0; k < 3; ++k) {
    box6[6 * i + k] = lo[k] - bump;
    box6[6 * i + 3 + k] = hi[k] + bump;
  }
}

inline dim3 grid1(int64_t n, int blk) { return dim3((unsigned)((n + blk - 1) / blk)); }
inline size_t lds_bytes(const LaunchCfg& c) { return sizeof(int32_t) * (size_t)c.stack * c.block; }

}  // namespace

void launch_find_element(const FindElementArgs& a, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_find_element, grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

template <int FO>
void launch_mesh_walk_t(const MeshWalkArgs& a, const LaunchCfg& c, const MeshWalkScratch* w, hipStream_t s)
{
  const int block = c.block > 256 ? 256 : c.block;  // k_mw_walk's launch bound
  const size_t sh = sizeof(int32_t) * ((size_t)c.stack + MW_PRIMS) * block;
  const dim3 g = grid1(a.n, block), b(block);
  if (!w) {  // everything inline, one particle per lane
    hipLaunchKernelGGL((k_mw_walk<FO, true>), g, b, sh, s, a, c.stack, nullptr, nullptr, nullptr, nullptr, nullptr);
    return;
  }
  // rounds of walk -> batched re-entry; the last walk re-enters inline, so
  // every particle ends whatever the number of re-entries
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, w->count, MW_COUNT_WORDS);
  const int32_t* in_list = nullptr;
  const unsigned int* in_count = nullptr;
  const size_t shp = sizeof(int32_t) * (size_t)c.stack * block;  // no re-entry primitives in the resident walk
  const int64_t gpt = c.mw_refill ? std::min<int64_t>(g.x, resident_grid(k_mw_walk_pt<FO>, block, shp)) : 0;
  for (int r = 0; r < MW_ROUNDS; ++r) {
    if (c.mw_refill)
      hipLaunchKernelGGL(k_mw_walk_pt<FO>, dim3((unsigned)std::max<int64_t>(gpt, 1)), b, shp, s, a, c.stack, w->save,
                         in_list, in_count, w->list[0], w->count + 2 * r, w->claim + r, c.mw_refill);
    else
      hipLaunchKernelGGL((k_mw_walk<FO, false>), g, b, sh, s, a, c.stack, w->save, in_list, in_count, w->list[0],
                         w->count + 2 * r);
    hipLaunchKernelGGL(k_mw_reenter, g, b, sh, s, a, c.stack, w->save, (const int32_t*)w->list[0],
                       (const unsigned int*)(w->count + 2 * r), w->list[1], w->count + 2 * r + 1);
    // the next walk reads list 1 and parks into list 0 again; the stream
    // orders each list's reads before it is rewritten
    in_list = w->list[1];
    in_count = w->count + 2 * r + 1;
  }
  hipLaunchKernelGGL((k_mw_walk<FO, true>), g, b, sh, s, a, c.stack, w->save, in_list, in_count, nullptr, nullptr);
}

void launch_mesh_walk(const MeshWalkArgs& a, int face_order_kind, const LaunchCfg& c, const MeshWalkScratch* w,
                      hipStream_t s)
{
  if (a.n <= 0) return;
  if (face_order_kind == FO_MOAB)
    launch_mesh_walk_t<FO_MOAB>(a, c, w, s);
  else if (face_order_kind == FO_MOCK)
    launch_mesh_walk_t<FO_MOCK>(a, c, w, s);
  else
    launch_mesh_walk_t<FO_TABLE>(a, c, w, s);
}

void launch_next_element(const NextElementArgs& a, hipStream_t s)
{
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_next_element, grid1(a.n, 256), dim3(256), 0, s, a);
}

template <int FO, int OUT>
void launch_segments_t(const SegmentsArgs& a, const LaunchCfg& c, hipStream_t s)
{
  if (a.volume_mode)
    hipLaunchKernelGGL((k_segments_vol<FO, OUT>), grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
  else
    hipLaunchKernelGGL((k_segments<FO, OUT>), grid1(a.n, c.block), dim3(c.block), lds_bytes(c), s, a, c.stack);
}

namespace {

__device__ __forceinline__ uint32_t spread10(uint32_t v)  // 10 bits -> every third bit of 30
{
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t q10(double x, float lo, float hi)
{
  const double w = (double)hi - (double)lo;
  double q = w > 0.0 ? (x - (double)lo) / w * 1024.0 : 0.0;
  q = q >= 0.0 ? (q > 1023.0 ? 1023.0 : q) : 0.0;  // NaN -> 0
  return (uint32_t)q;
}

// mode 1: Morton code of the start point (10 bits per axis); mode 2: the
// start cell and then the end cell (5 bits per axis each), so that
// neighbouring segments share most of their path
__global__ void k_segment_keys(SegmentsArgs a, int mode, uint32_t* keys, int32_t* idx)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  uint32_t m = 0;
  if (a.tree >= 1 && a.tree <= a.nelem_trees) {
    const ElemTreeDev& T = a.etable[a.tree];
    const uint32_t qs[3] = {q10(a.p0[3 * i], T.box_lo[0], T.box_hi[0]), q10(a.p0[3 * i + 1], T.box_lo[1], T.box_hi[1]),
                            q10(a.p0[3 * i + 2], T.box_lo[2], T.box_hi[2])};
    if (mode == 2) {
      const uint32_t qe[3] = {q10(a.p1[3 * i], T.box_lo[0], T.box_hi[0]),
                              q10(a.p1[3 * i + 1], T.box_lo[1], T.box_hi[1]),
                              q10(a.p1[3 * i + 2], T.box_lo[2], T.box_hi[2])};
      const uint32_t ms = spread10(qs[0] >> 5) | (spread10(qs[1] >> 5) << 1) | (spread10(qs[2] >> 5) << 2);
      const uint32_t me = spread10(qe[0] >> 5) | (spread10(qe[1] >> 5) << 1) | (spread10(qe[2] >> 5) << 2);
      m = (ms << 15) | me;
    } else {
      m = spread10(qs[0]) | (spread10(qs[1]) << 1) | (spread10(qs[2]) << 2);
    }
  }
  keys[i] = m;
  idx[i] = (int32_t)i;
}

}  // namespace

const int32_t* launch_segment_order(const SegmentsArgs& a, int mode, const RayOrderScratch& w, hipStream_t s)
{
  if (a.n <= 0) return nullptr;
  hipLaunchKernelGGL(k_segment_keys, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a, mode, w.keys_in,
                     w.idx_in);
  radix_sort_pairs(w.keys_in, w.idx_in, w.keys_out, w.idx_out, a.n, 30, (uint32_t*)w.temp, s);
  return w.idx_in;
}

void launch_segments(const SegmentsArgs& a, int face_order_kind, const LaunchCfg& c, hipStream_t s)
{
  if (a.n <= 0) return;
  const int out = a.seq_off ? OUT_SEQ : (a.tally_fx ? OUT_TALLY : OUT_NONE);
  auto go = [&](auto fo) {
    constexpr int FO = decltype(fo)::value;
    if (out == OUT_SEQ)
      launch_segments_t<FO, OUT_SEQ>(a, c, s);
    else if (out == OUT_TALLY)
      launch_segments_t<FO, OUT_TALLY>(a, c, s);
    else
      launch_segments_t<FO, OUT_NONE>(a, c, s);
  };
  if (face_order_kind == FO_MOAB)
    go(std::integral_constant<int, FO_MOAB>{});
  else if (face_order_kind == FO_MOCK)
    go(std::integral_constant<int, FO_MOCK>{});
  else
    go(std::integral_constant<int, FO_TABLE>{});
}

int face_order_kind(const int32_t* fo)
{
  static const int32_t moab[12] = {0, 1, 3, 1, 2, 3, 2, 0, 3, 0, 2, 1};
  static const int32_t mock[12] = {0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2};
  bool is_moab = true, is_mock = true;
  for (int i = 0; i < 12; ++i) {
    is_moab = is_moab && fo[i] == moab[i];
    is_mock = is_mock && fo[i] == mock[i];
  }
  return is_moab ? FO_MOAB : (is_mock ? FO_MOCK : FO_TABLE);
}

void launch_elem_boxes(const ElemRec* elems, int64_t n, double* box6, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_elem_boxes, grid1(n, 256), dim3(256), 0, s, elems, n, box6);
}

}  // namespace xdgh
