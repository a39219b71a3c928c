// xdg_device.h -- device-side geometry primitives and data layouts.
//
// Every FP64 expression keeps the reference's operation order (the whole
// library is compiled with -ffp-contract=off, so no FMA is formed implicitly)
// and therefore rounds bit-identically to the reference C++ code:
//   Vec3da dot/cross/normalize   include/xdg/vec3da.h:58-90
//   plucker_edge_test            include/xdg/geometry/plucker.h:45-62
//   plucker_ray_tri_intersect    include/xdg/geometry/plucker.h:64-161
//   face_normal                  src/mesh_manager_interface.cpp:216-220
//   closest_location_on_triangle src/geometry/closest.cpp:6-183
//   tetrahedron containment      src/tetrahedron_contain.cpp:11-43
// Explicit fma()/fmaf() appear only in conservative box tests, whose result
// decides which candidates are examined, never a returned value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xdgd {

// Tree data is read through the global address space: a pointer fetched
// from a table in memory is otherwise generic, and generic (flat) loads also
// count against the LDS counter, so each wait for a node would wait for the
// LDS stack traffic as well.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p)
{
  return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ T gload(const T* p)
{
  T v;
  __builtin_memcpy(&v, (const void*)as_global(p), sizeof(T));
  return v;
}

constexpr double INFTY = 1.7976931348623157e308;  // constants.h:15 (DBL_MAX)
constexpr double PLUCKER_ZERO_TOL = 20.0 * 2.220446049250313080847e-16;
constexpr double TINY_BIT = 1e-10;                // constants.h TINY_BIT
// constants.h:80: float(FLT_MIN * 1E5), promoted to double when compared
constexpr float MIN_RCP_INPUT = 0x1p-126f * 1E5;
constexpr int ID_NONE = -1;

// HitOrientation (include/xdg/shared_enums.h:11-15)
constexpr int ORIENT_ANY = -1;
constexpr int ORIENT_EXITING = 0;
constexpr int ORIENT_ENTERING = 1;

struct D3 {
  double x, y, z;
};

__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 sub(const D3& a, const D3& b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(const D3& a, const D3& b)
{
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double pick3(double x, double y, double z, int i)
{
  return i == 0 ? x : (i == 1 ? y : z);
}
__device__ __forceinline__ D3 neg(const D3& a) { return D3{-a.x, -a.y, -a.z}; }

// Vec3da::normalize (vec3da.h:61-66)
__device__ __forceinline__ D3 normalize(D3 a)
{
  double len = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  len = len < (double)MIN_RCP_INPUT ? (double)MIN_RCP_INPUT : len;
  a.x /= len;
  a.y /= len;
  a.z /= len;
  return a;
}

// first() plucker.h:35-43
// lexicographic (x, y, z) order, plucker.h:35-43 (written without branches:
// neighbouring lanes test different edges)
__device__ __forceinline__ bool first(const D3& a, const D3& b)
{
  const bool lx = a.x < b.x, gx = a.x > b.x, ly = a.y < b.y, gy = a.y > b.y, lz = a.z < b.z;
  return lx || (!gx && (ly || (!gy && lz)));
}

// plucker_edge_test plucker.h:45-62.  The reference evaluates
//   f:  dot(ray, cross(b - a, a)) + dot(rn, b - a)
//   !f: -(dot(ray, cross(a - b, b)) + dot(rn, a - b))
// with f = first(a, b).  a - b is exactly -(b - a) and negation passes
// exactly through every product and sum of the second form, so it equals
// dot(ray, cross(b - a, b)) + dot(rn, b - a) bit for bit (up to the sign of
// a zero, which the tolerance snap turns into +0.0 either way).  One
// expression with the endpoint chosen by select therefore serves both
// orders: a wave whose lanes disagree on the order runs no second branch.
// edge_test_f takes first(a, b) precomputed (element records carry it).
__device__ __forceinline__ double edge_test_f(const D3& va, const D3& vb, bool f, const D3& ray, const D3& rn)
{
  const D3 p = f ? va : vb;
  const D3 e = sub(vb, va);
  const D3 en = cross(e, p);
  double pip = dot(ray, en) + dot(rn, e);
  if (PLUCKER_ZERO_TOL > fabs(pip)) pip = 0.0;
  return pip;
}
__device__ __forceinline__ double edge_test(const D3& va, const D3& vb, const D3& ray, const D3& rn)
{
  return edge_test_f(va, vb, first(va, vb), ray, rn);
}

// largest-|dir| axis, first wins ties (plucker.h:137-144)
__device__ __forceinline__ int major_axis(const D3& d)
{
  int idx = 0;
  double m = 0.0;
  if (fabs(d.x) > m) { idx = 0; m = fabs(d.x); }
  if (fabs(d.y) > m) { idx = 1; m = fabs(d.y); }
  if (fabs(d.z) > m) { idx = 2; }
  return idx;
}

// Distance from the three Pluecker coordinates (plucker.h:123-160).
// Returns false for the coplanar / barycentric / range rejections.  o_i and
// d_i are the ray's origin and direction on the major axis idx.  Values on
// that axis are chosen by select, never by a dynamic index (which would
// force the vertices into scratch memory).
__device__ __forceinline__ bool plucker_finish(double c0, double c1, double c2, const D3& v0, const D3& v1,
                                               const D3& v2, double o_i, double d_i, int idx, double tmax,
                                               double tmin, double& t)
{
  if (0.0 == c0 && 0.0 == c1 && 0.0 == c2) return false;
  const double inverse_sum = 1.0 / (c0 + c1 + c2);
  const double w0 = c0 * inverse_sum;
  const double w1 = c1 * inverse_sum;
  const double w2 = c2 * inverse_sum;
  // only the major-axis component of the intersection point is used, so
  // the vertices' components on that axis are selected first
  const double ix = w0 * pick3(v2.x, v2.y, v2.z, idx) + w1 * pick3(v0.x, v0.y, v0.z, idx) +
                    w2 * pick3(v1.x, v1.y, v1.z, idx);
  double dist = (ix - o_i) / d_i;
  if (w2 < 0.0 || w0 < 0.0 || (w2 + w0) > 1.0) dist = -1.0;
  if (dist < tmin || dist > tmax) return false;
  t = dist;
  return true;
}

// plucker_ray_tri_intersect, unoriented mode (useOrientation = false), with
// the ray moment rb = cross(d, o), the major axis and its components
// precomputed per ray.
// first_bits: bit 0 first(v0, v1), bit 1 first(v1, v2), bit 2 first(v2, v0).
__device__ __forceinline__ bool plucker_any_f(const D3& v0, const D3& v1, const D3& v2, unsigned first_bits,
                                              const D3& d, const D3& rb, int idx, double o_i, double d_i,
                                              double tmax, double tmin, double& t)
{
  const double c0 = edge_test_f(v0, v1, first_bits & 1u, d, rb);
  const double c1 = edge_test_f(v1, v2, first_bits & 2u, d, rb);
  if ((0.0 < c0 && 0.0 > c1) || (0.0 > c0 && 0.0 < c1)) return false;
  const double c2 = edge_test_f(v2, v0, first_bits & 4u, d, rb);
  if ((0.0 < c1 && 0.0 > c2) || (0.0 > c1 && 0.0 < c2) || (0.0 < c0 && 0.0 > c2) || (0.0 > c0 && 0.0 < c2))
    return false;
  return plucker_finish(c0, c1, c2, v0, v1, v2, o_i, d_i, idx, tmax, tmin, t);
}

__device__ __forceinline__ unsigned tri_first_bits(const D3& v0, const D3& v1, const D3& v2)
{
  return (first(v0, v1) ? 1u : 0u) | (first(v1, v2) ? 2u : 0u) | (first(v2, v0) ? 4u : 0u);
}

__device__ __forceinline__ bool plucker_any(const D3& v0, const D3& v1, const D3& v2, const D3& d, const D3& rb,
                                            int idx, double o_i, double d_i, double tmax, double tmin, double& t)
{
  const double c0 = edge_test(v0, v1, d, rb);
  const double c1 = edge_test(v1, v2, d, rb);
  if ((0.0 < c0 && 0.0 > c1) || (0.0 > c0 && 0.0 < c1)) return false;
  const double c2 = edge_test(v2, v0, d, rb);
  if ((0.0 < c1 && 0.0 > c2) || (0.0 > c1 && 0.0 < c2) || (0.0 < c0 && 0.0 > c2) || (0.0 > c0 && 0.0 < c2))
    return false;
  return plucker_finish(c0, c1, c2, v0, v1, v2, o_i, d_i, idx, tmax, tmin, t);
}

// plucker_ray_tri_intersect, oriented mode (useOrientation = true)
__device__ __forceinline__ bool plucker_oriented(const D3& v0, const D3& v1, const D3& v2, const D3& d,
                                                 const D3& rb, int idx, double o_i, double d_i, double ori,
                                                 double tmax, double tmin, double& t)
{
  const double c0 = edge_test(v0, v1, d, rb);
  if (ori * c0 > 0) return false;
  const double c1 = edge_test(v1, v2, d, rb);
  if (ori * c1 > 0) return false;
  const double c2 = edge_test(v2, v0, d, rb);
  if (ori * c2 > 0) return false;
  return plucker_finish(c0, c1, c2, v0, v1, v2, o_i, d_i, idx, tmax, tmin, t);
}

// MeshManager::face_normal (mesh_manager_interface.cpp:216-220)
__device__ __forceinline__ D3 face_normal(const D3& v0, const D3& v1, const D3& v2)
{
  return normalize(cross(sub(v1, v0), sub(v2, v0)));
}

// closest_location_on_triangle (src/geometry/closest.cpp:54-183)
__device__ __forceinline__ D3 axpy(const D3& a, double f, const D3& b)
{
  return D3{a.x + f * b.x, a.y + f * b.y, a.z + f * b.z};
}
__device__ __forceinline__ D3 axmy(const D3& a, double f, const D3& b)
{
  return D3{a.x - f * b.x, a.y - f * b.y, a.z - f * b.z};
}
__device__ __forceinline__ D3 lin2(double f, const D3& a, double g, const D3& b)
{
  return D3{f * a.x + g * b.x, f * a.y + g * b.y, f * a.z + g * b.z};
}

__device__ inline D3 closest_on_triangle(const D3& v0, const D3& v1, const D3& v2, const D3& p)
{
  const D3 sv = sub(v1, v0);
  const D3 tv = sub(v2, v0);
  const D3 pv = sub(v0, p);
  const double ss = dot(sv, sv);
  const double st = dot(sv, tv);
  const double tt = dot(tv, tv);
  const double sp = dot(sv, pv);
  const double tp = dot(tv, pv);
  const double det = ss * tt - st * st;
  double s = st * tp - tt * sp;
  double t = st * sp - ss * tp;
  if (s + t < det) {
    if (s < 0) {
      if (t < 0) {  // region 4
        if (sp < 0) {
          if (-sp > ss) return v1;
          return axmy(v0, (sp / ss), sv);
        } else if (tp < 0) {
          if (-tp > tt) return v2;
          return axmy(v0, (tp / tt), tv);
        }
        return v0;
      }
      // region 3
      if (tp >= 0) return v0;
      if (-tp >= tt) return v2;
      return axmy(v0, (tp / tt), tv);
    } else if (t < 0) {  // region 5
      if (sp >= 0.0) return v0;
      if (-sp >= ss) return v1;
      return axmy(v0, (sp / ss), sv);
    }
    // region 0
    const double inv_det = 1.0 / det;
    s *= inv_det;
    t *= inv_det;
    return axpy(axpy(v0, s, sv), t, tv);
  }
  if (s < 0) {  // region 2
    s = st + sp;
    t = tt + tp;
    if (t > s) {
      const double num = t - s;
      const double den = ss - 2 * st + tt;
      if (num > den) return v1;
      const double q = num / den;
      return lin2(q, v1, (1 - q), v2);
    } else if (t <= 0) {
      return v2;
    } else if (tp >= 0) {
      return v0;
    }
    return axmy(v0, (tp / tt), tv);
  } else if (t < 0) {  // region 6
    t = st + tp;
    s = ss + sp;
    if (s > t) {
      const double num = t - s;
      const double den = tt - 2 * st + ss;
      if (num > den) return v2;
      t = num / den;
      s = 1 - t;
      return lin2(s, v1, t, v2);
    } else if (s <= 0) {
      return v1;
    } else if (sp >= 0) {
      return v0;
    }
    return axmy(v0, (sp / ss), sv);
  }
  // region 1
  const double num = tt + tp - st - sp;
  if (num <= 0) return v2;
  const double den = ss - 2 * st + tt;
  if (num >= den) return v1;
  const double q = num / den;
  return lin2(q, v1, (1 - q), v2);
}

// plucker_tet_containment_test (src/tetrahedron_contain.cpp:11-43):
// lambda123 = inverse(T) * (p - v0), T = [v1-v0, v2-v0, v3-v0] (columns);
// inverse = adjugate / det (element-wise); lambda0 = 1 - (l1 + l2 + l3).
__device__ __forceinline__ bool tet_contains(const D3& p, const D3& v0, const D3& v1, const D3& v2, const D3& v3)
{
  const D3 c0 = sub(v1, v0);
  const D3 c1 = sub(v2, v0);
  const D3 c2 = sub(v3, v0);
  const D3 r = sub(p, v0);
  const D3 a0 = D3{c1.y * c2.z - c2.y * c1.z, c2.y * c0.z - c0.y * c2.z, c0.y * c1.z - c1.y * c0.z};
  const D3 a1 = D3{c1.z * c2.x - c2.z * c1.x, c2.z * c0.x - c0.z * c2.x, c0.z * c1.x - c1.z * c0.x};
  const D3 a2 = D3{c1.x * c2.y - c2.x * c1.y, c2.x * c0.y - c0.x * c2.y, c0.x * c1.y - c1.x * c0.y};
  const double det = c0.x * a0.x + c0.y * a1.x + c0.z * a2.x;
  const D3 i0 = D3{a0.x / det, a0.y / det, a0.z / det};
  const D3 i1 = D3{a1.x / det, a1.y / det, a1.z / det};
  const D3 i2 = D3{a2.x / det, a2.y / det, a2.z / det};
  const double l1 = i0.x * r.x + i1.x * r.y + i2.x * r.z;
  const double l2 = i0.y * r.x + i1.y * r.y + i2.y * r.z;
  const double l3 = i0.z * r.x + i1.z * r.y + i2.z * r.z;
  const double l0 = 1.0 - (l1 + l2 + l3);
  const double hi = 1.0 + PLUCKER_ZERO_TOL;
  if (l0 < -PLUCKER_ZERO_TOL || l0 > hi) return false;
  if (l1 < -PLUCKER_ZERO_TOL || l1 > hi) return false;
  if (l2 < -PLUCKER_ZERO_TOL || l2 > hi) return false;
  if (l3 < -PLUCKER_ZERO_TOL || l3 > hi) return false;
  return true;
}

// ---------------------------------------------------------------------------
// data layouts in HBM
// ---------------------------------------------------------------------------

// BVH2 node with both children's boxes inline; one node fetch yields both
// child boxes.  Child encoding: c >= 0 internal node index; c < 0 leaf,
// ~c = (first << 3) | (count - 1) over the leaf-ordered primitive array.
// Per-item sequence hash (element walks, particle histories): h starts at
// HASH_SEED; each entry (id, length) does h = mix64(h ^ (uint32)id), then
// h = mix64(h ^ bits(length)); mix64 is the splitmix64 finaliser.
constexpr unsigned long long HASH_SEED = 0x6a09e667f3bcc909ULL;

__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

// ---------------------------------------------------------------------------
// Random numbers and the natural log used by the transport sampler.
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): counter (item_lo,
// item_hi, draw, 0), key (seed, stream); one block gives two 53-bit
// uniforms (w >> 11) * 2^-53 from the words (x0<<32|x1), (x2<<32|x3).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void uniform2(uint32_t seed, uint32_t stream, uint64_t item, uint32_t draw, double& ua,
                                         double& ub)
{
  uint32_t c0 = (uint32_t)item, c1 = (uint32_t)(item >> 32), c2 = draw, c3 = 0u;
  uint32_t k0 = seed, k1 = stream;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  const uint64_t a = ((uint64_t)c0 << 32) | c1;
  const uint64_t b = ((uint64_t)c2 << 32) | c3;
  ua = (double)(a >> 11) * 0x1.0p-53;
  ub = (double)(b >> 11) * 0x1.0p-53;
}

// Marsaglia (1972) isotropic direction as tools/ray_benchmark.h:24-40: one
// Philox block per attempt.
__device__ __forceinline__ bool iso_dir(uint32_t seed, uint32_t stream, uint64_t item, uint32_t& dc, D3& u)
{
  for (int a = 0; a < 1000; ++a) {
    double ua, ub;
    uniform2(seed, stream, item, dc++, ua, ub);
    const double x1 = ua * 2.0 - 1.0;
    const double x2 = ub * 2.0 - 1.0;
    const double s = x1 * x1 + x2 * x2;
    if (s > 0.0 && s < 1.0) {
      const double t = 2.0 * sqrt(1.0 - s);
      u = d3(x1 * t, x2 * t, 1.0 - 2.0 * s);
      return true;
    }
  }
  return false;
}

// ln(x) for positive x: x = m 2^e with m in (sqrt(1/2), sqrt(2)];
// ln(m) = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, atanh series to
// s^23 (Horner, plain IEEE ops); ln 2 split into a head with 32 significant
// bits (e * head is exact) and a tail.  Written out so the host sampler
// produces the same bits.
__device__ __forceinline__ double log_series(double x)
{
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
  if (x == __builtin_huge_val()) return x;
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff);
  if (e == 0) {
    x = x * 0x1.0p54;
    b = (uint64_t)__double_as_longlong(x);
    e = (int)((b >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = 1.0 / 21.0 + z * p;
  p = 1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = 1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = 1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = 1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = 1.0 / 3.0 + z * p;
  const double two_s = s + s;
  const double lnm = two_s + two_s * (z * p);
  const double de = (double)e;
  return de * 0x1.62e42fefa0000p-1 + (de * 0x1.cf79abc9e3b3ap-40 + lnm);
}

struct alignas(16) NodeD {  // plain FP64 layout, 112 B
  double b0[6];             // child 0: lo.x lo.y lo.z hi.x hi.y hi.z
  double b1[6];             // child 1
  int32_t c0, c1, pad0, pad1;
};

struct alignas(16) NodeF {  // mixed layout, 64 B: outward-rounded FP32 boxes
  float b0[6];
  float b1[6];
  int32_t c0, c1, pad0, pad1;
};

// 4-wide nodes (BVH4, collapsed from the binary LBVH): planes stored
// structure-of-arrays so one node gives all four child boxes.  Empty child
// slots hold an inverted box (lo = +inf, hi = -inf) that no ray can hit.
// Padded to whole 128-byte cache lines: one line per mixed node, two per
// plain node.
struct alignas(128) Node4F {  // mixed layout, 112 B used, 128 B stride
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t pad[4];
};
struct alignas(128) Node4D {  // plain FP64 layout, 208 B used, 256 B stride
  double lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t pad[12];
};

// Triangle record in leaf order, 80 B: FP64 vertices + ids.
struct alignas(16) TriRec {
  double v[9];
  int32_t face;   // mesh face id (the primitive id returned by queries)
  int32_t surf2;  // (surface id << 4) | first(v2, v0) << 3 | first(v1, v2) << 2 |
                  // first(v0, v1) << 1 | reverse-sense flag for this tree
};
constexpr int TRI_SURF_SHIFT = 4;

// Tetrahedron record indexed by element index, 128 B.
struct alignas(16) ElemRec {
  double v[12];         // vertices in connectivity order
  int32_t adj[4];       // neighbour across face i, -1 on the boundary
  uint32_t edge_first;  // bit k: first(v_a, v_b) for the k-th vertex pair (a, b) of
                        // (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) (plucker.h:35-43)
  int32_t conn[3];      // global ids of vertices 0..2 (diagnostics only)
};
static_assert(sizeof(ElemRec) == 128, "one element record per 128-byte cache line");

__host__ __device__ __forceinline__ int32_t leaf_code(int32_t first, int32_t count)
{
  return ~((first << 3) | (count - 1));
}
__host__ __device__ __forceinline__ int32_t leaf_first(int32_t code) { return (~code) >> 3; }
__host__ __device__ __forceinline__ int32_t leaf_count(int32_t code) { return ((~code) & 7) + 1; }

// Per-tree descriptor table entry (device-resident, indexed by tree id).
struct SurfTreeDev {
  const NodeD* nodes_d;
  const NodeF* nodes_f;
  const TriRec* tris;
  int32_t root;      // child-encoded root
  int32_t n_prims;
  float o_limit;     // mixed traversal is conservative for |o|_inf <= o_limit
  int32_t depth;     // max root-to-leaf depth
  double bump;
  // BVH4 over the same leaves (same leaf codes)
  const Node4D* nodes4_d;
  const Node4F* nodes4_f;
  int32_t root4;     // child-encoded root of the BVH4
  int32_t stack4;    // max traversal stack entries the BVH4 can need
  // face_normal of each triangle (leaf order, 3 doubles), computed on the
  // device at build time with the same expression as the reference
  const double* normals;
  // bounds of the tree's vertices, rounded outward to FP32 (ray ordering
  // keys, and the transport start stage's conservative volume prefilter)
  float box_lo[3];
  float box_hi[3];
};

struct ElemTreeDev {
  const NodeD* nodes_d;
  const NodeF* nodes_f;
  const int32_t* order;  // leaf-ordered element indices
  const ElemRec* elems;
  int32_t root;
  int32_t n_prims;
  int32_t depth;
  int32_t face_order[12];
  // implicit-complement re-entry (XDG::segments, xdg.cpp:164-185)
  int32_t ipc_tree;          // surface tree id, -1 if none
  const int32_t* bface_elem; // boundary face id -> element
  float box_lo[3];           // vertex bounds (walk ordering only)
  float box_hi[3];
};

}  // namespace xdgd
