// adjoint.hip -- reverse-mode derivative of the ray_fire hit distance.
//
// For a ray that hit triangle f = (v0, v1, v2) at distance t, the distance
// is the plane intersection t = n.(v0 - o) / (n.d) with n = e1 x e2,
// e1 = v1 - v0, e2 = v2 - v0 (the Pluecker test of triangle_intersect.cpp
// and plucker.h:64-161 decides which triangle and returns this t).  With
// D = n.d and q = v0 - (o + t d):
//   dt/do = -n / D            dt/dd = -t n / D
//   dt/dv1 = (e2 x q) / D     dt/dv2 = (q x e1) / D
//   dt/dv0 = (n - e2 x q - q x e1) / D
// One lane per ray scales these by dL/dt; vertex gradients are scattered
// with FP64 atomics (so their summation order varies run to run), ray
// gradients are written directly.  The sorted form first orders the rays
// by hit face (a stable radix sort), then sums each wave's run of rays on
// one face with shuffles, so one lane adds the run's nine vertex terms:
// about one atomic per face and wave instead of nine per ray.  No
// counterpart in the reference.

#include <hip/hip_runtime.h>

#include "internal.h"

using namespace xdgd;

namespace xdgh {
namespace {

__device__ __forceinline__ D3 ld3(const double* p) { return d3(p[0], p[1], p[2]); }
__device__ __forceinline__ D3 add3(const D3& a, const D3& b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }

// The gradient terms of ray i: false for a ray without a usable hit.
struct AdjTerms {
  int32_t i0, i1, i2;
  D3 g0, g1, g2;  // dL/d(v0, v1, v2)
};

__device__ __forceinline__ bool adjoint_ray(const AdjointArgs& a, int64_t i, AdjTerms& A)
{
  const int32_t f = a.face[i];
  const double ti = a.t[i];
  bool ok = f >= 0 && f < a.ntri && ti < INFTY;
  int32_t i0 = 0, i1 = 0, i2 = 0;
  D3 v0 = d3(0, 0, 0), e1 = d3(0, 0, 0), e2 = d3(0, 0, 0), nn = d3(0, 0, 0);
  const D3 o = ld3(a.org + 3 * i), d = ld3(a.dir + 3 * i);
  double den = 0.0;
  if (ok) {
    i0 = a.tris[3 * (int64_t)f];
    i1 = a.tris[3 * (int64_t)f + 1];
    i2 = a.tris[3 * (int64_t)f + 2];
    ok = i0 >= 0 && i0 < a.nv && i1 >= 0 && i1 < a.nv && i2 >= 0 && i2 < a.nv;
  }
  if (ok) {
    v0 = ld3(a.verts + 3 * (int64_t)i0);
    e1 = sub(ld3(a.verts + 3 * (int64_t)i1), v0);
    e2 = sub(ld3(a.verts + 3 * (int64_t)i2), v0);
    nn = cross(e1, e2);
    den = dot(nn, d);
    ok = den != 0.0;
  }
  if (!ok) {
    if (a.dldo) a.dldo[3 * i] = a.dldo[3 * i + 1] = a.dldo[3 * i + 2] = 0.0;
    if (a.dldd) a.dldd[3 * i] = a.dldd[3 * i + 1] = a.dldd[3 * i + 2] = 0.0;
    return false;
  }
  const double w = a.dldt[i] / den;
  const D3 q = d3(v0.x - (o.x + ti * d.x), v0.y - (o.y + ti * d.y), v0.z - (o.z + ti * d.z));
  const D3 a1 = cross(e2, q);
  const D3 a2 = cross(q, e1);
  A.i0 = i0;
  A.i1 = i1;
  A.i2 = i2;
  A.g1 = d3(w * a1.x, w * a1.y, w * a1.z);
  A.g2 = d3(w * a2.x, w * a2.y, w * a2.z);
  A.g0 = d3(w * (nn.x - a1.x - a2.x), w * (nn.y - a1.y - a2.y), w * (nn.z - a1.z - a2.z));
  if (a.dldo) {
    a.dldo[3 * i] = -w * nn.x;
    a.dldo[3 * i + 1] = -w * nn.y;
    a.dldo[3 * i + 2] = -w * nn.z;
  }
  if (a.dldd) {
    a.dldd[3 * i] = -w * ti * nn.x;
    a.dldd[3 * i + 1] = -w * ti * nn.y;
    a.dldd[3 * i + 2] = -w * ti * nn.z;
  }
  return true;
}

__device__ __forceinline__ void add_vertex(double* dldv, int32_t v, const D3& g)
{
  double* p = dldv + 3 * (int64_t)v;
  atomicAdd(p + 0, g.x);
  atomicAdd(p + 1, g.y);
  atomicAdd(p + 2, g.z);
}

__global__ void k_adjoint(AdjointArgs a)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  AdjTerms A;
  if (!adjoint_ray(a, i, A) || !a.dldv) return;
  add_vertex(a.dldv, A.i1, A.g1);
  add_vertex(a.dldv, A.i2, A.g2);
  add_vertex(a.dldv, A.i0, A.g0);
}

// sort key: the hit face, or ntri for a ray without one (sorts last)
__global__ void k_adjoint_keys(AdjointArgs a, uint32_t* keys, int32_t* idx)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int32_t f = a.face[i];
  keys[i] = (f >= 0 && f < a.ntri && a.t[i] < INFTY) ? (uint32_t)f : (uint32_t)a.ntri;
  idx[i] = (int32_t)i;
}

__device__ __forceinline__ D3 shfl_down3(const D3& v, int off)
{
  return d3(__shfl_down(v.x, off), __shfl_down(v.y, off), __shfl_down(v.z, off));
}

// rays in face order: lane l sums the terms of lanes l, l+1, ... of its run
// (equal face, contiguous after the sort) in six shuffle steps; the run's
// first lane in the wave adds them.  Every lane of the wave takes part in
// the shuffles (lanes past n carry no terms).
__global__ void k_adjoint_sorted(AdjointArgs a, const int32_t* order)
{
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  AdjTerms A;
  bool ok = false;
  if (j < a.n) ok = adjoint_ray(a, order[j], A);
  const int32_t face = ok ? a.face[order[j]] : -1 - lane;  // a lane without terms is a run of its own
  if (!ok) {
    A.i0 = A.i1 = A.i2 = 0;
    A.g0 = A.g1 = A.g2 = d3(0, 0, 0);
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const D3 h0 = shfl_down3(A.g0, off), h1 = shfl_down3(A.g1, off), h2 = shfl_down3(A.g2, off);
    const int32_t fo = __shfl_down(face, off);
    if (lane + off < 64 && fo == face) {
      A.g0 = add3(A.g0, h0);
      A.g1 = add3(A.g1, h1);
      A.g2 = add3(A.g2, h2);
    }
  }
  const int32_t fprev = __shfl_up(face, 1);
  const bool head = ok && (lane == 0 || fprev != face);
  if (head && a.dldv) {
    add_vertex(a.dldv, A.i1, A.g1);
    add_vertex(a.dldv, A.i2, A.g2);
    add_vertex(a.dldv, A.i0, A.g0);
  }
}

}  // namespace

void launch_adjoint(const AdjointArgs& a, int block, hipStream_t s)
{
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_adjoint, dim3((unsigned)((a.n + block - 1) / block)), dim3(block), 0, s, a);
}

void launch_adjoint_sorted(const AdjointArgs& a, const RayOrderScratch& w, hipStream_t s)
{
  if (a.n <= 0) return;
  const unsigned g = (unsigned)((a.n + 255) / 256);
  hipLaunchKernelGGL(k_adjoint_keys, dim3(g), dim3(256), 0, s, a, w.keys_in, w.idx_in);
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) <= a.ntri) ++bits;
  radix_sort_pairs(w.keys_in, w.idx_in, w.keys_out, w.idx_out, a.n, bits, (uint32_t*)w.temp, s);
  hipLaunchKernelGGL(k_adjoint_sorted, dim3(g), dim3(256), 0, s, a, w.idx_in);
}

}  // namespace xdgh
