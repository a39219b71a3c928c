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
// gradients are written directly.  No counterpart in the reference.

#include <hip/hip_runtime.h>

#include "internal.h"

using namespace xdgd;

namespace xdgh {
namespace {

__device__ __forceinline__ D3 ld3(const double* p) { return d3(p[0], p[1], p[2]); }

__global__ void k_adjoint(AdjointArgs a)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
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
    return;
  }
  const double w = a.dldt[i] / den;
  const D3 q = d3(v0.x - (o.x + ti * d.x), v0.y - (o.y + ti * d.y), v0.z - (o.z + ti * d.z));
  const D3 a1 = cross(e2, q);
  const D3 a2 = cross(q, e1);
  if (a.dldv) {
    double* g0 = a.dldv + 3 * (int64_t)i0;
    double* g1 = a.dldv + 3 * (int64_t)i1;
    double* g2 = a.dldv + 3 * (int64_t)i2;
    atomicAdd(g1 + 0, w * a1.x);
    atomicAdd(g1 + 1, w * a1.y);
    atomicAdd(g1 + 2, w * a1.z);
    atomicAdd(g2 + 0, w * a2.x);
    atomicAdd(g2 + 1, w * a2.y);
    atomicAdd(g2 + 2, w * a2.z);
    atomicAdd(g0 + 0, w * (nn.x - a1.x - a2.x));
    atomicAdd(g0 + 1, w * (nn.y - a1.y - a2.y));
    atomicAdd(g0 + 2, w * (nn.z - a1.z - a2.z));
  }
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
}

}  // namespace

void launch_adjoint(const AdjointArgs& a, int block, hipStream_t s)
{
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_adjoint, dim3((unsigned)((a.n + block - 1) / block)), dim3(block), 0, s, a);
}

}  // namespace xdgh
