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

template <int MODE, bool MIXED, class STK>
__device__ __forceinline__ void trace_s(const SurfTreeDev& T, const RayD& ray, int orientation, const int32_t* ex,
                                        int k, STK& S, HitState& h)
{
  if (MIXED && mixed_safe(T, ray))
    traverse4_s<MODE, true, false, false>(T, ray, orientation, ex, k, S, h);
  else
    traverse4_s<MODE, false, false, false>(T, ray, orientation, ex, k, S, h);
}

// Whole histories keep more state live than one query: workgroups of at
// most 256 lanes let the kernel use more than 128 VGPRs instead of spilling
// to scratch; three waves per SIMD (168 VGPRs) is the target.
template <bool MIXED>
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
      hix = (int64_t)cur++;
      have = true;
      // Particle::initialize (particle_sim.h:34-40) with a sampled start
      item = R.first + (uint64_t)hix;
      dc = 0;
      double ua, ub, uc, unused;
      uniform2(P.seed, P.stream, item, dc++, ua, ub);
      uniform2(P.seed, P.stream, item, dc++, uc, unused);
      r = d3(P.lo[0] + (P.hi[0] - P.lo[0]) * ua, P.lo[1] + (P.hi[1] - P.lo[1]) * ub,
             P.lo[2] + (P.hi[2] - P.lo[2]) * uc);
      st = TR_ALIVE;
      ev = 0;
      hh = HASH_SEED;
      hn = 0;
      if (!iso_dir(P.seed, P.stream, item, dc, u)) st = TR_ERROR;
      // XDG::find_volume (xdg.cpp:113-128) with the particle's direction
      vol = P.ipc;
      if (st == TR_ALIVE) {
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
  if (mixed)
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport<true>, block, L.lds_bytes));
  else
    XDG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport<false>, block, L.lds_bytes));
  if (nb <= 0) throw HipError{"transport launch shape does not fit one CU (block " + std::to_string(block) + ")"};
  L.grid = (int64_t)nb * cu;
  return L;
}

void launch_transport(const TransportDev& P, const TransportRun& R, bool mixed, const TransportLaunch& L,
                      hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, R.next, (int64_t)1, 0ull);
  if (P.nvol > 0)
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((P.nvol + 255) / 256)), dim3(256), 0, s, R.tally_fx,
                       (int64_t)P.nvol, 0ull);
  if (R.n <= 0) return;
  const int64_t need = (R.n + (int64_t)L.block * R.chunk - 1) / ((int64_t)L.block * R.chunk);
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min(L.grid, need));
  if (mixed)
    hipLaunchKernelGGL(k_transport<true>, dim3(g), dim3(L.block), L.lds_bytes, s, P, R);
  else
    hipLaunchKernelGGL(k_transport<false>, dim3(g), dim3(L.block), L.lds_bytes, s, P, R);
}

}  // namespace xdgh
