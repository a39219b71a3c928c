// internal.h -- host-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "xdg_device.h"

namespace xdgh {

struct HipError {
  std::string msg;
};

#define XDG_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::xdgh::HipError{std::string(#expr) + ": " + hipGetErrorString(_e)};         \
  } while (0)

// Owned device allocation (freed by the owner).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

void* dev_alloc(size_t bytes, std::vector<DevBuf>& owned);
void dev_free_all(std::vector<DevBuf>& owned);

// Binary-tree topology kept after a build so the boxes can be refitted for
// moved primitives (same connectivity): Karras children, parents and leaf
// ranges, the refit arrival flags and the internal node boxes.
struct LbvhTopo {
  int32_t* left = nullptr;
  int32_t* right = nullptr;
  int32_t* pint = nullptr;   // parent of internal node i (-1 for the root)
  int32_t* pleaf = nullptr;  // parent of leaf position s
  int32_t* rfirst = nullptr;
  int32_t* rlast = nullptr;
  unsigned int* flags = nullptr;
  double* ibox = nullptr;    // 6 per internal node
};

// Result of one LBVH build over n primitive boxes.
struct LbvhOut {
  xdgd::NodeD* nodes_d = nullptr;  // n-1 slots (unreachable slots left unused)
  xdgd::NodeF* nodes_f = nullptr;
  int32_t* order = nullptr;        // leaf-ordered primitive indices (n)
  int32_t root = 0;
  int32_t depth = 0;
  int64_t n_nodes = 0;
  int32_t leaf_max = 2;
  double max_abs = 0.0;            // max |coordinate| over all boxes
  LbvhTopo topo;                   // kept only when a refit is wanted
};

// Stable LSD radix sort (4-bit digits, hand-written) of (key, int32 value)
// pairs by the low `bits` bits of the keys.  The result lands in keys/vals;
// k2/v2 are scratch of n entries and hist holds radix_sort_hist_words(n)
// words.  Asynchronous on s; allocates nothing.
int64_t radix_sort_hist_words(int64_t n);
void radix_sort_pairs(uint64_t* keys, int32_t* vals, uint64_t* k2, int32_t* v2, int64_t n, int bits, uint32_t* hist,
                      hipStream_t s);
void radix_sort_pairs(uint32_t* keys, int32_t* vals, uint32_t* k2, int32_t* v2, int64_t n, int bits, uint32_t* hist,
                      hipStream_t s);
// exclusive scan of a[0..m) in place by one workgroup; the sum goes to
// *total (nullable).  Asynchronous on s.
void launch_scan_excl(uint32_t* a, int64_t m, uint32_t* total, hipStream_t s);

// box6: device array of n boxes (lo.xyz, hi.xyz), already dilated.  The
// whole build runs on the device: 63-bit Morton codes, a stable LSD radix
// sort, the hierarchy (the Karras radix tree over the Morton order, or with
// sah a top-down binned-SAH split of that order, level by level), a
// bottom-up refit and a structural check by breadth-first levels; only
// counts and the max |coordinate| are read back.
void lbvh_build(const double* box6, int64_t n, int leaf_max, hipStream_t s, std::vector<DevBuf>& owned,
                LbvhOut& out, bool keep_topology = false, bool sah = false);
// New boxes for the same primitives (same order and connectivity): node
// boxes recomputed bottom-up, topology unchanged (needs out.topo).
void lbvh_refit(const double* box6, int64_t n, LbvhOut& out, hipStream_t s);

// 4-wide tree over the same leaves, collapsed from a binary LBVH on the
// device, level by level.
struct Bvh4Out {
  xdgd::Node4D* nodes_d = nullptr;
  xdgd::Node4F* nodes_f = nullptr;
  int32_t root = 0;
  int32_t stack = 0;  // max stack entries any traversal can need
  int64_t n_nodes = 0;
  int32_t* src = nullptr;  // 4 per node: binary (node << 1 | side) each child box comes from, -1 empty
};
void bvh4_collapse(const LbvhOut& bin, int64_t n, hipStream_t s, std::vector<DevBuf>& owned, Bvh4Out& out,
                   bool keep_sources = false);
// Child boxes of the BVH4 copied again from the (refitted) binary nodes.
void bvh4_refit(const LbvhOut& bin, const Bvh4Out& b4, hipStream_t s);

// ---- surface queries (surface_query.hip) ---------------------------------
struct RayFireArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;        // table holds ids 1..ntrees
  int32_t tree;          // used when tree_per_ray == nullptr
  const int32_t* tree_per_ray;
  int64_t n;
  const double* org;
  const double* dir;
  const double* tmax;    // nullable
  int orientation;
  const int32_t* excl;   // nullable, k per ray
  int k;
  double* t_out;
  int32_t* surf_out;
  int32_t* face_out;
  // mixed precision only (nullable): rays whose origin lies beyond the tree's
  // FP32-rounding limit are queued here and traced by a second FP64 pass,
  // which keeps FP64 box state out of the mixed kernel's registers
  int32_t* fb_list;
  unsigned int* fb_count;
  // persistent kernels (nullable): next unclaimed ray of the batch
  unsigned long long* next_ray;
  // nullable: slot j of the launch traces ray order[j] (results still land
  // at the ray's own index)
  const int32_t* order;
};

// ---- ray ordering (ray_order.hip) -----------------------------------------
struct RayOrderArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int32_t tree;
  const int32_t* tree_per_ray;
  int64_t n;
  const double* org;
  const double* dir;
};
struct RayOrderScratch {
  uint32_t* keys_in;
  uint32_t* keys_out;
  int32_t* idx_in;
  int32_t* idx_out;
  void* temp;
  size_t temp_bytes;
};
size_t ray_order_temp_bytes(int64_t n);
// mode 1: (tree, octant, origin Morton); mode 2: adds one coarse direction
// bit per axis.  Returns the permutation (w.idx_out), valid on stream s.
const int32_t* launch_ray_order(const RayOrderArgs& a, int mode, const RayOrderScratch& w, hipStream_t s);

struct LaunchCfg {
  int block = 64;
  int stack = 64;     // LDS stack entries per lane
  int defer = 1;  // mixed ray_fire: FP64 pass for far-origin rays
  int wide = 1;   // surface queries traverse the BVH4
  // resident-grid kernels with per-lane work refill: bit 0 plain ray_fire,
  // bit 1 mixed ray_fire, bit 2 find_volume
  int persistent = 7;
  int order = 0;  // ray_fire coherence ordering mode (0 = off)
  int spec = 1;   // ray_fire BVH4 node loop parks leaves until the wave holds one each
  int refill = 44; // persistent ray_fire: idle lanes of a wave needed before a refill
  int fv_refill = 0;  // persistent find_volume: the same threshold (0: 56 mixed, 40 plain)
  int adj_sort = 1;   // adjoint batches of >= 65536 rays: ordered by hit face, vertex sums per wave
  int mw_wave = 1;  // mesh walk: batched re-entry rays (wavefront) instead of inline ones
  int mw_refill = 32;  // mesh walk rounds on a resident grid, lanes claiming particles in batches of this many (0: off)
  int seg_order = 0;  // segment walks in Morton order of the start points
  int tr_start = 2;   // transport starts: 0 in the event kernel, 1 separate kernels, 2 separate and sorted
  int tr_slice = 22;  // transport, staged form: log2 of the histories per slice (0.8-1.3 GB of state at 22)
  int tr_items = 8;   // transport, staged form: start-stage item buffer entries per history
  int tr_rounds = 0;  // transport, staged form: sorted one-event rounds before the event kernel takes the rest
  int tr_events = 1;  // transport, staged form: events by the traversal-level kernel (1) or whole-event kernel (0)
};

void launch_ray_fire(const RayFireArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);
// counters[12] (device, zeroed): rays, V, F, E, C, N, hits, mixed fallbacks,
// wave node iterations, wave triangle iterations, per-wave max V, max F
void launch_ray_fire_stats(const RayFireArgs& a, bool mixed, const LaunchCfg& c, unsigned long long* counters,
                           hipStream_t s);

struct PivArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int32_t tree;
  int64_t n;
  const double* pts;
  const double* dir;  // nullable -> default direction
  double dflt[3];
  const int32_t* excl;
  int k;
  uint8_t* inside;
};
void launch_point_in_volume(const PivArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);

struct FindVolumeArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int nvol;
  const int32_t* vol_trees;  // device, nvol
  const int32_t* vol_ids;    // device, nvol
  int32_t ipc;
  int64_t n;
  const double* pts;
  const double* dir;  // nullable -> default direction
  double dflt[3];
  int32_t* vol_out;
  // persistent launch (nullable): point counter; mixed precision also queues
  // points whose origin is beyond a tree's FP32 limit for an FP64 pass
  unsigned long long* next;
  int32_t* fb_list;
  unsigned int* fb_count;
};
void launch_find_volume(const FindVolumeArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);

// Overlap check (tools/overlap.cpp): every volume of a list, no early exit
// for points; first hit over the list for edge rays.
struct VolumesArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int nvol;
  const int32_t* vol_trees;  // device, nvol
  const int32_t* vol_ids;    // device, nvol
  int64_t n;
  const double* pts;         // points / edge-ray origins
  const double* dir;         // one direction per item
  // point_in_volumes
  int max_k;
  int32_t* count_out;
  int32_t* ids_out;          // max_k per point
  // edge rays
  const double* len;         // distance limit per ray
  const int32_t* skip;       // two volume ids per ray that are not tested
  double* t_out;
  int32_t* surf_out;
  int32_t* vol_out;
};
void launch_point_in_volumes(const VolumesArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);
void launch_edge_rays(const VolumesArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);

struct OccludedArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int32_t tree;
  int64_t n;
  const double* org;
  const double* dir;
  uint8_t* hit;
  double* dist;  // nullable
};
void launch_occluded(const OccludedArgs& a, bool mixed, const LaunchCfg& c, hipStream_t s);

struct ClosestArgs {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int32_t tree;
  int64_t n;
  const double* pts;
  double* dist;
  int32_t* face;
};
void launch_closest(const ClosestArgs& a, const LaunchCfg& c, hipStream_t s);

// ---- batched particle random walk (transport.hip) -------------------------
// Device-resident problem (arrays owned by the context's transport handle).
struct TransportDev {
  const xdgd::SurfTreeDev* table;
  int32_t ntrees;
  int32_t nvol;
  const int32_t* vol_tree;  // slot -> surface tree id
  const int32_t* vol_id;    // slot -> volume id (find_volume order)
  const int32_t* id2slot;   // volume id -> slot (-1), ids [0, n_ids)
  int32_t n_ids;
  int32_t ipc;
  int32_t nsurf;            // sense tables indexed by surface id [0, nsurf)
  const int32_t* sense_fwd;
  const int32_t* sense_rev;
  uint32_t seed, stream;
  double lo[3], hi[3];
  double mfp, p_absorb;
  int32_t max_events;
  const uint8_t* surf_bc;   // nullable: boundary condition per surface id [0, nsurf)
  double* pre_box;          // nvol x (lo.xyz, hi.xyz): start-stage prefilter boxes, written per launch
  uint32_t* slot_cnt;       // nvol + 1: start-stage items per volume slot, then their offsets (+ total)
  uint32_t* slot_cur;       // nvol: scatter cursors
};
// Start state of one slice of histories, written by the start stage and
// read by the event kernel (component-major: x, y, z, ux, uy, uz blocks of
// cap doubles each), and the order in which both run them.
struct TransportStage {
  double* ru;
  int32_t* vol;
  uint32_t* dc;   // Philox counter after the start draws
  uint8_t* st;
  uint32_t* key;  // coherence key of the start
  int32_t* idx;   // processing order: slice-relative history indices
  uint32_t* k2;   // sort scratch
  int32_t* v2;
  uint32_t* hist;  // radix_sort_hist_words(cap) words
  int32_t* found;  // lowest volume slot found to contain the start (INT32_MAX: none yet)
  uint8_t* over;   // the history's candidate items did not fit the item buffer
  int2* items;     // (history, volume slot) point_in_volume items, grouped by slot
  int64_t item_cap;
  unsigned long long* masks;  // candidate masks, mask_words x cap (chunk-major, by list position)
  int mask_words;             // chunks of 64 volume slots, or 0 when the masks are not kept
  // event state between rounds
  int32_t* ev;
  unsigned long long* hh;
  uint8_t* hn;
  int32_t* hist8;  // exclusion history, 8 x cap (entry-major)
  uint32_t* live;  // live histories at the head of idx (one word)
  unsigned long long* claim;  // item claim counter of the start stage's persistent kernel
  int64_t cap;    // histories per slice
};
// per history, beside the sort histogram and the 8-byte items: position and
// direction, hash, 8 four-byte arrays, the exclusion history, 3 bytes
constexpr size_t TRANSPORT_STAGE_BYTES = 6 * 8 + 8 + 8 * 4 + 8 * 4 + 3;
// One launch over histories [first, first + n).
struct TransportRun {
  uint64_t first;
  int64_t n;
  unsigned long long* next;      // claim counter (zeroed by the launcher)
  unsigned long long* tally_fx;  // per slot, int64 two's complement, scale 2^32
  int32_t* events;               // nullable, per history
  uint8_t* status;               // nullable
  uint64_t* hash;                // nullable
  int chunk;                     // histories a lane claims at a time
  int lds_stack;                 // stack entries kept in LDS
  int stack_cap;                 // total stack entries (LDS + global)
  int32_t* spill;                // global stack share: (cap - lds) * spill_lanes
  int64_t spill_lanes;
  int lds_tally;                 // per-workgroup LDS tally (nvol small enough)
  int refill;                    // idle lanes a wave waits for before starting histories
  int staged;                    // starts as separate kernels per slice of stage.cap histories (2: sorted)
  TransportStage stage;
  int64_t start_grid;            // resident workgroups of the start stage's point_in_volume kernel
  int rounds;                    // staged form: sorted one-event rounds before the event kernel
  int events_pt;                 // staged form: traversal-level event kernel (k_transport_pt)
};
struct TransportLaunch {
  int block;
  int64_t grid;     // resident workgroups of the whole-event kernel (at least grid_pt)
  int64_t grid_pt;  // of the traversal-level event kernel
  size_t lds_bytes;
};
// Shape of the persistent grid for (block, lds_stack, nvol): resident
// workgroups on the whole device.
TransportLaunch transport_shape(bool mixed, int block, int lds_stack, int nvol);
void launch_transport(const TransportDev& P, const TransportRun& R, bool mixed, const TransportLaunch& L,
                      hipStream_t s);
constexpr int TRANSPORT_LDS_TALLY_MAX = 4096;
// start stage: its traversal stack lives wholly in LDS, at most this much
// per workgroup; 0 = no block fits the stack
constexpr size_t TRANSPORT_START_LDS_MAX = 64 * 1024;
int transport_start_block(int stack_cap);
constexpr int TRANSPORT_MASK_WORDS_MAX = 16;  // candidate masks kept for up to 1024 volume slots
int64_t transport_start_grid(bool mixed, int stack_cap);

// ---- adjoint of the ray distance (adjoint.hip) ----------------------------
struct AdjointArgs {
  int64_t n;
  const double* org;
  const double* dir;
  const double* t;
  const int32_t* face;
  const double* dldt;
  int64_t nv;
  const double* verts;
  int64_t ntri;
  const int32_t* tris;
  double* dldv;  // nullable, accumulated
  double* dldo;  // nullable, written
  double* dldd;  // nullable, written
};
void launch_adjoint(const AdjointArgs& a, int block, hipStream_t s);
// the same with the rays ordered by hit face first (w: ray-order scratch of
// at least a.n entries); the vertex sums run per wave and face
void launch_adjoint_sorted(const AdjointArgs& a, const RayOrderScratch& w, hipStream_t s);

// ---- element queries (element_query.hip) ---------------------------------
struct FindElementArgs {
  const xdgd::ElemTreeDev* etable;
  int32_t nelem_trees;
  int32_t tree;
  int64_t n;
  const double* pts;
  int32_t* elem_out;
};
void launch_find_element(const FindElementArgs& a, const LaunchCfg& c, hipStream_t s);

struct NextElementArgs {
  const xdgd::ElemTreeDev* etable;
  int32_t nelem_trees;
  int32_t tree;
  int64_t n;
  const int32_t* elem;
  const double* r;
  const double* u;
  int32_t* next_out;
  double* dist_out;
};
void launch_next_element(const NextElementArgs& a, hipStream_t s);

struct SegmentsArgs {
  const xdgd::ElemTreeDev* etable;
  int32_t nelem_trees;
  const xdgd::SurfTreeDev* stable;
  int32_t nsurf_trees;
  int32_t tree;
  int64_t n;
  const double* p0;
  const double* p1;
  unsigned long long* tally_fx;  // nullable, per element
  int32_t* count;
  uint64_t* hash;
  const int64_t* seq_off;  // nullable
  int32_t* seq_elem;
  double* seq_len;
  uint32_t* status;  // device word: bit0 = step cap hit, bit1 = no element behind the entry hit
  // XDG::segments(volume, start, end): walks only the element tree and
  // enters it through surface tree surf_tree (ENTERING)
  int volume_mode;
  int32_t surf_tree;
  // nullable: slot j of the launch walks segment order[j] (outputs still
  // land at the segment's own index)
  const int32_t* order;
};
void launch_segments(const SegmentsArgs& a, int face_order_kind, const LaunchCfg& c, hipStream_t s);
// Walk order of a segment batch (stable, so equal keys keep the caller's
// order): mode 1 by the 30-bit Morton code of the start point inside the
// element tree's bounds, mode 2 by start cell then end cell (15 bits each).
// Returns the permutation (in w), valid on stream s.
const int32_t* launch_segment_order(const SegmentsArgs& a, int mode, const RayOrderScratch& w, hipStream_t s);

// Element-by-element particle walk (tools/walk_elements.h:53-117).
struct MeshWalkArgs {
  const xdgd::ElemTreeDev* etable;
  int32_t nelem_trees;
  const xdgd::SurfTreeDev* stable;
  int32_t nsurf_trees;
  int32_t tree;
  uint64_t first;
  int64_t n;
  uint32_t seed, stream;
  double lo[3], hi[3];  // start sampling box
  double mfp;
  int32_t max_events;
  double* distance;               // nullable outputs, one per particle
  int32_t* events;
  uint8_t* status;
  uint64_t* hash;
  unsigned long long* tally_fx;   // nullable, per element, scale 2^40
};
// A particle's state between the kernels of the wavefront mesh walk.
struct MwSave {
  double r[3], u[3], dist;
  unsigned long long hh;
  int32_t e, nev;
  uint32_t dc;
  int32_t pad;
};
constexpr int MW_ROUNDS = 3;  // walk / batched re-entry rounds before the inline fallback
struct MeshWalkScratch {
  MwSave* save;               // n particles
  int32_t* list[2];           // parked and re-entered particle lists, n each
  unsigned int* count;        // 2 * MW_ROUNDS list counters, then the claim counters
  unsigned long long* claim;  // MW_ROUNDS claim counters of the resident-grid walk (inside count's words)
};
constexpr int MW_COUNT_WORDS = 2 * MW_ROUNDS + 2 + 2 * MW_ROUNDS;  // 32-bit words zeroed per launch
// w (nullable): scratch of the wavefront walk; null runs every particle
// inline, one per lane
void launch_mesh_walk(const MeshWalkArgs& a, int face_order_kind, const LaunchCfg& c, const MeshWalkScratch* w,
                      hipStream_t s);
// 0 = MOAB canonical, 1 = MeshMock, 2 = other (read from the tree)
int face_order_kind(const int32_t* face_order);

// device evaluation of one geometry primitive per item (parity checks)
void launch_eval(int kind, int64_t n, const double* in, double* out, hipStream_t s);

// element prim boxes (per-element dilation, tetrahedron_contain.cpp:47-63)
void launch_elem_boxes(const xdgd::ElemRec* elems, int64_t n, double* box6, hipStream_t s);
// triangle prim boxes (+- bump, triangle_intersect.cpp:36-50)
void launch_tri_boxes(const double* verts9, int64_t n, double bump, double* box6, hipStream_t s);
// per-triangle face normals (leaf order, 3 doubles each)
void launch_tri_normals(const xdgd::TriRec* tris, int64_t n, double* out, hipStream_t s);
// leaf-ordered triangle records
void launch_write_tris(const double* verts9, const int32_t* face, const int32_t* surf, const uint8_t* flip,
                       const int32_t* order, int64_t n, xdgd::TriRec* out, hipStream_t s);

// Resident workgroups of a kernel on the whole device for (block, lds).
template <class K>
inline int64_t resident_grid(K kernel, int block, size_t lds)
{
  static int cu = 0;
  if (cu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
      cu = 256;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, lds) != hipSuccess || nb <= 0) nb = 1;
  return (int64_t)nb * cu;
}

}  // namespace xdgh
