/*
 * xdg_hip.h -- C-ABI of the MI355X geometry-query backend (libxdg_hip.so).
 *
 * This is the drop-in boundary beside XDG's Embree and GPRT backends.  Each
 * entry point replaces one method of the reference's abstract RayTracer
 * (include/xdg/ray_tracing_interface.h:16-147) or the facade callers on the
 * hot path (src/xdg.cpp), batched: one call processes n queries.  A C++
 * RayTracer-shaped class over this ABI is in include/xdg_hip.hpp.
 *
 * Conventions
 *   - Plain pointers and sizes only.  "host" arrays are read during the call;
 *     "device" arrays are device pointers (hipMalloc'd or any allocator on
 *     the same device), read/written asynchronously on `stream`
 *     (NULL = the context's stream).  Results are ready after
 *     xdg_hip_synchronize(ctx, stream).
 *   - Every function returns an int status: XDG_HIP_OK (0) or a negative
 *     code; xdg_hip_last_error(ctx) describes the last failure.  No C++
 *     exception crosses the ABI.
 *   - Calls on one context from several host threads are serialized.
 *   - Streams: queries on different streams may overlap, except that the
 *     context's shared scratch (the persistent kernels' work counters, the
 *     mixed-precision deferral list, the ray-ordering buffers, the mesh
 *     walk's particle store) serves one launch at a time.  The library
 *     orders those launches itself: a ray_fire, find_volume or wavefront
 *     mesh walk on stream B waits (hipStreamWaitEvent) for the last such
 *     launch on stream A.  Under graph capture that ordering is the
 *     caller's.  A growth of the scratch (a larger batch than before)
 *     synchronizes the device.
 *   - Types follow include/xdg/constants.h: ids are int32, ID_NONE = -1,
 *     INFTY = DBL_MAX.  Tree ids start at 1 (ray_tracing_interface.cpp:10-18).
 *   - Fixed-point tallies (order independent, so bit-exact for any launch
 *     shape): tet-walk and mesh-walk tallies are int64 sum(llrint(len *
 *     2^40)) per element, exact to 2^-40 and safe up to 2^62 / 2^40 =
 *     4.19e6 length units per element; transport tallies are int64
 *     sum(llrint(len * 2^32)) per volume slot, safe up to 2^30 = 1.07e9.
 *     After each tallying call a check kernel sets status bit
 *     XDG_HIP_STATUS_TALLY_RANGE when any entry of the tally array has
 *     reached 2^62 in magnitude (or wrapped negative), before 2^63 wraps.
 *   - Tie rule (the only intended difference from Embree's visit-order
 *     dependence, triangle_intersect.cpp:78): among bitwise-equal distances
 *     the smallest face id wins; find_element returns the lowest containing
 *     element index.
 */
#ifndef XDG_HIP_H
#define XDG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XDG_HIP_OK 0
#define XDG_HIP_ERR_INVALID (-1)     /* bad argument */
#define XDG_HIP_ERR_NOT_FOUND (-2)   /* unknown tree id (std::map::at throws in the reference) */
#define XDG_HIP_ERR_HIP (-3)         /* HIP runtime error */
#define XDG_HIP_ERR_NO_DEVICE (-4)   /* no usable GPU */
#define XDG_HIP_ERR_CAPPED (-5)      /* a walk hit its step bound (results kept) */

/* precision of box traversal */
#define XDG_HIP_PLAIN 0 /* FP64 boxes + FP64 Pluecker */
#define XDG_HIP_MIXED 1 /* outward-rounded FP32 boxes + FP64 Pluecker */

/* HitOrientation (include/xdg/shared_enums.h:11-15) */
#define XDG_HIP_ANY (-1)
#define XDG_HIP_EXITING 0
#define XDG_HIP_ENTERING 1

#define XDG_HIP_MAX_EXCLUDE 1024 /* exclusion slots per ray */

typedef struct xdg_hip_ctx xdg_hip_ctx;

typedef struct {
  int64_t n_prims;   /* triangles or elements */
  int64_t n_nodes;   /* BVH2 node slots */
  int32_t depth;     /* max root-to-leaf depth */
  int32_t root;      /* child-encoded root */
  int64_t bytes;     /* device bytes held by the tree */
  double bump;       /* box dilation used */
  double o_limit;    /* mixed traversal is exact for |origin|_inf <= o_limit */
  int64_t n_nodes4;  /* 4-wide nodes (surface trees) */
  int32_t stack4;    /* traversal stack entries the 4-wide tree can need */
  int32_t reserved;
} xdg_hip_tree_info;

/* ---- lifecycle --------------------------------------------------------- */

/* Replaces constructing EmbreeRayTracer (src/embree/ray_tracer.cpp:15-24)
 * and RayTracer::init (ray_tracing_interface.h:22). */
int xdg_hip_create(xdg_hip_ctx** ctx, int device);
int xdg_hip_destroy(xdg_hip_ctx* ctx);
const char* xdg_hip_last_error(const xdg_hip_ctx* ctx);
/* Launch shape for the traversal kernels (block: 64..1024, multiple of 64;
 * stack: LDS entries per lane, >= the deepest tree, <= 64). */
int xdg_hip_set_launch(xdg_hip_ctx* ctx, int block, int stack);
int xdg_hip_get_launch(xdg_hip_ctx* ctx, int* block, int* stack);
/* Traversal variant: wide != 0 uses the 4-wide tree (default), else the
 * binary tree it is collapsed from; defer != 0 (default) traces far-origin mixed ray_fire rays in
 * a separate FP64 pass; persistent selects resident-grid kernels whose
 * lanes claim new work from a counter as they finish (bit 0 plain
 * ray_fire, bit 1 mixed ray_fire, bit 2 find_volume; default 7).  Results
 * are identical either way. */
int xdg_hip_set_traversal(xdg_hip_ctx* ctx, int wide, int defer, int persistent);
/* Tuning options by name (no counterpart in the reference; results are
 * identical for every setting):
 *   "block"       workgroup size, multiple of 64 (default 64)
 *   "stack"       minimum LDS stack entries per lane (default 1; raised to
 *                 what each tree needs)
 *   "persistent"  resident-grid kernels: bit 0 plain ray_fire, bit 1 mixed
 *                 ray_fire, bit 2 find_volume (default 7)
 *   "ray_order"   ray_fire batches of >= 65536 rays: 0 caller's order,
 *                 1 sorted by (tree, direction octant, origin Morton code),
 *                 2 as 1 plus one coarse direction bit per axis (0)
 *   "speculative" ray_fire BVH4 node loop keeps descending with a parked
 *                 leaf until every lane of the wave holds one (1)
 *   "wide"        BVH4 traversal (1) or the binary tree (0)
 *   "defer"       mixed ray_fire traces far-origin rays in an FP64 pass (1)
 *   "leaf_size"   max triangles per leaf for trees built afterwards (2)
 *   "build"       hierarchy of trees built afterwards: top-down binned SAH
 *                 (1, default; ranges above 32768 primitives split at the
 *                 Morton bit) or the Karras radix tree over the Morton
 *                 order (0).  Only speed differs, never results
 *   "refill"      persistent ray_fire: idle lanes a wave waits for before
 *                 it claims new work (1..64, default 44)
 *   "fv_refill"   the same for the persistent find_volume (1..64; 0, the
 *                 default: 56 mixed, 40 plain)
 *   "adjoint_sort" adjoint batches of >= 65536 rays with dldv: rays ordered
 *                 by hit face and each wave's run on one face summed before
 *                 the vertex atomics (1, default), or nine atomics per ray (0);
 *                 dldv's FP64 summation order differs, ray gradients do not
 *   "mw_wave"     mesh walk: particles that leave the mesh are parked and
 *                 their re-entry rays traced together in a separate kernel
 *                 (1), or each lane traces its own at once (0) (default 1)
 *   "mw_refill"   mesh walk with mw_wave 1: the parking rounds run on a
 *                 resident grid whose lanes claim the next particles once
 *                 this many lanes of a wave are idle (1..64, default 32),
 *                 or one particle per lane (0)
 *   "refit"       surface trees built while set keep their topology, so
 *                 xdg_hip_refit_surface_tree can move their vertices (0)
 *   "segment_order" segment batches of >= 65536: walked in the Morton order
 *                 of their start points (1) or in the caller's order (0)
 *   "transport_start" transport starts (sampling, find_volume): inside the
 *                 event kernel (0), as separate kernels in history order
 *                 (1), or as separate kernels on histories sorted by start
 *                 cell and direction, with find_volume as point_in_volume
 *                 items grouped by volume (2, default)
 *   "transport_slice" staged starts run in slices of 2^value histories
 *                 (6..24, default 22: 0.8 to 1.3 GB of state per slice)
 *   "transport_items" start-stage item buffer entries per history (1..64,
 *                 default 8); histories whose items do not fit run the
 *                 volume loop themselves
 *   "transport_events" staged form: events (and the start stage's
 *                 point_in_volume items) by kernels whose traversal is a
 *                 per-lane state machine, so a lane whose ray has ended goes
 *                 on to its next event without waiting for the wave's
 *                 longest traversal (1, default), or by whole-event kernels
 *                 (0)
 *   "transport_rounds" staged form: one-event rounds over the live
 *                 histories, sorted by volume, cell and direction before each
 *                 round, before the event kernel takes the rest (0..64,
 *                 default 0)
 * Unknown names give XDG_HIP_ERR_NOT_FOUND, out-of-range values
 * XDG_HIP_ERR_INVALID. */
int xdg_hip_set_option(xdg_hip_ctx* ctx, const char* name, int value);
int xdg_hip_get_option(xdg_hip_ctx* ctx, const char* name, int* value);
int xdg_hip_synchronize(xdg_hip_ctx* ctx, void* stream);

/* ---- tree construction (host arrays in, device tree out) ---------------- */

/* RayTracer::create_surface_tree (ray_tracing_interface.h:51; Embree:
 * ray_tracer.cpp:50-105).  One tree = all triangles of one volume, in the
 * caller's order: tris[3*i..] index verts; face_ids[i] is the mesh face id
 * returned by queries; surf_ids[i] the surface id (0 <= id < 2^27); flip[i] = 1 when this
 * volume is the surface's reverse-sense parent (triangle_intersect.cpp:84).
 * bump: box dilation, normally max(1e-6 * volume box chord, 1e-3)
 * (ray_tracing_interface.cpp:20-24). */
int xdg_hip_build_surface_tree(xdg_hip_ctx* ctx, int64_t n_verts, const double* verts, int64_t n_tris,
                               const int32_t* tris, const int32_t* face_ids, const int32_t* surf_ids,
                               const uint8_t* flip, double bump, int32_t* tree_out);

/* RayTracer::create_element_tree (ray_tracing_interface.h:63; Embree:
 * ray_tracer.cpp:150-188) for tetrahedra.  tets: 4 vertex indices per
 * element (element id = index); adj: neighbour across face f (-1 boundary);
 * face_order: 12 ints, the local vertex triples of faces 0..3 (MOAB
 * canonical {0,1,3},{1,2,3},{2,0,3},{0,2,1}, include/xdg/moab/
 * direct_access.h:181-183). */
int xdg_hip_build_element_tree(xdg_hip_ctx* ctx, int64_t n_verts, const double* verts, int64_t n_tets,
                               const int32_t* tets, const int32_t* adj, const int32_t* face_order,
                               int32_t* tree_out);

/* Attach the implicit-complement surface tree and the boundary face ->
 * element map (MeshManager::get_boundary_face_element) used by the
 * re-entry branch of segments (src/xdg.cpp:164-185).  bface_elem is
 * indexed by the face ids of the surface tree. */
int xdg_hip_attach_boundary(xdg_hip_ctx* ctx, int32_t element_tree, int32_t surface_tree, int64_t n_faces,
                            const int32_t* bface_elem);

/* New coordinates for the vertices of a surface tree built while the
 * "refit" option was set (same n_verts, triangles and ids as the build;
 * verts is a DEVICE array of 3 n_verts doubles): triangle records, normals,
 * box dilation (bump >= 0 replaces it, bump < 0 keeps it), both node layouts
 * and the mixed-precision limit are recomputed on the device with the
 * topology kept.  Queries stay exact for any vertex motion (boxes are
 * recomputed, only their tightness changes).  Synchronous.  No counterpart
 * in the reference (it rebuilds); for callers that move vertices with fixed
 * connectivity, such as the differentiable-surface harness. */
int xdg_hip_refit_surface_tree(xdg_hip_ctx* ctx, int32_t tree, int64_t n_verts, const double* verts, double bump);

int xdg_hip_get_tree_info(xdg_hip_ctx* ctx, int32_t tree, int element, xdg_hip_tree_info* info);
int xdg_hip_num_trees(xdg_hip_ctx* ctx, int32_t* surface_trees, int32_t* element_trees);
/* Frees a surface tree's device memory; its id is not reused and later
 * queries on it return XDG_HIP_ERR_NOT_FOUND.  (The reference keeps trees
 * for the RayTracer's lifetime; this serves callers that rebuild geometry,
 * such as the differentiable-surface harness.) */
int xdg_hip_release_surface_tree(xdg_hip_ctx* ctx, int32_t tree);

/* ---- queries (device arrays) -------------------------------------------- */

/* RayTracer::ray_fire (ray_tracing_interface.h:81-86; Embree
 * ray_tracer.cpp:290-326).  tree_per_ray (device, nullable) overrides
 * `tree` per ray.  tmax (device, nullable): dist_limit per ray (INFTY when
 * NULL).  excl (device, nullable): k <= XDG_HIP_MAX_EXCLUDE (1024) face
 * ids per ray, read in place (-1 = empty).
 * Miss: t = INFTY, surf = face = -1.  The reference appends the hit face to
 * the caller's exclusion vector: face_out carries it. */
int xdg_hip_ray_fire(xdg_hip_ctx* ctx, int32_t tree, const int32_t* tree_per_ray, int64_t n, const double* org,
                     const double* dir, const double* tmax, int orientation, const int32_t* excl, int k,
                     int precision, double* t_out, int32_t* surf_out, int32_t* face_out, void* stream);

/* xdg_hip_ray_fire without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  No counterpart in the reference.  Same arguments,
 * results and kernels as xdg_hip_ray_fire; its scratch (work counter and
 * the mixed kernel's deferral list) is one array per GPU in the library,
 * 16 MiB, held until the process ends, instead of context buffers grown on
 * demand.  n must not exceed XDG_HIP_NOALLOC_MAX_RAYS (XDG_HIP_ERR_INVALID,
 * nothing launched); rays are traced in the caller's order (the
 * "ray_order" option is ignored).  All calls on one GPU share the array:
 * calls that overlap on one GPU (other streams, other contexts) race, so
 * order them.  Counters are cleared by kernels, not memsets, so the call
 * may be captured. */
#define XDG_HIP_NOALLOC_MAX_RAYS 4194304
int xdg_hip_ray_fire_noalloc(xdg_hip_ctx* ctx, int32_t tree, const int32_t* tree_per_ray, int64_t n,
                             const double* org, const double* dir, const double* tmax, int orientation,
                             const int32_t* excl, int k, int precision, double* t_out, int32_t* surf_out,
                             int32_t* face_out, void* stream);

/* Instrumented xdg_hip_ray_fire (same traversal, no outputs) for roofline
 * accounting; synchronous.  counters (host, XDG_HIP_NUM_COUNTERS): rays, BVH
 * nodes visited, triangle tests, Pluecker edge tests, tests reaching the
 * distance computation, normal evaluations, hits, mixed-precision fallbacks
 * to FP64 boxes, node-loop iterations executed per wave (summed over waves),
 * triangle iterations executed per wave, node visits of the longest ray of
 * each wave (summed over waves), triangle tests of the longest ray of each
 * wave.  No counterpart in the reference (Embree exposes no counters). */
#define XDG_HIP_NUM_COUNTERS 12
int xdg_hip_ray_fire_stats(xdg_hip_ctx* ctx, int32_t tree, const int32_t* tree_per_ray, int64_t n,
                           const double* org, const double* dir, const double* tmax, int orientation,
                           const int32_t* excl, int k, int precision, uint64_t* counters);

/* RayTracer::point_in_volume (ray_tracing_interface.h:76-79; Embree
 * ray_tracer.cpp:259-287).  dir NULL -> (1/sqrt 2, 1/sqrt 2, 0). */
int xdg_hip_point_in_volume(xdg_hip_ctx* ctx, int32_t tree, int64_t n, const double* pts, const double* dir,
                            const int32_t* excl, int k, int precision, uint8_t* inside_out, void* stream);

/* XDG::find_volume (src/xdg.cpp:113-128): the first volume in the given
 * order (skipping ipc_volume) whose point_in_volume is true, else
 * ipc_volume.  vol_trees/vol_ids are host arrays of nvol entries. */
int xdg_hip_find_volume(xdg_hip_ctx* ctx, int nvol, const int32_t* vol_trees, const int32_t* vol_ids,
                        int32_t ipc_volume, int64_t n, const double* pts, const double* dir, int precision,
                        int32_t* vol_out, void* stream);

/* ---- overlap check (tools/overlap.cpp) ------------------------------------
 * check_location_for_overlap (overlap.cpp:9-61) asks point_in_volume of
 * every volume of the model at each location.  For each location i (device
 * pts[3n], dir[3n]: its own ray direction, no exclusions):
 * count_out[i] = number of volumes in the list whose point_in_volume is
 * true, ids_out[max_k*i ...] = the first max_k of them (volume ids, list
 * order).  vol_trees/vol_ids are host arrays of nvol entries (the model's
 * volumes, the implicit complement included as MeshManager::volumes() has
 * it). */
int xdg_hip_point_in_volumes(xdg_hip_ctx* ctx, int nvol, const int32_t* vol_trees, const int32_t* vol_ids, int64_t n,
                             const double* pts, const double* dir, int max_k, int precision, int32_t* count_out,
                             int32_t* ids_out, void* stream);

/* check_along_edge (overlap.cpp:229-265): for each edge ray i (device
 * org[3n], dir[3n], len[n]), XDG::ray_fire(vol, org, dir, len) (EXITING, no
 * exclusions) over the volumes of the list in order, skipping the two
 * volume ids skip[2i], skip[2i+1] (the parents of the edge's surface); the
 * first volume with a hit ends the search.  Outputs: t_out (INFTY when no
 * volume is hit), surf_out (hit surface or ID_NONE), vol_out (the volume
 * whose tree was hit, or ID_NONE). */
int xdg_hip_edge_rays(xdg_hip_ctx* ctx, int nvol, const int32_t* vol_trees, const int32_t* vol_ids, int64_t n,
                      const double* org, const double* dir, const double* len, const int32_t* skip, int precision,
                      double* t_out, int32_t* surf_out, int32_t* vol_out, void* stream);

/* RayTracer::occluded (ray_tracing_interface.h:112-115; Embree
 * ray_tracer.cpp:347-370).  dist_out (nullable): -INFTY on hit, INFTY on
 * miss, as the reference's `dist` argument. */
int xdg_hip_occluded(xdg_hip_ctx* ctx, int32_t tree, int64_t n, const double* org, const double* dir,
                     int precision, uint8_t* hit_out, double* dist_out, void* stream);

/* RayTracer::closest (ray_tracing_interface.h:109-110; Embree
 * ray_tracer.cpp:328-345): distance and the nearest FACE id. */
int xdg_hip_closest(xdg_hip_ctx* ctx, int32_t tree, int64_t n, const double* pts, double* dist_out,
                    int32_t* face_out, void* stream);

/* RayTracer::find_element(tree, point) (ray_tracing_interface.h:98-107;
 * Embree ray_tracer.cpp:226-257). */
int xdg_hip_find_element(xdg_hip_ctx* ctx, int32_t element_tree, int64_t n, const double* pts, int32_t* elem_out,
                         void* stream);

/* MeshManager::next_element (mesh_manager_interface.cpp:150-200), batched:
 * no exit face -> next = -1, dist = INFTY (the reference indexes face -1). */
int xdg_hip_next_element(xdg_hip_ctx* ctx, int32_t element_tree, int64_t n, const int32_t* elem, const double* r,
                         const double* u, int32_t* next_out, double* dist_out, void* stream);

/* XDG::segments(start, end) (src/xdg.cpp:142-198) over the element tree,
 * fused with the per-element track-length tally of tally_segments
 * (tools/tally_segments.h:51-83).  Per segment: count of (element, length)
 * entries (zero-length entries kept, mesh_manager_interface.cpp:125) and a
 * 64-bit sequence hash (h = 0x6a09e667f3bcc909; per entry
 * h = mix(h ^ elem); h = mix(h ^ bits(len)); mix = splitmix64 finaliser).
 * tally_fx (device, nullable, one per element): += llrint(len * 2^40).
 * seq_off (device, nullable, n+1 prefix offsets) with seq_elem/seq_len:
 * explicit sequences.  Asynchronous on stream; a walk that hits its step
 * bound (2^16 steps) sets bit 0 of the status word (xdg_hip_status). */
int xdg_hip_tally_segments(xdg_hip_ctx* ctx, int32_t element_tree, int64_t n, const double* p0, const double* p1,
                           int64_t* tally_fx, int32_t* count_out, uint64_t* hash_out, const int64_t* seq_off,
                           int32_t* seq_elem, double* seq_len, void* stream);

/* XDG::segments(volume, start, end) (src/xdg.cpp:200-227), outputs as in
 * xdg_hip_tally_segments.  element_tree is the volume's element tree and
 * surface_tree its surface tree.  A start outside the mesh enters it along
 * the ENTERING hit on the volume's surface (no hit: no entries); the walk
 * is MeshManager::walk_elements(elem, start, end)
 * (mesh_manager_interface.cpp:139-147) from the entry point, with no
 * implicit-complement re-entry.  When no element lies behind the entry hit
 * (the reference warns) the segment has no entries and bit 1 of the status
 * word is set. */
int xdg_hip_tally_segments_volume(xdg_hip_ctx* ctx, int32_t element_tree, int32_t surface_tree, int64_t n,
                                  const double* p0, const double* p1, int64_t* tally_fx, int32_t* count_out,
                                  uint64_t* hash_out, const int64_t* seq_off, int32_t* seq_elem, double* seq_len,
                                  void* stream);

/* The element-by-element particle walk of tools/walk_elements.h:53-117
 * (walk_elements), one particle per lane, particles [first, first + n).
 * Start: points uniform in [lo, hi] (host arrays of 3) until find_element
 * finds an element (1000 tries), an isotropic direction, normalized.  Each
 * event: MeshManager::next_element, a collision distance -log(1-u)*mfp; a
 * collision scatters isotropically, otherwise the particle moves into the
 * next element; a particle leaving the mesh fires at the implicit
 * complement (EXITING, that event's hit faces excluded) and re-enters at
 * find_element(hit + u*TINY_BIT), or ends.  Random numbers: Philox4x32-10,
 * key (seed, stream_id), counter (particle, 0, draw, 0), in place of the
 * tool's drand48 and rand_dir.  Nullable device outputs per particle:
 * distance travelled (f64, re-entry flights included), events (i32),
 * status (XDG_HIP_TR_LEAKED: left the mesh; XDG_HIP_TR_CAPPED: max_events
 * reached; XDG_HIP_TR_LOST: more than 64 re-entry rays in one event;
 * XDG_HIP_TR_ERROR: no start found), a hash over the (element, length)
 * pieces; tally_fx (nullable, per element) += llrint(len * 2^40). */
int xdg_hip_mesh_walk(xdg_hip_ctx* ctx, int32_t element_tree, uint64_t first, int64_t n, uint32_t seed,
                      uint32_t stream_id, const double* lo, const double* hi, double mfp, int32_t max_events,
                      double* distance_out, int32_t* events_out, uint8_t* status_out, uint64_t* hash_out,
                      int64_t* tally_fx, void* stream);

/* Synchronize the device, return and clear the status word (bits below). */
#define XDG_HIP_STATUS_STEP_CAP 1u      /* a segment walk hit its step bound */
#define XDG_HIP_STATUS_NO_ENTRY_ELEM 2u /* a volume segment found no element behind its entry hit */
#define XDG_HIP_STATUS_TALLY_RANGE 4u   /* a fixed-point tally reached 2^62 in magnitude */
int xdg_hip_status(xdg_hip_ctx* ctx, uint32_t* flags);

/* Evaluate one device geometry primitive per item, for parity checks
 * against the reference's own functions (device in/out arrays):
 *   0 plucker_ray_tri_intersect (plucker.h:64-161): 19 doubles
 *     (v0 v1 v2 o d tmax tmin use_orientation orientation) -> (hit, t)
 *   1 face_normal (mesh_manager_interface.cpp:216-220): 9 -> 3
 *   2 closest_location_on_triangle (closest.cpp:54-183): 12 -> 3
 *   3 plucker_tet_containment_test (tetrahedron_contain.cpp:11-43):
 *     15 (p v0 v1 v2 v3) -> 1
 *   4 sqrt: 1 -> 1;  5 division a/b: 2 -> 1 */
int xdg_hip_eval_primitive(xdg_hip_ctx* ctx, int kind, int64_t n, const double* in, double* out, void* stream);

/* ---- batched particle random walk ----------------------------------------
 * Replaces the serial history loop of tools/particle_sim.h:136-154
 * (transport_particles) and its Particle methods (:34-130): each history
 * samples a start uniform in [lo, hi] and an isotropic direction, finds its
 * volume (XDG::find_volume, xdg.cpp:113-128), then repeats ray_fire
 * (EXITING, excluding the last 8 faces hit since the last collision,
 * ray_tracer.cpp:324), an exponential flight -log(1-u)*mfp, a track-length
 * tally in the current volume, and either a collision (absorbed with
 * probability p_absorb, else a new isotropic direction) or a surface
 * crossing (MeshManager::next_volume, mesh_manager_interface.cpp:202-214).
 * A particle entering volume ipc leaks (ipc_graveyard).  A surface
 * crossing first reads the surface's boundary condition
 * (MeshManager::get_surface_property, mesh_manager_interface.cpp:99-104;
 * Particle::cross_surface, particle_sim.h:85-130): vacuum ends the history,
 * reflecting mirrors the direction about face_normal of the hit face
 * (u - n * (2 (n.u) / |n|), normalized; XDG::surface_normal, xdg.cpp:272-290)
 * and keeps only that face in the history, transmission (the default)
 * moves to the next volume.  Random numbers:
 * Philox4x32-10, key (seed, stream), counter (history, 0, draw, 0).
 * Additions to the reference driver: Philox, sampled starts, absorption,
 * the event cap.  One persistent launch runs all histories; lanes claim
 * `chunk` histories at a time from a device counter. */
typedef struct {
  int32_t nvol;               /* volumes with trees, in find_volume order */
  const int32_t* vol_ids;     /* host, nvol */
  const int32_t* vol_trees;   /* host, nvol: surface tree of each volume */
  int32_t ipc;                /* implicit complement volume id */
  int32_t nsurf_ids;          /* sense tables cover surface ids [0, nsurf_ids) */
  const int32_t* sense_fwd;   /* host: forward parent volume per surface id */
  const int32_t* sense_rev;   /* host: reverse parent volume per surface id */
  uint32_t seed, stream;
  double lo[3], hi[3];        /* start box */
  double mfp;                 /* mean free path */
  double p_absorb;            /* absorption probability per collision */
  int32_t max_events;         /* event cap per history */
  const uint8_t* surf_bc;     /* host, nullable, nsurf_ids: XDG_HIP_BC_* per
                                 surface id (NULL: all transmission) */
} xdg_hip_transport_problem;

/* Surface boundary conditions */
#define XDG_HIP_BC_TRANSMISSION 0
#define XDG_HIP_BC_VACUUM 1
#define XDG_HIP_BC_REFLECTING 2

/* Per-history status codes */
#define XDG_HIP_TR_ABSORBED 1
#define XDG_HIP_TR_LEAKED 2
#define XDG_HIP_TR_LOST 3     /* ray_fire missed (reference: fatal_error) */
#define XDG_HIP_TR_STUCK 4    /* hit at distance 0 (reference: fatal_error) */
#define XDG_HIP_TR_CAPPED 5   /* max_events reached */
#define XDG_HIP_TR_ERROR 6    /* topology error (volume not a parent, ...) */
#define XDG_HIP_TR_VACUUM 7   /* crossed a vacuum surface */

int xdg_hip_transport_create(xdg_hip_ctx* ctx, const xdg_hip_transport_problem* p, int32_t* handle);
int xdg_hip_transport_destroy(xdg_hip_ctx* ctx, int32_t handle);
/* Launch configuration (0 keeps a value): workgroup size (64..256), traversal stack
 * entries held in LDS (the rest spill to a global area; capped at what the
 * trees need), histories claimed per lane at a time, idle lanes a wave
 * waits for before it starts new histories (1..64).  stack_need = entries
 * the deepest tree of the problem can need. */
int xdg_hip_transport_set_config(xdg_hip_ctx* ctx, int32_t handle, int block, int lds_stack, int chunk, int refill);
int xdg_hip_transport_get_config(xdg_hip_ctx* ctx, int32_t handle, int* block, int* lds_stack, int* chunk,
                                 int* refill, int* stack_need);
/* Histories [first, first + n): device outputs tally_fx (nvol, per volume
 * slot, int64 two's complement of sum(llrint(len * 2^32))), and nullable
 * per-history events (int32), status (uint8), sequence hash (uint64) over
 * (volume, length) pairs.  Asynchronous on stream; allocates nothing after
 * the first run of a configuration, so it may be captured. */
int xdg_hip_transport_run(xdg_hip_ctx* ctx, int32_t handle, uint64_t first, int64_t n, int precision,
                          uint64_t* tally_fx, int32_t* events, uint8_t* status, uint64_t* hash, void* stream);
/* The same launch captured into an instantiated HIP graph (opaque handle):
 * replay with xdg_hip_graph_launch, free with xdg_hip_graph_destroy (before
 * xdg_hip_destroy).  The graph holds device pointers of the capture time:
 * the output arrays (kept alive by the caller), the tree table and trees,
 * and the transport's buffers.  Building, releasing or attaching any tree
 * on the context, destroying the transport, or a launch that needs a
 * larger stack spill area (transport_set_config / autotune followed by a
 * run) or a larger start buffer (more histories per slice, or the
 * "transport_slice" / "transport_items" options raised, followed by a run)
 * invalidates it: xdg_hip_graph_launch then returns
 * XDG_HIP_ERR_INVALID without launching; capture again. */
int xdg_hip_transport_capture(xdg_hip_ctx* ctx, int32_t handle, uint64_t first, int64_t n, int precision,
                              uint64_t* tally_fx, int32_t* events, uint8_t* status, uint64_t* hash, void** graph);
int xdg_hip_graph_launch(xdg_hip_ctx* ctx, void* graph, void* stream);
int xdg_hip_graph_destroy(xdg_hip_ctx* ctx, void* graph);
/* Times block x lds_stack x chunk x refill candidates on histories
 * [first, first + sample) and keeps the fastest configuration
 * (synchronous). */
int xdg_hip_transport_autotune(xdg_hip_ctx* ctx, int32_t handle, uint64_t first, int64_t sample, int precision,
                               double* best_ms);

/* ---- adjoint of the ray distance ------------------------------------------
 * No counterpart in the reference.  For the rays of a ray_fire batch (t,
 * face as returned; face ids must index the rows of tris), accumulates
 * dldv[3 nv] += sum_i dldt[i] dt_i/dverts (FP64 atomics) and writes
 * dldo[3 n] = dldt dt/do, dldd[3 n] = dldt dt/dd (zero for misses), with
 * t = n.(v0 - o)/(n.d), n = (v1 - v0) x (v2 - v0).  All arrays device,
 * dldv/dldo/dldd nullable.  Asynchronous on stream.  Batches of >= 65536
 * rays with dldv use the context's ray-order scratch (see the
 * "adjoint_sort" option; ordered across streams like the other users). */
int xdg_hip_ray_distance_adjoint(xdg_hip_ctx* ctx, int64_t n, const double* org, const double* dir,
                                 const double* t, const int32_t* face, const double* dldt, int64_t nv,
                                 const double* verts, int64_t ntri, const int32_t* tris, double* dldv,
                                 double* dldo, double* dldd, void* stream);

/* ---- device memory helpers for callers without their own allocator ----- */
int xdg_hip_malloc(xdg_hip_ctx* ctx, size_t bytes, void** ptr);
int xdg_hip_free(xdg_hip_ctx* ctx, void* ptr);
int xdg_hip_memcpy_htod(xdg_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int xdg_hip_memcpy_dtoh(xdg_hip_ctx* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* XDG_HIP_H */
