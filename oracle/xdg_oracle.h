/*
 * xdg_oracle.h -- CPU FP64 oracle for the XDG geometry-query hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (xdg_amd/, libxdg_hip.so) never links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain C with the same
 * floating-point operation order (no FMA contraction: build with
 * -ffp-contract=off, no -march, no -ffast-math).  Citations are to the
 * reference tree (Waqar-ukaea/xdg).
 *
 * Parity status per function (what pins it; see tests/test_oracle_*.py):
 *   xo_plucker_ray_tri_intersect  golden vectors from the compiled reference
 *                                 header (random cases, and the vertex and
 *                                 edge-midpoint rays of an icosphere, where
 *                                 the reference leaks two rays) + MeshMock
 *                                 known answers + closed form
 *   xo_face_normal                golden vectors (compiled face_common.h)
 *   xo_closest_location_on_triangle / xo_determine_region
 *                                 golden vectors (compiled closest.cpp) +
 *                                 test_closest.cpp known answers
 *   xo_tet_contains               reference booleans (test_tet_containment.cpp,
 *                                 test_tet_intersection.cpp) + barycentric
 *                                 identities (lambda(sum w_i v_i) = w to 1e-9,
 *                                 vertices -> unit vectors, vertex permutations
 *                                 permute lambda; a planted transposed inverse
 *                                 fails).  The lambda bits at face ties depend
 *                                 on the absent `linalg` library's inverse():
 *                                 PARITY UNPINNED at exact-tie bits.
 *   xo_next_element / xo_walk_elements
 *                                 golden sequences from the compiled
 *                                 mesh_manager_interface.cpp on MeshMock
 *   xo_ray_fire / xo_point_in_volume / xo_occluded / xo_closest
 *                                 MeshMock known answers (test_ray_fire.cpp,
 *                                 test_point_in_volume.cpp, test_occluded.cpp,
 *                                 test_closest.cpp) + brute force == BVH
 *   xo_find_element               analytic Kuhn-cell location on the
 *                                 unjittered cube (unique element away from
 *                                 faces; lowest index among the exact ties of
 *                                 dyadic points on shared faces, edges,
 *                                 vertices) + tree == brute force on a
 *                                 jittered mesh, points a few ulps outside
 *                                 the boundary included; planted mistakes
 *                                 (first-visited element, undilated boxes)
 *                                 fail (tests/test_oracle_find_element.py)
 *   xo_segments                   sum(len) == chord invariant (random, and
 *                                 along lattice edges, planes and vertices of
 *                                 an unjittered Kuhn mesh) + MeshMock walk
 *                                 golden sequences
 *   xo_segments_volume            walk_elements(elem, start, end) sequences of
 *                                 the compiled reference on MeshMock (starts
 *                                 inside and outside the mesh); slab closed
 *                                 forms of the entry on the unit cube (sum to
 *                                 |end - entry|, to the chord, none on a miss)
 *   xo_philox4x32_10              Random123 known-answer vectors
 *   xo_log                        glibc log() within 2 ulp
 *   xo_transport                  no reference output exists
 *                                 (particle_sim.h has no absorption and uses
 *                                 drand48); pinned by the homogeneous-medium
 *                                 closed forms (mean track mfp/p, mean
 *                                 collisions 1/p), region-volume tally
 *                                 fractions and thread-count independence;
 *                                 boundary conditions: a reflecting cube
 *                                 unfolds each history to a straight line
 *                                 (events = integer planes crossed + 1), a
 *                                 reflecting sphere keeps the mean track
 *                                 mfp/p with no leak, vacuum spheres tally
 *                                 nothing for outside starts
 *   xo_mesh_walk                  no reference output exists (drand48,
 *                                 rand_dir); pinned by straight flights
 *                                 without collisions: slab exit distance of
 *                                 the cube, and exit + gap + chord through a
 *                                 second block (re-entry); tally conservation;
 *                                 thread-count independence
 *   xo_adjoint                    central finite differences of the
 *                                 oracle's own ray_fire distances
 *   pyoracle.OverlapQueries       the overlap check's queries composed of
 *                                 xo_point_in_volume and xo_ray_fire; pinned
 *                                 by a brute-force Moeller-Trumbore
 *                                 classification of every vertex location
 *                                 on four box models
 */
#ifndef XDG_ORACLE_H
#define XDG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XO_ID_NONE (-1)

/* HitOrientation (include/xdg/shared_enums.h:11-15) */
#define XO_ANY (-1)
#define XO_EXITING 0
#define XO_ENTERING 1

/* ---- primitive geometry ------------------------------------------------ */

/* plucker_ray_tri_intersect (include/xdg/geometry/plucker.h:64-161).
 * v: 9 doubles (v0,v1,v2).  Returns hit flag, writes t (0.0 on miss). */
int xo_plucker_ray_tri_intersect(const double v[9], const double o[3],
                                 const double d[3], double tmax, double tmin,
                                 int use_orientation, int orientation,
                                 double* t);

/* plucker_edge_test (plucker.h:45-62) */
double xo_plucker_edge_test(const double va[3], const double vb[3],
                            const double ray[3], const double ray_normal[3]);

/* MeshManager::face_normal (src/mesh_manager_interface.cpp:216-220) */
void xo_face_normal(const double v[9], double n[3]);

/* closest_location_on_triangle / determine_region (src/geometry/closest.cpp) */
void xo_closest_location_on_triangle(const double v[9], const double p[3],
                                     double out[3]);
int xo_determine_region(const double v[9], const double p[3]);

/* plucker_tet_containment_test (src/tetrahedron_contain.cpp:11-43);
 * v: 12 doubles (v0..v3).  lambda_out (nullable) receives l0..l3. */
int xo_tet_contains(const double p[3], const double v[12], double* lambda_out);

/* ---- surface trees ------------------------------------------------------ */

typedef struct xo_tree xo_tree;

/* One surface tree = the triangles of one volume (EmbreeRayTracer::
 * create_surface_tree, src/embree/ray_tracer.cpp:50-105).  verts9: 9 doubles
 * per triangle; flip[i] = 1 when this tree is the reverse-sense parent of the
 * triangle's surface (triangle_intersect.cpp:84).  bump: box dilation
 * (ray_tracing_interface.cpp:20-24).  bvh != 0 builds a host BVH (exact same
 * results as brute force; boxes are conservative). */
xo_tree* xo_tree_create(int64_t n, const double* verts9, const int32_t* face_id,
                        const int32_t* surf_id, const uint8_t* flip,
                        double bump, int bvh);
void xo_tree_destroy(xo_tree* t);
int64_t xo_tree_size(const xo_tree* t);

/* EmbreeRayTracer::ray_fire (src/embree/ray_tracer.cpp:290-326) + callback
 * TriangleIntersectionFunc (src/triangle_intersect.cpp:52-109).
 * tmax may be NULL (INFTY).  excl: k slots per ray, -1 = empty (NULL = none).
 * Miss: t = DBL_MAX, surf = face = -1.  Tie rule: among bitwise-equal t the
 * smallest face id wins.  nthreads <= 0 uses OpenMP default. */
void xo_ray_fire(const xo_tree* t, int64_t n, const double* org,
                 const double* dir, const double* tmax, int orientation,
                 const int32_t* excl, int k, int brute, int nthreads,
                 double* t_out, int32_t* surf_out, int32_t* face_out);

/* EmbreeRayTracer::point_in_volume (ray_tracer.cpp:259-287).  dir NULL uses
 * (1/sqrt(2), 1/sqrt(2), 0). */
void xo_point_in_volume(const xo_tree* t, int64_t n, const double* pts,
                        const double* dir, const int32_t* excl, int k,
                        int brute, int nthreads, uint8_t* inside);

/* EmbreeRayTracer::occluded (ray_tracer.cpp:347-370) +
 * TriangleOcclusionFunc (triangle_intersect.cpp:139-161).
 * dist_out: -DBL_MAX on hit, DBL_MAX on miss. */
void xo_occluded(const xo_tree* t, int64_t n, const double* org,
                 const double* dir, int brute, int nthreads, uint8_t* hit,
                 double* dist_out);

/* EmbreeRayTracer::closest (ray_tracer.cpp:328-345) +
 * TriangleClosestFunc (triangle_intersect.cpp:111-137).  Tie rule: smallest
 * face id among bitwise-equal distances.  Miss: DBL_MAX, -1. */
void xo_closest(const xo_tree* t, int64_t n, const double* pts, int brute,
                int nthreads, double* dist_out, int32_t* face_out);

/* XDG::find_volume (src/xdg.cpp:113-128): first volume (in the given order,
 * skipping ipc) whose point_in_volume is true, else ipc. */
void xo_find_volume(const xo_tree* const* trees, const int32_t* vol_ids,
                    int nvol, int32_t ipc, int64_t n, const double* pts,
                    const double* dirs, int brute, int nthreads,
                    int32_t* vol_out);

/* ---- tetrahedral meshes -------------------------------------------------- */

typedef struct xo_tetmesh xo_tetmesh;

/* face_order: 12 ints, the 4 faces of a tet as local vertex triples
 * (MOAB canonical {0,1,3},{1,2,3},{2,0,3},{0,2,1}:
 *  include/xdg/moab/direct_access.h:181-183; MeshMock
 *  {0,1,2},{0,2,3},{0,3,1},{1,3,2}: tests/mesh_mock.h:108-115).
 * adj: 4 per tet, neighbour across face i, -1 on the boundary. */
xo_tetmesh* xo_tetmesh_create(int64_t nv, const double* verts, int64_t ne,
                              const int32_t* tets, const int32_t* adj,
                              const int32_t face_order[12], int bvh);
void xo_tetmesh_destroy(xo_tetmesh* m);

/* RayTracer::find_element over the element tree (ray_tracer.cpp:226-257,
 * tetrahedron_contain.cpp:47-114): lowest element index whose containment
 * test passes, else -1. */
void xo_find_element(const xo_tetmesh* m, int64_t n, const double* pts,
                     int brute, int nthreads, int32_t* elem_out);

/* Planted mistakes, for tests that show a check can fail (0 = none, the
 * default).  XO_FAULT_FIRST_VISITED: the element-tree find_element returns
 * the first containing element in tree-visit order instead of the lowest
 * index.  XO_FAULT_NO_DILATION: element trees created while it is set use
 * undilated element boxes.  XO_FAULT_TET_TRANSPOSE: tet_contains
 * multiplies by the transpose of the inverse.  Process-wide; not thread-safe against queries
 * running at the same time. */
#define XO_FAULT_NONE 0
#define XO_FAULT_FIRST_VISITED 1
#define XO_FAULT_NO_DILATION 2
#define XO_FAULT_TET_TRANSPOSE 3 /* tet_contains applies the transposed inverse */
void xo_set_fault(int code);

/* MeshManager::next_element (mesh_manager_interface.cpp:150-200).
 * No exit face -> next = -1, dist = DBL_MAX (reference: UB). */
void xo_next_element(const xo_tetmesh* m, int32_t elem, const double r[3],
                     const double u[3], int32_t* next, double* dist);

/* MeshManager::walk_elements (mesh_manager_interface.cpp:107-136).
 * Writes up to cap entries, returns the number produced (may exceed cap). */
int64_t xo_walk_elements(const xo_tetmesh* m, int32_t elem,
                         const double r[3], const double u[3],
                         double distance, int64_t cap, int32_t* elems,
                         double* lens);

/* XDG::segments(start, end) (src/xdg.cpp:142-198) with the implicit-
 * complement re-entry branch.  ipc: surface tree of the mesh boundary
 * triangles as seen from the implicit complement (may be NULL: then the
 * branch returns, as a ray_fire miss does).  bface_elem[face_id] = element
 * adjacent to that boundary face (get_boundary_face_element).
 * Per segment: count, 64-bit sequence hash (xo_seq_hash), and optionally
 * fixed-point tallies (tally_fx, scale 2^40, int64 per element) and the
 * explicit sequences (seq_off: prefix offsets, NULL = skip). */
void xo_segments(const xo_tetmesh* m, const xo_tree* ipc,
                 const int32_t* bface_elem, int64_t n, const double* p0,
                 const double* p1, int nthreads, int32_t* count,
                 uint64_t* hash, int64_t* tally_fx, const int64_t* seq_off,
                 int32_t* seq_elem, double* seq_len);

/* XDG::segments(volume, start, end) (src/xdg.cpp:200-227).  vol: the
 * volume's surface tree; outputs as in xo_segments; flags[i] (nullable):
 * bit 0 = step bound hit, bit 1 = entering hit with no element behind it. */
void xo_segments_volume(const xo_tetmesh* m, const xo_tree* vol, int64_t n,
                        const double* p0, const double* p1, int nthreads,
                        int32_t* count, uint64_t* hash, int64_t* tally_fx,
                        const int64_t* seq_off, int32_t* seq_elem,
                        double* seq_len, uint32_t* flags);

/* Sequence hash used for per-segment parity: h0 = XO_HASH_SEED; per entry
 * h = mix64(h ^ (uint32)elem); h = mix64(h ^ bits(len)).  mix64 is the
 * splitmix64 finaliser. */
#define XO_HASH_SEED 0x6a09e667f3bcc909ULL
uint64_t xo_mix64(uint64_t x);

/* ---- random numbers (Philox4x32-10, Salmon et al. SC'11) --------------- */
void xo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
/* Two 53-bit uniforms in [0,1) from one Philox block of counter
 * (item_lo, item_hi, draw, 0), key (seed, stream). */
void xo_uniform2(uint32_t seed, uint32_t stream, uint64_t item, uint32_t draw,
                 double* ua, double* ub);
/* Natural log from the exponent/mantissa split and an atanh series;
 * identical bits on host and device (plain IEEE ops, no FMA). */
double xo_log(double x);

/* ---- batched particle random walk ---------------------------------------- */

/* Transport_particles (tools/particle_sim.h:136-154) as a batch of
 * independent histories, with these additions: Philox streams instead of
 * drand48, starts uniform in a box with isotropic directions, absorption
 * with probability p_absorb at each collision, leakage when a particle
 * enters the implicit complement (ipc_graveyard), an event cap.
 *
 * Per history h (Philox counter item = first + i, key (seed, stream); the
 * draw counter dc starts at 0 and advances one block at a time):
 *   start      blocks 0,1: r = lo + (hi-lo)*(ua0, ub0, ua1); dc = 2
 *   direction  Marsaglia rejection, one block per attempt
 *   volume     XDG::find_volume(r, u) over the given volume order
 *   event      ray_fire(tree(vol), r, u, INFTY, EXITING, history) with the
 *              last XO_TRANSPORT_HISTORY hit faces since the last collision
 *              excluded (the hit face is appended, ray_tracer.cpp:324);
 *              one block: dcol = -log(1 - ua) * mfp (xo_log); s = min(dcol,
 *              t_hit); r += s * u; tally[vol] += s; if dcol < t_hit:
 *              absorbed when ub < p_absorb, else new direction, history
 *              cleared; otherwise the hit surface's boundary condition
 *              (particle_sim.h:85-130): vacuum ends the history
 *              (XO_TR_VACUUM); reflecting mirrors u about face_normal of the
 *              hit face (u - n * (2 (n.u) / |n|), then normalize) and keeps
 *              only that face in the history; transmission moves to
 *              vol = next_volume(vol, surf), leaked when that is ipc.
 * Outputs per history: event count, status (XO_TR_*), sequence hash over
 * (vol, s) as in xo_segments.  tally_fx[slot] += llrint(s * 2^32) (int64,
 * order independent). */
#define XO_TRANSPORT_HISTORY 8
#define XO_TR_ALIVE 0
#define XO_TR_ABSORBED 1
#define XO_TR_LEAKED 2
#define XO_TR_LOST 3
#define XO_TR_STUCK 4
#define XO_TR_CAPPED 5
#define XO_TR_ERROR 6
#define XO_TR_VACUUM 7
/* surface boundary conditions (MeshManager::get_surface_property,
 * mesh_manager_interface.cpp:99-104; default "transmission") */
#define XO_BC_TRANSMISSION 0
#define XO_BC_VACUUM 1
#define XO_BC_REFLECTING 2

typedef struct {
  int nvol;                   /* volumes with surface trees, find order */
  const xo_tree* const* trees;
  const int32_t* vol_ids;
  int32_t ipc;
  int nsurf_ids;              /* sense tables are indexed by surface id */
  const int32_t* sense_fwd;
  const int32_t* sense_rev;
  uint32_t seed, stream;
  double lo[3], hi[3];
  double mfp, p_absorb;
  int32_t max_events;
  const uint8_t* surf_bc;     /* per surface id (nsurf_ids), NULL = all
                                 transmission */
} xo_transport_problem;

void xo_transport(const xo_transport_problem* p, uint64_t first, int64_t n,
                  int nthreads, int64_t* tally_fx, int32_t* events,
                  uint8_t* status, uint64_t* hash);

/* Element-by-element particle walk: tools/walk_elements.h:53-117 with
 * Philox draws in place of drand48 / rand_dir.  Per particle (item = first
 * + i, key (seed, stream), draw counter from 0):
 *   start      blocks (2 per attempt): r = lo + (hi-lo)*(ua0, ub0, ua1) until
 *              find_element(r) finds an element (at most
 *              XO_MESH_WALK_STARTS attempts, else XO_TR_ERROR)
 *   direction  Marsaglia rejection, one block per attempt, then
 *              Vec3da::normalize
 *   event      (next, exit) = next_element(e, r, u); one block: coll =
 *              -log(1 - ua) * mfp; coll < exit: r += u*coll, new direction
 *              (not normalized); else r += u*exit, e = next.  While e is
 *              ID_NONE: ray_fire(ipc, r, u, INFTY, EXITING, prims) (the hit
 *              face is appended), stop on a miss, else r += u*t and e =
 *              find_element(r + u*TINY_BIT); prims cleared after the event.
 * Outputs per particle: total distance, events, status (XO_TR_LEAKED: left
 * the mesh; XO_TR_CAPPED at max_events; XO_TR_LOST: more than
 * XO_MESH_WALK_REENTRIES re-entry rays in one event), a hash over the
 * (element, length) pieces as in xo_segments; tally_fx (nullable) +=
 * llrint(len * 2^40) per element. */
#define XO_MESH_WALK_STARTS 1000
#define XO_MESH_WALK_REENTRIES 64
void xo_mesh_walk(const xo_tetmesh* m, const xo_tree* ipc, const double lo[3],
                  const double hi[3], uint32_t seed, uint32_t stream,
                  double mfp, int32_t max_events, uint64_t first, int64_t n,
                  int nthreads, double* distance, int32_t* events,
                  uint8_t* status, uint64_t* hash, int64_t* tally_fx);

/* ---- adjoint of the ray distance ----------------------------------------- */

/* For rays that hit triangle f = face[i] of the mesh (verts, tris) at
 * distance t[i], t = n.(v0 - o) / (n.d) with n = (v1 - v0) x (v2 - v0), so
 * with D = n.d and q = v0 - (o + t d):
 *   dt/do = -n / D,  dt/dd = -t n / D,
 *   dt/dv1 = (e2 x q) / D,  dt/dv2 = (q x e1) / D,
 *   dt/dv0 = (n - e2 x q - q x e1) / D      (e1 = v1 - v0, e2 = v2 - v0).
 * dldv (3 nv) accumulates sum_i dldt[i] dt_i/dv serially; dldo, dldd (3 n)
 * are written (zero for misses).  No reference counterpart. */
void xo_adjoint(int64_t nv, const double* verts, int64_t ntri,
                const int32_t* tris, int64_t n, const double* org,
                const double* dir, const double* t, const int32_t* face,
                const double* dldt, double* dldv, double* dldo, double* dldd);

#ifdef __cplusplus
}
#endif

#endif /* XDG_ORACLE_H */
