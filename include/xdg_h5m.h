/*
 * xdg_h5m.h -- DAGMC/MOAB .h5m geometry files, read into and written from
 * the flat arrays the xdg_hip tree builders take (libxdg_h5m.so, host only).
 *
 * Replaces the file-ingestion half of the reference's MOABMeshManager
 * (src/moab/mesh_manager.cpp:37-107 init, :489-590 parse_metadata) and the
 * storage its MBDirectAccess reads (include/xdg/moab/direct_access.h:
 * 185-349), for the data the hot path needs: vertex coordinates, the
 * triangles of each surface set, the tetrahedra of each volume set, the
 * GEOM_SENSE_2 (forward, reverse) parent volumes of each surface, and the
 * NAME of each group set ("mat:...", "boundary:...") with its members.  Tag
 * names follow include/xdg/moab/tag_conventions.h: GEOM_DIMENSION,
 * GLOBAL_ID, CATEGORY, NAME, GEOM_SENSE_2.
 *
 * File layout: MOAB's native HDF5 layout under /tstt (nodes/coordinates,
 * elements/<type>/connectivity, sets/list|contents|children|parents,
 * sparse tags/<name>/id_list|values), with one file id space: nodes first,
 * then elements, then sets.  No reference .h5m file exists to check this
 * against, so byte-level compatibility with MOAB-written files is unpinned;
 * what is tested is that models round-trip through write and read unchanged.
 *
 * Conventions: ids are the sets' GLOBAL_IDs (int32); vertex and element
 * indices are 0-based positions in the arrays; -1 = none.  Functions return
 * 0 or a negative code; xdg_h5m_last_error() describes the last failure on
 * the calling thread.
 */
#ifndef XDG_H5M_H
#define XDG_H5M_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XDG_H5M_OK 0
#define XDG_H5M_ERR_IO (-1)      /* HDF5 open/read/write failure */
#define XDG_H5M_ERR_FORMAT (-2)  /* missing or inconsistent /tstt data */
#define XDG_H5M_ERR_INVALID (-3) /* bad argument */

#define XDG_H5M_NAME_SIZE 32 /* NAME and CATEGORY tag values (tag_conventions.h) */

typedef struct xdg_h5m_model xdg_h5m_model;

typedef struct {
  int64_t n_verts;
  int64_t n_tris;    /* triangles over all surfaces */
  int64_t n_tets;    /* tetrahedra over all volumes */
  int32_t n_surfaces;
  int32_t n_volumes;
  int32_t n_groups;
  int32_t reserved;
  int64_t n_group_members;
} xdg_h5m_counts;

/* Reads a file.  Surfaces and volumes are the entity sets with
 * GEOM_DIMENSION 2 and 3, in file order (as MOABMeshManager::init lists
 * them); groups are the sets with CATEGORY "Group". */
int xdg_h5m_read(const char* path, xdg_h5m_model** out);
void xdg_h5m_free(xdg_h5m_model* m);
int xdg_h5m_get_counts(const xdg_h5m_model* m, xdg_h5m_counts* c);
/* Caller-allocated arrays, sized from the counts (any may be NULL):
 *   verts[3 n_verts]; vert_ids[n_verts]          file ids of the vertices
 *   surf_ids[n_surfaces]; surf_tri_off[n_surfaces + 1]
 *   tris[3 n_tris]  vertex indices, surface by surface; tri_ids[n_tris] file ids
 *   surf_sense[2 n_surfaces]  GLOBAL_IDs of the forward and reverse parent
 *                             volumes (-1 none: the implicit complement's side)
 *   vol_ids[n_volumes]; vol_tet_off[n_volumes + 1]
 *   tets[4 n_tets] vertex indices, volume by volume; tet_ids[n_tets] file ids */
int xdg_h5m_get_mesh(const xdg_h5m_model* m, double* verts, int64_t* vert_ids, int32_t* surf_ids,
                     int64_t* surf_tri_off, int32_t* tris, int64_t* tri_ids, int32_t* surf_sense, int32_t* vol_ids,
                     int64_t* vol_tet_off, int32_t* tets, int64_t* tet_ids);
/* Groups: names[XDG_H5M_NAME_SIZE * n_groups] (NUL-padded),
 * member_off[n_groups + 1], member_dim[n_group_members] (2 surface,
 * 3 volume), member_id[n_group_members] (GLOBAL_IDs). */
int xdg_h5m_get_groups(const xdg_h5m_model* m, char* names, int64_t* member_off, int32_t* member_dim,
                       int32_t* member_id);

/* Writes a model (same arrays as above; surfaces' triangles and volumes'
 * tetrahedra index verts).  Groups may be absent (n_groups = 0). */
int xdg_h5m_write(const char* path, int64_t n_verts, const double* verts, int32_t n_surfaces,
                  const int32_t* surf_ids, const int64_t* surf_tri_off, const int32_t* tris,
                  const int32_t* surf_sense, int32_t n_volumes, const int32_t* vol_ids, const int64_t* vol_tet_off,
                  const int32_t* tets, int32_t n_groups, const char* names, const int64_t* member_off,
                  const int32_t* member_dim, const int32_t* member_id);

const char* xdg_h5m_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* XDG_H5M_H */
