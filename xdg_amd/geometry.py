"""Synthetic DAGMC-style geometry for the geometry-query workloads.

A SurfaceModel mirrors what the reference's MeshManager exposes to a
RayTracer (include/xdg/mesh_manager_interface.h): vertices, triangle faces
grouped into surfaces, surface->(forward, reverse) parent volumes, and the
ordered volume list with an implicit complement.  A TetMesh mirrors the
element side (tet connectivity, face adjacency, boundary faces).

All coordinates are float64 and are produced once on the host; the device
path and the CPU oracle receive the same arrays bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

from . import philox

# include/xdg/moab/direct_access.h:181-183 (MOAB canonical tet face order)
MOAB_TET_FACES = np.array([[0, 1, 3], [1, 2, 3], [2, 0, 3], [0, 2, 1]],
                          dtype=np.int32)
# tests/mesh_mock.h:108-115 (MeshMock tet face order)
MOCK_TET_FACES = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 1], [1, 3, 2]],
                          dtype=np.int32)

# include/xdg/constants.h:80 -- float(FLT_MIN * 1e5), promoted to double
MIN_RCP_INPUT = float(np.float32(float(np.finfo(np.float32).tiny) * 1e5))
FP_BOX_TOL = 1e-6          # constants.h FP_BOX_TOL
NUMERICAL_PRECISION = 1e-3  # ray_tracing_interface.h:146


# ---------------------------------------------------------------------------
# surface models
# ---------------------------------------------------------------------------

def box_bump(lo, hi) -> float:
    """max(BoundingBox::dilation(), numerical_precision): the chord of the
    box times FP_BOX_TOL, at least 1e-3 (ray_tracing_interface.cpp:20-24)."""
    w = np.asarray(hi, dtype=np.float64) - np.asarray(lo, dtype=np.float64)
    chord = math.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
    return max(chord * FP_BOX_TOL, NUMERICAL_PRECISION)


def vertex_bump(verts) -> float:
    """Bump of a volume whose surfaces use exactly these vertices (the box
    starts all-zero, as MeshManager::volume_bounding_box does)."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    if v.size == 0:
        return box_bump(np.zeros(3), np.zeros(3))
    return box_bump(np.minimum(0.0, v.min(axis=0)), np.maximum(0.0, v.max(axis=0)))


@dataclass
class SurfaceModel:
    verts: np.ndarray                 # (nv, 3) float64
    tris: np.ndarray                  # (nt, 3) int32, face id == row index
    tri_surf: np.ndarray              # (nt,) int32 surface id per face
    surfaces: list                    # surface ids
    senses: dict                      # surf -> (forward_vol, reverse_vol)
    volumes: list                     # volume ids in registration order
    ipc: int = -1                     # implicit complement volume id
    _surf_faces: dict = field(default_factory=dict, repr=False)

    def surface_faces(self, s: int) -> np.ndarray:
        if s not in self._surf_faces:
            self._surf_faces[s] = np.nonzero(self.tri_surf == s)[0].astype(np.int32)
        return self._surf_faces[s]

    def volume_surfaces(self, vol: int) -> list:
        return [s for s in self.surfaces if vol in self.senses[s]]

    def tree_arrays(self, vol: int):
        """Triangles of one volume in registration order
        (EmbreeRayTracer::create_surface_tree, ray_tracer.cpp:50-105):
        verts9 (n, 9), face ids, surface ids, flip (reverse sense)."""
        faces, surfs, flips = [], [], []
        for s in self.volume_surfaces(vol):
            f = self.surface_faces(s)
            fwd, rev = self.senses[s]
            faces.append(f)
            surfs.append(np.full(f.size, s, dtype=np.int32))
            flips.append(np.full(f.size, 1 if vol == rev else 0, dtype=np.uint8))
        faces = np.concatenate(faces) if faces else np.zeros(0, np.int32)
        surfs = np.concatenate(surfs) if surfs else np.zeros(0, np.int32)
        flips = np.concatenate(flips) if flips else np.zeros(0, np.uint8)
        v9 = np.ascontiguousarray(
            self.verts[self.tris[faces]].reshape(-1, 9), dtype=np.float64)
        return v9, faces.astype(np.int32), surfs, flips

    def global_tree_arrays(self):
        """Triangles of every surface once, never flipped
        (EmbreeRayTracer::create_global_surface_tree, ray_tracer.cpp:190-206;
        the global tree is no volume's reverse tree), with the largest
        volume bump for the boxes."""
        faces = [self.surface_faces(s) for s in self.surfaces]
        faces = np.concatenate(faces) if faces else np.zeros(0, np.int32)
        v9 = np.ascontiguousarray(self.verts[self.tris[faces]].reshape(-1, 9), dtype=np.float64)
        bump = max([self.bump(v) for v in self.volumes if v != self.ipc] or [1e-3])
        return (v9, faces.astype(np.int32), self.tri_surf[faces].astype(np.int32),
                np.zeros(faces.size, np.uint8), bump)

    def volume_bbox(self, vol: int):
        """MeshManager::volume_bounding_box (mesh_manager_interface.cpp:237-246):
        starts from an all-zero box, so it always contains the origin."""
        lo = np.zeros(3)
        hi = np.zeros(3)
        for s in self.volume_surfaces(vol):
            f = self.surface_faces(s)
            if f.size == 0:
                continue
            p = self.verts[self.tris[f]].reshape(-1, 3)
            lo = np.minimum(lo, p.min(axis=0))
            hi = np.maximum(hi, p.max(axis=0))
        return lo, hi

    def bump(self, vol: int) -> float:
        """RayTracer::bounding_box_bump (ray_tracing_interface.cpp:20-24)."""
        lo, hi = self.volume_bbox(vol)
        return box_bump(lo, hi)

    def next_volume(self, vol: int, surf: int) -> int:
        """MeshManager::next_volume (mesh_manager_interface.cpp:202-214)."""
        f, r = self.senses[surf]
        if f == vol:
            return r
        if r == vol:
            return f
        raise ValueError(f"volume {vol} is not a parent of surface {surf}")


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
                 dtype=np.int64)
    return v, f


def _unit(v):
    n = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    return v / n[:, None]


def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Icosphere with 20*4^k faces and 10*4^k+2 vertices, CCW seen from
    outside (outward normals)."""
    v, f = _icosahedron()
    v = _unit(v)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (v.shape[0] + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (v.shape[0] + 1)
        b = uniq % (v.shape[0] + 1)
        mid = _unit(0.5 * (v[a] + v[b]))
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1)])
    # enforce outward orientation
    p = v[f]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    c = p.mean(axis=1)
    bad = (n * c).sum(axis=1) < 0
    f[bad] = f[bad][:, [0, 2, 1]]
    c = np.asarray(center, dtype=np.float64)
    return (v * radius + c).astype(np.float64), f.astype(np.int32)


def box_surface(lo, hi):
    """12 outward-facing triangles of an axis-aligned box."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    v = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]],
                  [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
    f = np.array([[0, 2, 1], [0, 3, 2],      # z = lo
                  [4, 5, 6], [4, 6, 7],      # z = hi
                  [0, 1, 5], [0, 5, 4],      # y = lo
                  [3, 7, 6], [3, 6, 2],      # y = hi
                  [0, 4, 7], [0, 7, 3],      # x = lo
                  [1, 2, 6], [1, 6, 5]],     # x = hi
                 dtype=np.int32)
    return v, f


def _assemble(parts):
    """parts: list of (verts, tris, surf_id) -> stacked arrays."""
    vs, ts, ss = [], [], []
    off = 0
    for v, t, s in parts:
        vs.append(v)
        ts.append(t + off)
        ss.append(np.full(t.shape[0], s, dtype=np.int32))
        off += v.shape[0]
    return (np.concatenate(vs).astype(np.float64),
            np.concatenate(ts).astype(np.int32), np.concatenate(ss))


def _two_volume_model(parts, senses) -> SurfaceModel:
    verts, tris, ts = _assemble(parts)
    surfs = sorted(senses)
    return SurfaceModel(verts, tris, ts, surfs, dict(senses), [1, 2, 3], ipc=3)


def overlapping_boxes() -> SurfaceModel:
    """Two boxes that overlap (the analogue of overlap.h5m in the
    reference's tests/test_overlap_check.cpp): A = [0,2] x [0,1.9] x [0,2.2]
    and B = [1,3] x [1.3,2.75] x [0.7,3.1]; A's corner (2,1.9,2.2) lies inside
    B and B's corner (1,1.3,0.7) inside A.  Volumes 1, 2, implicit
    complement 3.  Side lengths are chosen so that no ray along +-(1,1,1)
    from a vertex meets a box edge or corner."""
    av, at = box_surface([0.0] * 3, [2.0, 1.9, 2.2])
    bv, bt = box_surface([1.0, 1.3, 0.7], [3.0, 2.75, 3.1])
    return _two_volume_model([(av, at, 1), (bv, bt, 2)], {1: (1, 3), 2: (2, 3)})


def imprinted_boxes() -> SurfaceModel:
    """Two boxes [0,1.1] x [0,0.7] x [0,0.45] and [1.1,2.3] x [0,0.7] x
    [0,0.45] sharing the face x = 1.1 as one surface with senses (1, 2)
    (no_overlap_imp.h5m): no overlap."""
    av, at = box_surface([0.0] * 3, [1.1, 0.7, 0.45])
    bv, bt = box_surface([1.1, 0.0, 0.0], [2.3, 0.7, 0.45])
    # box_surface rows 8,9: x = lo; rows 10,11: x = hi
    a_out = at[np.r_[0:10]]
    shared = at[10:12]
    b_out = np.concatenate([bt[0:8], bt[10:12]])
    return _two_volume_model([(av, a_out, 1), (av, shared, 2), (bv, b_out, 3)],
                             {1: (1, 3), 2: (1, 2), 3: (2, 3)})


def enclosed_boxes() -> SurfaceModel:
    """A box [1.05,2] x [1.2,2.13] x [0.9,1.61] (volume 2) inside a box
    [0,4] x [0,3.7] x [0,4.4] (volume 1) whose surface does not bound it
    (enclosed.h5m): volume 2 overlaps volume 1."""
    av, at = box_surface([0.0] * 3, [4.0, 3.7, 4.4])
    bv, bt = box_surface([1.05, 1.2, 0.9], [2.0, 2.13, 1.61])
    return _two_volume_model([(av, at, 1), (bv, bt, 2)], {1: (1, 3), 2: (2, 3)})


def crossed_bars() -> SurfaceModel:
    """Two bars crossing like a plus sign (overlap-edge.h5m): A = [-2,2] x
    [-0.5,0.5] x [-0.1,0.1], B = [-0.5,0.5] x [-2,2] x [-0.2,0.2].  No vertex
    of one lies inside the other, but A's long edges pass through B."""
    av, at = box_surface([-2.0, -0.5, -0.1], [2.0, 0.5, 0.1])
    bv, bt = box_surface([-0.5, -2.0, -0.2], [0.5, 2.0, 0.2])
    return _two_volume_model([(av, at, 1), (bv, bt, 2)], {1: (1, 3), 2: (2, 3)})


def single_sphere(subdiv: int = 5, radius: float = 10.0) -> SurfaceModel:
    """One closed icosphere volume (id 1) plus its implicit complement (2)."""
    v, t = icosphere(subdiv, radius)
    verts, tris, ts = _assemble([(v, t, 1)])
    return SurfaceModel(verts, tris, ts, [1], {1: (1, 2)}, [1, 2], ipc=2)


def nested_spheres(subdiv: int = 6, radii=(5.0, 10.0, 15.0),
                   half: float = 20.0) -> SurfaceModel:
    """Spheres r0 < r1 < r2 inside the cube [-half, half]^3.  Volumes:
    1 inside r0, 2 between r0/r1, 3 between r1/r2, 4 between r2 and the
    cube, 5 the implicit complement outside the cube."""
    parts = [icosphere(subdiv, r) + (i + 1,) for i, r in enumerate(radii)]
    bv, bt = box_surface([-half] * 3, [half] * 3)
    parts.append((bv, bt, len(radii) + 1))
    verts, tris, ts = _assemble(parts)
    nsurf = len(radii) + 1
    senses = {s: (s, s + 1) for s in range(1, nsurf + 1)}
    vols = list(range(1, nsurf + 2))
    return SurfaceModel(verts, tris, ts, list(range(1, nsurf + 1)), senses,
                        vols, ipc=nsurf + 1)


def sphere_lattice(n: int = 8, subdiv: int = 4, radius: float = 0.8,
                   pitch: float = 2.0) -> SurfaceModel:
    """n^3 icospheres (centres pitch/2 + pitch*i) in the box [0, n*pitch]^3.
    Volumes 1..n^3 are the spheres, n^3+1 the box region outside them,
    n^3+2 the implicit complement outside the box."""
    sv, st = icosphere(subdiv, 1.0)
    parts = []
    sid = 1
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = np.array([pitch / 2 + pitch * i, pitch / 2 + pitch * j,
                              pitch / 2 + pitch * k])
                parts.append((sv * radius + c, st, sid))
                sid += 1
    L = n * pitch
    bv, bt = box_surface([0.0] * 3, [L] * 3)
    box_sid = sid
    parts.append((bv, bt, box_sid))
    verts, tris, ts = _assemble(parts)
    box_vol = n ** 3 + 1
    ipc = n ** 3 + 2
    senses = {s: (s, box_vol) for s in range(1, n ** 3 + 1)}
    senses[box_sid] = (box_vol, ipc)
    return SurfaceModel(verts, tris, ts, list(range(1, box_sid + 1)), senses,
                        list(range(1, ipc + 1)), ipc=ipc)


def perturbed_sphere(subdiv: int = 6, radius: float = 10.0,
                     amp: float = 0.05) -> SurfaceModel:
    """Icosphere with r = R (1 + amp sin(3 theta) cos(2 phi))."""
    v, t = icosphere(subdiv, 1.0)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = radius * (1.0 + amp * np.sin(3.0 * theta) * np.cos(2.0 * phi))
    v = v * r[:, None]
    verts, tris, ts = _assemble([(v, t, 1)])
    return SurfaceModel(verts, tris, ts, [1], {1: (1, 2)}, [1, 2], ipc=2)


# ---------------------------------------------------------------------------
# MeshMock (tests/mesh_mock.h:246-319): 1 volume, 6 surfaces x 2 triangles,
# 12 tets sharing the box centre.
# ---------------------------------------------------------------------------

MOCK_BOX = (-2.0, -3.0, -4.0, 5.0, 6.0, 7.0)


def mesh_mock_arrays():
    x0, y0, z0, x1, y1, z1 = MOCK_BOX
    verts = np.array([
        [x1, y0, z1], [x1, y1, z1], [x0, y1, z1], [x0, y0, z1],
        [x1, y0, z0], [x1, y1, z0], [x0, y1, z0], [x0, y0, z0],
        [(x0 + x1) * 0.5, (y0 + y1) * 0.5, (z0 + z1) * 0.5]],
        dtype=np.float64)
    tets = np.array([[0, 1, 2, 8], [0, 2, 3, 8], [4, 6, 5, 8], [4, 7, 6, 8],
                     [0, 5, 1, 8], [0, 4, 5, 8], [2, 6, 7, 8], [2, 7, 3, 8],
                     [0, 7, 4, 8], [0, 3, 7, 8], [1, 5, 6, 8], [1, 6, 2, 8]],
                    dtype=np.int32)
    tris = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6],
                     [0, 5, 1], [0, 4, 5], [2, 6, 7], [2, 7, 3],
                     [1, 5, 6], [1, 6, 2], [0, 7, 4], [0, 3, 7]],
                    dtype=np.int32)
    adj = np.array([[-1, 1, 4, 11], [-1, 9, 0, 7], [-1, 5, 3, 10],
                    [-1, 2, 8, 6], [-1, 0, 5, 10], [-1, 4, 8, 2],
                    [-1, 7, 11, 3], [-1, 1, 6, 9], [-1, 5, 9, 3],
                    [-1, 8, 1, 7], [-1, 11, 4, 2], [-1, 0, 10, 6]],
                   dtype=np.int32)
    return verts, tris, tets, adj


def mesh_mock_model(with_ipc: bool = False) -> SurfaceModel:
    """MeshMock as a surface model: volume 0, surfaces 0..5 (2 faces each).
    with_ipc adds the implicit complement (volume 1) as
    create_implicit_complement (mesh_manager_interface.cpp:19-45) does."""
    verts, tris, _, _ = mesh_mock_arrays()
    ts = np.repeat(np.arange(6, dtype=np.int32), 2)
    senses = {s: (0, 1 if with_ipc else -1) for s in range(6)}
    vols = [0, 1] if with_ipc else [0]
    return SurfaceModel(verts, tris, ts, list(range(6)), senses, vols,
                        ipc=1 if with_ipc else -1)


# ---------------------------------------------------------------------------
# tetrahedral meshes
# ---------------------------------------------------------------------------

@dataclass
class TetMesh:
    verts: np.ndarray        # (nv, 3) float64
    tets: np.ndarray         # (ne, 4) int32
    adj: np.ndarray          # (ne, 4) int32, neighbour across face i or -1
    face_order: np.ndarray   # (4, 3) int32 local face triples
    bfaces: np.ndarray       # (nb, 3) int32 boundary triangles (outward)
    bface_elem: np.ndarray   # (nb,) int32 element owning each boundary face

    def boundary_model(self) -> SurfaceModel:
        """Boundary faces as one surface (id 1) of the mesh volume (1) with
        the implicit complement (2) on the other side."""
        ts = np.ones(self.bfaces.shape[0], dtype=np.int32)
        return SurfaceModel(self.verts, self.bfaces, ts, [1], {1: (1, 2)},
                            [1, 2], ipc=2)


def tet_adjacency(tets: np.ndarray, face_order: np.ndarray, nv: int):
    """Face adjacency by sorting face vertex triples; returns (adj,
    boundary faces (outward, in face_order winding), owning elements)."""
    ne = tets.shape[0]
    fv = tets[:, face_order]                       # (ne, 4, 3)
    s = np.sort(fv.reshape(-1, 3), axis=1).astype(np.int64)
    bits = max(1, int(nv).bit_length())
    key = (s[:, 0] << (2 * bits)) | (s[:, 1] << bits) | s[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    adj = np.full(ne * 4, -1, dtype=np.int32)
    same = ks[1:] == ks[:-1]
    a = order[:-1][same]
    b = order[1:][same]
    adj[a] = (b // 4).astype(np.int32)
    adj[b] = (a // 4).astype(np.int32)
    bmask = adj < 0
    bidx = np.nonzero(bmask)[0]
    bfaces = fv.reshape(-1, 3)[bidx].astype(np.int32)
    belem = (bidx // 4).astype(np.int32)
    return adj.reshape(ne, 4), bfaces, belem


def two_blocks(n: int = 8, gap: float = 0.5, jitter: float = 0.2) -> TetMesh:
    """Two Kuhn unit cubes, the second shifted by 1 + gap along x, as one
    mesh: a particle that leaves the first block can re-enter the mesh in
    the second through the implicit complement."""
    a = kuhn_cube(n, jitter, seed=4)
    b = kuhn_cube(n, jitter, seed=5)
    nv, ne = a.verts.shape[0], a.tets.shape[0]
    bv = b.verts.copy()
    bv[:, 0] += 1.0 + gap
    badj = np.where(b.adj >= 0, b.adj + ne, -1).astype(np.int32)
    return TetMesh(np.concatenate([a.verts, bv]),
                   np.concatenate([a.tets, b.tets + nv]).astype(np.int32),
                   np.concatenate([a.adj, badj]).astype(np.int32), a.face_order,
                   np.concatenate([a.bfaces, b.bfaces + nv]).astype(np.int32),
                   np.concatenate([a.bface_elem, b.bface_elem + ne]).astype(np.int32))


def tet_signed_volumes(verts, tets):
    p = verts[tets]
    e1 = p[:, 1] - p[:, 0]
    e2 = p[:, 2] - p[:, 0]
    e3 = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", np.cross(e1, e2), e3) / 6.0


def kuhn_cube(n: int = 64, jitter: float = 0.2, seed: int = 4) -> TetMesh:
    """Unit cube split into n^3 hexes x 6 Kuhn tets, positively oriented so
    that the MOAB canonical faces point outward; interior vertices jittered
    uniformly by +-jitter*h per coordinate (Philox, stream AUX)."""
    h = 1.0 / n
    m = n + 1
    idx = np.arange(m ** 3)
    ix = idx % m
    iy = (idx // m) % m
    iz = idx // (m * m)
    verts = np.stack([ix / n, iy / n, iz / n], axis=1).astype(np.float64)
    if jitter > 0:
        interior = ((ix > 0) & (ix < n) & (iy > 0) & (iy < n) &
                    (iz > 0) & (iz < n))
        ii = np.nonzero(interior)[0].astype(np.uint64)
        ua, ub = philox.uniform2(seed, philox.STREAM_AUX, ii, 0)
        uc, _ = philox.uniform2(seed, philox.STREAM_AUX, ii, 1)
        amp = jitter * h
        verts[ii.astype(np.int64), 0] += (ua * 2.0 - 1.0) * amp
        verts[ii.astype(np.int64), 1] += (ub * 2.0 - 1.0) * amp
        verts[ii.astype(np.int64), 2] += (uc * 2.0 - 1.0) * amp

    hx, hy, hz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing="ij")
    hx = hx.transpose(2, 1, 0).ravel()
    hy = hy.transpose(2, 1, 0).ravel()
    hz = hz.transpose(2, 1, 0).ravel()

    def vid(dx, dy, dz):
        return (hx + dx) + m * ((hy + dy) + m * (hz + dz))

    unit = {0: (1, 0, 0), 1: (0, 1, 0), 2: (0, 0, 1)}
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        a = unit[p[0]]
        b = tuple(unit[p[0]][k] + unit[p[1]][k] for k in range(3))
        v0 = vid(0, 0, 0)
        v1 = vid(*a)
        v2 = vid(*b)
        v3 = vid(1, 1, 1)
        # det[e_a, e_b, e_c] = parity of the permutation
        inv = sum(1 for i in range(3) for j in range(i + 1, 3) if p[i] > p[j])
        if inv % 2:
            v1, v2 = v2, v1
        tets.append(np.stack([v0, v1, v2, v3], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    vol = tet_signed_volumes(verts, tets)
    if not np.all(vol > 0):
        raise RuntimeError("kuhn_cube: non-positive tetrahedron volume")
    adj, bfaces, belem = tet_adjacency(tets, MOAB_TET_FACES, verts.shape[0])
    return TetMesh(verts, tets, adj, MOAB_TET_FACES.copy(), bfaces, belem)


def mock_tetmesh() -> TetMesh:
    verts, tris, tets, adj = mesh_mock_arrays()
    _, bfaces, belem = tet_adjacency(tets, MOCK_TET_FACES, verts.shape[0])
    return TetMesh(verts, tets, adj, MOCK_TET_FACES.copy(), bfaces, belem)
