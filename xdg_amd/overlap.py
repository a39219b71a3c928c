"""Overlap check of a surface model on the GPU (tools/overlap.cpp).

check_instance_for_overlaps (overlap.cpp:63-205) looks for places that lie
inside more than one volume:

* vertex check: every vertex of every surface triangle (surfaces in model
  order, a vertex once per triangle that has it), moved 1e-9 along
  d = normalize(0.1, 0.1, 0.1) and then 2e-9 back along -d; at each of the
  two locations point_in_volume of every volume of the model, the implicit
  complement included, with the location's direction
  (check_location_for_overlap, overlap.cpp:9-61);
* edge check (optional): one ray along each triangle edge, length = the
  edge length (return_ray_queries, overlap.cpp:207-227), fired EXITING
  against every volume except the two parents of the edge's surface; the
  first volume hit gives an overlap between the hit surface's forward
  parent and the edge surface's forward parent (check_along_edge,
  overlap.cpp:229-265).

The result is the reference's OverlapMap: {frozenset(volume ids): location},
built in the serial order of the reference loops (one thread: each later
location with the same volume set replaces the earlier one).  The point
queries and edge rays run as two batched kernels
(xdg_hip_point_in_volumes, xdg_hip_edge_rays); only the map is built on the
host.
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as G

BUMP = 1e-9  # overlap.cpp:15


def overlap_direction() -> np.ndarray:
    """Direction dir = {0.1, 0.1, 0.1}; dir.normalize() (overlap.cpp:89-90)."""
    d = np.array([0.1, 0.1, 0.1])
    ln = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    ln = G.MIN_RCP_INPUT if ln < G.MIN_RCP_INPUT else ln
    return np.array([d[0] / ln, d[1] / ln, d[2] / ln])


def model_vertices(model) -> np.ndarray:
    """allVerts of check_instance_for_overlaps (overlap.cpp:76-87): per
    surface, per face, its three vertices."""
    faces = [model.surface_faces(s) for s in model.surfaces]
    faces = np.concatenate(faces) if faces else np.zeros(0, np.int32)
    return model.verts[model.tris[faces]].reshape(-1, 3)


def vertex_locations(model):
    """The two locations per vertex and their directions, in check order
    (vertex i -> rows 2i, 2i+1).  loc += dir * bump; dir *= -1;
    loc += dir * 2.0 * bump (overlap.cpp:15-18, 37-39)."""
    v = model_vertices(model)
    d = overlap_direction()
    n = v.shape[0]
    loc1 = v + d * BUMP
    nd = d * -1.0
    loc2 = loc1 + (nd * 2.0) * BUMP
    locs = np.empty((2 * n, 3))
    dirs = np.empty((2 * n, 3))
    locs[0::2], locs[1::2] = loc1, loc2
    dirs[0::2], dirs[1::2] = d, nd
    return locs, dirs


def edge_queries(model):
    """return_ray_queries (overlap.cpp:207-227) for every face of every
    surface, edges in (face, k) order: origin v_k, direction
    (v_{k+1} - v_k) / |v_{k+1} - v_k|, length |v_{k+1} - v_k|.
    Returns (org, dir, length, surface id per ray)."""
    orgs, ends, surfs = [], [], []
    for s in model.surfaces:
        f = model.surface_faces(s)
        tri = model.verts[model.tris[f]]  # (m, 3, 3)
        orgs.append(tri.reshape(-1, 3))
        ends.append(tri[:, [1, 2, 0]].reshape(-1, 3))
        surfs.append(np.full(3 * f.size, s, np.int32))
    if not surfs:
        z = np.zeros((0, 3))
        return z, z, np.zeros(0), np.zeros(0, np.int32)
    a, b = np.concatenate(orgs), np.concatenate(ends)
    d = b - a
    ln = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    d = d / ln[:, None]  # dir /= edgeLength: component-wise division
    return a, d, ln, np.concatenate(surfs)


class GpuQueries:
    """The two batched queries on a context: trees = {volume id: surface
    tree id} for every volume of the model (the implicit complement
    included), in model order."""

    def __init__(self, ctx, model, trees, precision=1, max_k=4):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.vols = [v for v in model.volumes if v in trees]
        self.trees = [trees[v] for v in self.vols]
        self.precision = precision
        self.max_k = max_k

    def _dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda")

    def point_in_volumes(self, pts, dirs):
        """(count, sets): sets[i] = volume ids containing location i."""
        cnt, ids = self.ctx.point_in_volumes(self.trees, self.vols, self._dev(pts),
                                             self._dev(dirs), self.max_k,
                                             precision=self.precision)
        cnt = cnt.cpu().numpy()
        ids = ids.cpu().numpy()
        many = np.nonzero(cnt > self.max_k)[0]
        full = {}
        if many.size:
            c2, i2 = self.ctx.point_in_volumes(self.trees, self.vols,
                                               self._dev(pts[many]),
                                               self._dev(dirs[many]),
                                               len(self.vols),
                                               precision=self.precision)
            i2 = i2.cpu().numpy()
            for j, r in enumerate(many):
                full[int(r)] = i2[j, :cnt[r]]
        return cnt, ids, full

    def edge_rays(self, org, dirs, length, skip):
        t, s, v = self.ctx.edge_rays(self.trees, self.vols, self._dev(org),
                                     self._dev(dirs), self._dev(length),
                                     self._dev(skip), precision=self.precision)
        return t.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy()


def check_instance_for_overlaps(queries, model, check_edges=False):
    """OverlapMap of the model: {frozenset(volume ids): (x, y, z)}.
    queries: GpuQueries (or any object with the same two methods)."""
    overlap = {}
    locs, dirs = vertex_locations(model)
    if locs.shape[0]:
        cnt, ids, full = queries.point_in_volumes(locs, dirs)
        for i in np.nonzero(cnt > 1)[0]:
            s = full.get(int(i))
            if s is None:
                s = ids[i, :cnt[i]]
            overlap[frozenset(int(x) for x in s)] = tuple(locs[i])
    if not check_edges:
        return overlap
    org, d, ln, esurf = edge_queries(model)
    if org.shape[0] == 0:
        return overlap
    skip = np.array([model.senses[int(s)] for s in esurf], np.int32).reshape(-1, 2)
    t, hs, _ = queries.edge_rays(org, d, ln, skip)
    seen = set()
    last = None
    for i in np.nonzero(hs >= 0)[0]:
        p = (org[i, 0] + t[i] * d[i, 0], org[i, 1] + t[i] * d[i, 1],
             org[i, 2] + t[i] * d[i, 2])
        if p not in seen:
            seen.add(p)
            last = p
        vol_hit = model.senses[int(hs[i])][0]
        overlap[frozenset((int(vol_hit), int(skip[i, 0])))] = last
    return overlap
