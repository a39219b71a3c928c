"""DAGMC/MOAB .h5m geometry files for the backend (ctypes over
libxdg_h5m.so, include/xdg_h5m.h).

read_model(path) returns the file's surfaces, volumes, senses, tetrahedra
and groups as a SurfaceModel plus per-volume tetrahedra, completed with the
implicit complement the way MeshManager::create_implicit_complement does it
(src/mesh_manager_interface.cpp:20-45): a new volume id (one past the
largest), added as the missing parent of every surface that has only one.
write_model(path, model, ...) writes the inverse (the implicit complement is
not stored: DAGMC files do not hold it).

Group names follow the reference's metadata parsing (parse_metadata,
src/moab/mesh_manager.cpp:489-590): "key:value" tokens with keys mat,
material, boundary, temp; boundary_conditions() maps "boundary:vacuum" /
"boundary:reflecting" groups onto surface ids for the transport kernel.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import geometry as G

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libxdg_h5m.so")
NAME_SIZE = 32

_vp = C.c_void_p
_LIB = None


class H5mError(RuntimeError):
    pass


class Counts(C.Structure):
    _fields_ = [("n_verts", C.c_int64), ("n_tris", C.c_int64), ("n_tets", C.c_int64),
                ("n_surfaces", C.c_int32), ("n_volumes", C.c_int32), ("n_groups", C.c_int32),
                ("reserved", C.c_int32), ("n_group_members", C.c_int64)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(needs the HDF5 C library)")
        L = C.CDLL(LIB_PATH)
        L.xdg_h5m_read.argtypes = [C.c_char_p, C.POINTER(_vp)]
        L.xdg_h5m_free.argtypes = [_vp]
        L.xdg_h5m_free.restype = None
        L.xdg_h5m_get_counts.argtypes = [_vp, C.POINTER(Counts)]
        L.xdg_h5m_get_mesh.argtypes = [_vp] + [_vp] * 11
        L.xdg_h5m_get_groups.argtypes = [_vp] * 5
        L.xdg_h5m_write.argtypes = ([C.c_char_p, C.c_int64, _vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32,
                                     _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp, _vp])
        L.xdg_h5m_last_error.restype = C.c_char_p
        L.xdg_h5m_last_error.argtypes = []
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise H5mError(f"xdg_h5m error {rc}: {lib().xdg_h5m_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


class H5mModel:
    """A model read from a file: the surface model (implicit complement
    added), the tetrahedra of each volume, and the groups."""

    def __init__(self, model, vol_tets, tet_ids, tri_ids, groups):
        self.model = model            # G.SurfaceModel
        self.vol_tets = vol_tets      # {volume id: (k, 4) int32 vertex indices}
        self.tet_ids = tet_ids        # {volume id: (k,) file ids}
        self.tri_ids = tri_ids        # (nt,) file ids of the faces (row order of model.tris)
        self.groups = groups          # [(name, [(dim, id), ...])]

    def boundary_conditions(self):
        """{surface id: XDG_HIP_BC_*} from "boundary:..." groups."""
        codes = {"vacuum": 1, "reflecting": 2, "reflective": 2, "transmission": 0}
        bc = {}
        for name, members in self.groups:
            tok = name.split(":")
            for k in range(0, len(tok) - 1, 2):
                if tok[k] == "boundary" and tok[k + 1].lower() in codes:
                    for dim, mid in members:
                        if dim == 2:
                            bc[mid] = codes[tok[k + 1].lower()]
        return bc


def read_model(path) -> H5mModel:
    h = _vp()
    _check(lib().xdg_h5m_read(os.fsencode(path), C.byref(h)))
    try:
        c = Counts()
        _check(lib().xdg_h5m_get_counts(h, C.byref(c)))
        verts = np.zeros((c.n_verts, 3))
        vert_ids = np.zeros(c.n_verts, np.int64)
        surf_ids = np.zeros(c.n_surfaces, np.int32)
        surf_off = np.zeros(c.n_surfaces + 1, np.int64)
        tris = np.zeros((c.n_tris, 3), np.int32)
        tri_ids = np.zeros(c.n_tris, np.int64)
        sense = np.zeros((c.n_surfaces, 2), np.int32)
        vol_ids = np.zeros(c.n_volumes, np.int32)
        vol_off = np.zeros(c.n_volumes + 1, np.int64)
        tets = np.zeros((c.n_tets, 4), np.int32)
        tet_ids = np.zeros(c.n_tets, np.int64)
        _check(lib().xdg_h5m_get_mesh(h, _p(verts), _p(vert_ids), _p(surf_ids), _p(surf_off), _p(tris),
                                      _p(tri_ids), _p(sense), _p(vol_ids), _p(vol_off), _p(tets), _p(tet_ids)))
        names = np.zeros(NAME_SIZE * max(c.n_groups, 1), np.uint8)
        moff = np.zeros(c.n_groups + 1, np.int64)
        mdim = np.zeros(max(c.n_group_members, 1), np.int32)
        mid = np.zeros(max(c.n_group_members, 1), np.int32)
        _check(lib().xdg_h5m_get_groups(h, _p(names), _p(moff), _p(mdim), _p(mid)))
    finally:
        lib().xdg_h5m_free(h)
    tri_surf = np.repeat(surf_ids, np.diff(surf_off)).astype(np.int32)
    vols = [int(v) for v in vol_ids]
    ipc = (max(vols) + 1) if vols else 1
    senses = {}
    for s, (f, r) in zip(surf_ids.tolist(), sense.tolist()):
        senses[int(s)] = (f if f >= 0 else ipc, r if r >= 0 else ipc)
    model = G.SurfaceModel(verts, tris, tri_surf, [int(s) for s in surf_ids], senses, vols + [ipc], ipc=ipc)
    vol_tets = {v: tets[vol_off[i]:vol_off[i + 1]].copy() for i, v in enumerate(vols)}
    vtid = {v: tet_ids[vol_off[i]:vol_off[i + 1]].copy() for i, v in enumerate(vols)}
    groups = []
    for g in range(c.n_groups):
        raw = bytes(names[NAME_SIZE * g:NAME_SIZE * (g + 1)])
        name = raw.split(b"\0", 1)[0].decode()
        members = [(int(mdim[k]), int(mid[k])) for k in range(moff[g], moff[g + 1])]
        groups.append((name, members))
    return H5mModel(model, vol_tets, vtid, tri_ids, groups)


def write_model(path, model, vol_tets=None, groups=None):
    """Writes model's surfaces and volumes (not its implicit complement),
    vol_tets {volume id: (k, 4) vertex indices into model.verts}, groups
    [(name, [(dim, id), ...])]."""
    vols = [v for v in model.volumes if v != model.ipc]
    surfs = list(model.surfaces)
    verts = np.ascontiguousarray(model.verts, np.float64)
    tri_rows = [model.surface_faces(s) for s in surfs]
    tris = np.ascontiguousarray(np.concatenate([model.tris[r] for r in tri_rows]) if surfs
                                else np.zeros((0, 3)), np.int32)
    surf_off = np.concatenate([[0], np.cumsum([len(r) for r in tri_rows])]).astype(np.int64)
    sense = np.array([[f if f != model.ipc else -1, r if r != model.ipc else -1]
                      for f, r in (model.senses[s] for s in surfs)], np.int32).reshape(-1, 2)
    vol_tets = vol_tets or {}
    tl = [np.asarray(vol_tets.get(v, np.zeros((0, 4))), np.int32).reshape(-1, 4) for v in vols]
    tets = np.ascontiguousarray(np.concatenate(tl) if tl else np.zeros((0, 4)), np.int32)
    vol_off = np.concatenate([[0], np.cumsum([len(t) for t in tl])]).astype(np.int64)
    groups = groups or []
    names = np.zeros(NAME_SIZE * max(len(groups), 1), np.uint8)
    moff, mdim, mid = [0], [], []
    for g, (name, members) in enumerate(groups):
        b = name.encode()[:NAME_SIZE - 1]
        names[NAME_SIZE * g:NAME_SIZE * g + len(b)] = np.frombuffer(b, np.uint8)
        for dim, m in members:
            mdim.append(dim)
            mid.append(m)
        moff.append(len(mid))
    moff = np.array(moff, np.int64)
    mdim = np.array(mdim or [0], np.int32)
    mid = np.array(mid or [0], np.int32)
    sid = np.array(surfs, np.int32)
    vid = np.array(vols, np.int32)
    _check(lib().xdg_h5m_write(os.fsencode(path), verts.shape[0], _p(verts), len(surfs), _p(sid), _p(surf_off),
                               _p(tris), _p(sense), len(vols), _p(vid), _p(vol_off), _p(tets), len(groups),
                               _p(names), _p(moff), _p(mdim), _p(mid)))
