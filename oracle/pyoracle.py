"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)
_u64 = C.POINTER(C.c_uint64)
_vp = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # XDG_ORACLE_LIB: another build of the same source (the sanitizer
        # build of tests/test_oracle_sanitized.py)
        path = os.environ.get("XDG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.xo_plucker_ray_tri_intersect.restype = C.c_int
        L.xo_plucker_ray_tri_intersect.argtypes = [_d, _d, _d, C.c_double,
                                                   C.c_double, C.c_int, C.c_int,
                                                   _d]
        L.xo_plucker_edge_test.restype = C.c_double
        L.xo_plucker_edge_test.argtypes = [_d, _d, _d, _d]
        L.xo_face_normal.argtypes = [_d, _d]
        L.xo_closest_location_on_triangle.argtypes = [_d, _d, _d]
        L.xo_determine_region.restype = C.c_int
        L.xo_determine_region.argtypes = [_d, _d]
        L.xo_tet_contains.restype = C.c_int
        L.xo_tet_contains.argtypes = [_d, _d, _d]
        L.xo_tree_create.restype = _vp
        L.xo_tree_create.argtypes = [C.c_int64, _d, _i32, _i32, _u8, C.c_double,
                                     C.c_int]
        L.xo_tree_destroy.argtypes = [_vp]
        L.xo_ray_fire.argtypes = [_vp, C.c_int64, _d, _d, _d, C.c_int, _i32,
                                  C.c_int, C.c_int, C.c_int, _d, _i32, _i32]
        L.xo_point_in_volume.argtypes = [_vp, C.c_int64, _d, _d, _i32, C.c_int,
                                         C.c_int, C.c_int, _u8]
        L.xo_occluded.argtypes = [_vp, C.c_int64, _d, _d, C.c_int, C.c_int,
                                  _u8, _d]
        L.xo_closest.argtypes = [_vp, C.c_int64, _d, C.c_int, C.c_int, _d, _i32]
        L.xo_find_volume.argtypes = [C.POINTER(_vp), _i32, C.c_int, C.c_int32,
                                     C.c_int64, _d, _d, C.c_int, C.c_int, _i32]
        L.xo_tetmesh_create.restype = _vp
        L.xo_tetmesh_create.argtypes = [C.c_int64, _d, C.c_int64, _i32, _i32,
                                        _i32, C.c_int]
        L.xo_tetmesh_destroy.argtypes = [_vp]
        L.xo_find_element.argtypes = [_vp, C.c_int64, _d, C.c_int, C.c_int,
                                      _i32]
        L.xo_next_element.argtypes = [_vp, C.c_int32, _d, _d, _i32, _d]
        L.xo_set_fault.restype = None
        L.xo_set_fault.argtypes = [C.c_int]
        L.xo_walk_elements.restype = C.c_int64
        L.xo_walk_elements.argtypes = [_vp, C.c_int32, _d, _d, C.c_double,
                                       C.c_int64, _i32, _d]
        L.xo_segments.argtypes = [_vp, _vp, _i32, C.c_int64, _d, _d, C.c_int,
                                  _i32, _u64, _i64, _i64, _i32, _d]
        L.xo_segments_volume.argtypes = [_vp, _vp, C.c_int64, _d, _d, C.c_int,
                                         _i32, _u64, _i64, _i64, _i32, _d,
                                         C.POINTER(C.c_uint32)]
        L.xo_mesh_walk.argtypes = [_vp, _vp, _d, _d, C.c_uint32, C.c_uint32,
                                   C.c_double, C.c_int32, C.c_uint64,
                                   C.c_int64, C.c_int, _d, _i32, _u8, _u64,
                                   _i64]
        L.xo_mix64.restype = C.c_uint64
        L.xo_mix64.argtypes = [C.c_uint64]
        L.xo_philox4x32_10.argtypes = [C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]
        L.xo_uniform2.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64,
                                  C.c_uint32, _d, _d]
        L.xo_log.restype = C.c_double
        L.xo_log.argtypes = [C.c_double]
        L.xo_adjoint.argtypes = [C.c_int64, _d, C.c_int64, _i32, C.c_int64,
                                 _d, _d, _d, _i32, _d, _d, _d, _d]
        L.xo_transport.argtypes = [C.POINTER(TransportProblem), C.c_uint64,
                                   C.c_int64, C.c_int, C.POINTER(C.c_int64),
                                   _i32, _u8, C.POINTER(C.c_uint64)]
        _LIB = L
    return _LIB


class TransportProblem(C.Structure):
    _fields_ = [("nvol", C.c_int), ("trees", C.POINTER(C.c_void_p)),
                ("vol_ids", C.POINTER(C.c_int32)), ("ipc", C.c_int32),
                ("nsurf_ids", C.c_int), ("sense_fwd", C.POINTER(C.c_int32)),
                ("sense_rev", C.POINTER(C.c_int32)), ("seed", C.c_uint32),
                ("stream", C.c_uint32), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("mfp", C.c_double),
                ("p_absorb", C.c_double), ("max_events", C.c_int32),
                ("surf_bc", C.POINTER(C.c_uint8))]

BC_TRANSMISSION, BC_VACUUM, BC_REFLECTING = 0, 1, 2
TR_ABSORBED, TR_LEAKED, TR_LOST, TR_STUCK, TR_CAPPED, TR_ERROR, TR_VACUUM = range(1, 8)


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32a(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- primitives -----------------------------------------------------------

def plucker(v9, o, d, tmax=np.finfo(np.float64).max, tmin=0.0,
            use_orientation=False, orientation=0):
    v9, o, d = _f64(v9), _f64(o), _f64(d)
    t = C.c_double(0.0)
    hit = lib().xo_plucker_ray_tri_intersect(_p(v9, _d), _p(o, _d), _p(d, _d),
                                             float(tmax), float(tmin),
                                             int(use_orientation),
                                             int(orientation), C.byref(t))
    return bool(hit), t.value


def edge_test(va, vb, ray, rn):
    return lib().xo_plucker_edge_test(_p(_f64(va), _d), _p(_f64(vb), _d),
                                      _p(_f64(ray), _d), _p(_f64(rn), _d))


def face_normal(v9):
    n = np.zeros(3)
    lib().xo_face_normal(_p(_f64(v9), _d), _p(n, _d))
    return n


def closest_point(v9, p):
    out = np.zeros(3)
    lib().xo_closest_location_on_triangle(_p(_f64(v9), _d), _p(_f64(p), _d),
                                          _p(out, _d))
    return out


def region(v9, p):
    return lib().xo_determine_region(_p(_f64(v9), _d), _p(_f64(p), _d))


FAULT_NONE = 0
FAULT_FIRST_VISITED = 1
FAULT_NO_DILATION = 2
FAULT_TET_TRANSPOSE = 3


class planted_fault:
    """Context manager that plants a mistake in the oracle (tests that show
    a check fails on it).  Element trees built inside the block keep the
    boxes they were built with."""

    def __init__(self, code):
        self.code = int(code)

    def __enter__(self):
        lib().xo_set_fault(self.code)
        return self

    def __exit__(self, *exc):
        lib().xo_set_fault(FAULT_NONE)
        return False


def tet_contains(p, v12):
    lam = np.zeros(4)
    r = lib().xo_tet_contains(_p(_f64(p), _d), _p(_f64(v12), _d), _p(lam, _d))
    return bool(r), lam


def mix64(x):
    return lib().xo_mix64(C.c_uint64(x))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().xo_philox4x32_10(c, k, o)
    return list(o)


def uniform2(seed, stream, item, draw):
    a, b = C.c_double(), C.c_double()
    lib().xo_uniform2(seed, stream, item, draw, C.byref(a), C.byref(b))
    return a.value, b.value


def log(x):
    return lib().xo_log(float(x))


# ---- surface trees --------------------------------------------------------

class Tree:
    """A surface tree (one volume's triangles)."""

    def __init__(self, verts9, face_id, surf_id, flip, bump, bvh=True):
        self.v9 = _f64(verts9).reshape(-1, 9)
        self.face = _i32a(face_id)
        self.surf = _i32a(surf_id)
        self.flip = np.ascontiguousarray(flip, dtype=np.uint8)
        self.n = self.v9.shape[0]
        self.h = lib().xo_tree_create(self.n, _p(self.v9, _d),
                                      _p(self.face, _i32), _p(self.surf, _i32),
                                      _p(self.flip, _u8), float(bump), int(bvh))

    @classmethod
    def from_model(cls, model, vol, bvh=True):
        v9, f, s, fl = model.tree_arrays(vol)
        return cls(v9, f, s, fl, model.bump(vol), bvh)

    def __del__(self):
        if getattr(self, "h", None):
            lib().xo_tree_destroy(self.h)
            self.h = None

    def ray_fire(self, org, dirs, tmax=None, orientation=0, excl=None,
                 brute=False, nthreads=0):
        org, dirs = _f64(org).reshape(-1, 3), _f64(dirs).reshape(-1, 3)
        n = org.shape[0]
        t = np.zeros(n)
        s = np.zeros(n, np.int32)
        f = np.zeros(n, np.int32)
        k = 0
        if excl is not None:
            excl = _i32a(excl).reshape(n, -1)
            k = excl.shape[1]
        tm = None if tmax is None else _f64(np.broadcast_to(tmax, (n,)))
        lib().xo_ray_fire(self.h, n, _p(org, _d), _p(dirs, _d), _p(tm, _d),
                          orientation, _p(excl, _i32), k, int(brute), nthreads,
                          _p(t, _d), _p(s, _i32), _p(f, _i32))
        return t, s, f

    def point_in_volume(self, pts, dirs=None, excl=None, brute=False,
                        nthreads=0):
        pts = _f64(pts).reshape(-1, 3)
        n = pts.shape[0]
        dd = None if dirs is None else _f64(dirs).reshape(-1, 3)
        k = 0
        if excl is not None:
            excl = _i32a(excl).reshape(n, -1)
            k = excl.shape[1]
        out = np.zeros(n, np.uint8)
        lib().xo_point_in_volume(self.h, n, _p(pts, _d), _p(dd, _d),
                                 _p(excl, _i32), k, int(brute), nthreads,
                                 _p(out, _u8))
        return out.astype(bool)

    def occluded(self, org, dirs, brute=False, nthreads=0):
        org, dirs = _f64(org).reshape(-1, 3), _f64(dirs).reshape(-1, 3)
        n = org.shape[0]
        hit = np.zeros(n, np.uint8)
        dist = np.zeros(n)
        lib().xo_occluded(self.h, n, _p(org, _d), _p(dirs, _d), int(brute),
                          nthreads, _p(hit, _u8), _p(dist, _d))
        return hit.astype(bool), dist

    def closest(self, pts, brute=False, nthreads=0):
        pts = _f64(pts).reshape(-1, 3)
        n = pts.shape[0]
        dist = np.zeros(n)
        face = np.zeros(n, np.int32)
        lib().xo_closest(self.h, n, _p(pts, _d), int(brute), nthreads,
                         _p(dist, _d), _p(face, _i32))
        return dist, face


def find_volume(trees, vol_ids, ipc, pts, dirs, brute=False, nthreads=0):
    pts, dirs = _f64(pts).reshape(-1, 3), _f64(dirs).reshape(-1, 3)
    n = pts.shape[0]
    arr = (_vp * len(trees))(*[t.h for t in trees])
    ids = _i32a(vol_ids)
    out = np.zeros(n, np.int32)
    lib().xo_find_volume(arr, _p(ids, _i32), len(trees), ipc, n, _p(pts, _d),
                         _p(dirs, _d), int(brute), nthreads, _p(out, _i32))
    return out


# ---- tet meshes -----------------------------------------------------------

class TetMesh:
    def __init__(self, mesh, bvh=True, with_ipc=True):
        self.mesh = mesh
        self.verts = _f64(mesh.verts)
        self.tets = _i32a(mesh.tets)
        self.adj = _i32a(mesh.adj)
        self.fo = _i32a(mesh.face_order).reshape(12)
        self.h = lib().xo_tetmesh_create(self.verts.shape[0], _p(self.verts, _d),
                                         self.tets.shape[0], _p(self.tets, _i32),
                                         _p(self.adj, _i32), _p(self.fo, _i32),
                                         int(bvh))
        self.ipc = None
        self.bface_elem = _i32a(mesh.bface_elem)
        if with_ipc and mesh.bfaces.shape[0] > 0:
            bm = mesh.boundary_model()
            self.ipc = Tree.from_model(bm, bm.ipc, bvh=bvh)

    def __del__(self):
        if getattr(self, "h", None):
            lib().xo_tetmesh_destroy(self.h)
            self.h = None

    def find_element(self, pts, brute=False, nthreads=0):
        pts = _f64(pts).reshape(-1, 3)
        out = np.zeros(pts.shape[0], np.int32)
        lib().xo_find_element(self.h, pts.shape[0], _p(pts, _d), int(brute),
                              nthreads, _p(out, _i32))
        return out

    def next_element(self, elem, r, u):
        nx = C.c_int32()
        d = C.c_double()
        lib().xo_next_element(self.h, int(elem), _p(_f64(r), _d),
                              _p(_f64(u), _d), C.byref(nx), C.byref(d))
        return nx.value, d.value

    def walk_elements(self, elem, r, u, distance, cap=1 << 16):
        e = np.zeros(cap, np.int32)
        ln = np.zeros(cap)
        n = lib().xo_walk_elements(self.h, int(elem), _p(_f64(r), _d),
                                   _p(_f64(u), _d), float(distance), cap,
                                   _p(e, _i32), _p(ln, _d))
        return e[:min(n, cap)].copy(), ln[:min(n, cap)].copy()

    def segments(self, p0, p1, tally=False, sequences=None, nthreads=0):
        """Returns (count, hash[, tally_fx][, (offsets, elems, lens)]).
        sequences: None or per-segment capacity array for explicit output."""
        p0, p1 = _f64(p0).reshape(-1, 3), _f64(p1).reshape(-1, 3)
        n = p0.shape[0]
        cnt = np.zeros(n, np.int32)
        hsh = np.zeros(n, np.uint64)
        tal = np.zeros(self.tets.shape[0], np.int64) if tally else None
        off = se = sl = None
        if sequences is not None:
            cap = np.asarray(sequences, dtype=np.int64)
            off = np.zeros(n + 1, np.int64)
            off[1:] = np.cumsum(cap)
            se = np.zeros(int(off[-1]), np.int32)
            sl = np.zeros(int(off[-1]))
        lib().xo_segments(self.h, self.ipc.h if self.ipc else None,
                          _p(self.bface_elem, _i32), n, _p(p0, _d), _p(p1, _d),
                          nthreads, _p(cnt, _i32), _p(hsh, _u64),
                          _p(tal, _i64), _p(off, _i64), _p(se, _i32),
                          _p(sl, _d))
        out = [cnt, hsh]
        if tally:
            out.append(tal)
        if sequences is not None:
            out.append((off, se, sl))
        return tuple(out)

    def mesh_walk(self, first, n, seed, stream, lo, hi, mfp, max_events=100000,
                  tally=False, nthreads=0):
        """xo_mesh_walk (tools/walk_elements.h:53-117): per particle
        distance, events, status, hash[, per-element tally_fx]."""
        dist = np.zeros(n)
        ev = np.zeros(n, np.int32)
        st = np.zeros(n, np.uint8)
        hs = np.zeros(n, np.uint64)
        tal = np.zeros(self.tets.shape[0], np.int64) if tally else None
        lo_a, hi_a = _f64(lo), _f64(hi)
        lib().xo_mesh_walk(self.h, self.ipc.h if self.ipc else None,
                           _p(lo_a, _d), _p(hi_a, _d), seed, stream, float(mfp),
                           int(max_events), int(first), int(n), nthreads,
                           _p(dist, _d), _p(ev, _i32), _p(st, _u8),
                           _p(hs, _u64), _p(tal, _i64))
        out = [dist, ev, st, hs]
        if tally:
            out.append(tal)
        return tuple(out)

    def segments_volume(self, vol_tree, p0, p1, tally=False, sequences=None,
                        nthreads=0):
        """XDG::segments(volume, start, end).  vol_tree: the volume's
        surface Tree (None: starts outside the mesh give no entries).
        Returns (count, hash, flags[, tally_fx][, (offsets, elems, lens)])."""
        p0, p1 = _f64(p0).reshape(-1, 3), _f64(p1).reshape(-1, 3)
        n = p0.shape[0]
        cnt = np.zeros(n, np.int32)
        hsh = np.zeros(n, np.uint64)
        flg = np.zeros(n, np.uint32)
        tal = np.zeros(self.tets.shape[0], np.int64) if tally else None
        off = se = sl = None
        if sequences is not None:
            cap = np.asarray(sequences, dtype=np.int64)
            off = np.zeros(n + 1, np.int64)
            off[1:] = np.cumsum(cap)
            se = np.zeros(int(off[-1]), np.int32)
            sl = np.zeros(int(off[-1]))
        lib().xo_segments_volume(self.h, vol_tree.h if vol_tree else None, n,
                                 _p(p0, _d), _p(p1, _d), nthreads,
                                 _p(cnt, _i32), _p(hsh, _u64), _p(tal, _i64),
                                 _p(off, _i64), _p(se, _i32), _p(sl, _d),
                                 flg.ctypes.data_as(C.POINTER(C.c_uint32)))
        out = [cnt, hsh, flg]
        if tally:
            out.append(tal)
        if sequences is not None:
            out.append((off, se, sl))
        return tuple(out)


# ---- batched particle random walk -----------------------------------------

def sense_tables(model):
    """Forward / reverse parent volume per surface id (-1 where absent)."""
    ns = max(model.senses) + 1
    fwd = np.full(ns, -1, np.int32)
    rev = np.full(ns, -1, np.int32)
    for sid, (f, r) in model.senses.items():
        fwd[sid] = f
        rev[sid] = r
    return fwd, rev


def bc_table(model, bc):
    """Per-surface-id boundary condition codes from {surface id: code}
    (None: every surface transmits)."""
    if not bc:
        return None
    ns = max(max(model.senses), max(bc)) + 1
    t = np.zeros(ns, np.uint8)
    for sid, code in bc.items():
        t[sid] = code
    return t


def transport(model, trees, vol_ids, seed, stream, lo, hi, mfp, p_absorb,
              max_events, first, n, nthreads=0, bc=None):
    """xo_transport over histories first .. first+n-1.  trees: Tree per
    entry of vol_ids (find_volume order).  bc: {surface id: BC_*} (others
    transmit).  Returns tally_fx (int64 per slot, scale 2^32), events,
    status, hash."""
    ids = _i32a(vol_ids)
    fwd, rev = sense_tables(model)
    bct = bc_table(model, bc)
    if bct is not None and bct.size > fwd.size:
        pad = np.full(bct.size - fwd.size, -1, np.int32)
        fwd, rev = np.concatenate([fwd, pad]), np.concatenate([rev, pad])
    arr = (_vp * len(trees))(*[t.h for t in trees])
    pr = TransportProblem()
    pr.nvol = len(trees)
    pr.trees = C.cast(arr, C.POINTER(C.c_void_p))
    pr.vol_ids = _p(ids, _i32)
    pr.ipc = int(model.ipc)
    pr.nsurf_ids = fwd.size
    pr.sense_fwd = _p(fwd, _i32)
    pr.sense_rev = _p(rev, _i32)
    pr.seed = seed
    pr.stream = stream
    pr.lo = (C.c_double * 3)(*[float(x) for x in lo])
    pr.hi = (C.c_double * 3)(*[float(x) for x in hi])
    pr.mfp = float(mfp)
    pr.p_absorb = float(p_absorb)
    pr.max_events = int(max_events)
    pr.surf_bc = _p(bct, _u8)
    tally = np.zeros(len(trees), np.int64)
    events = np.zeros(n, np.int32)
    status = np.zeros(n, np.uint8)
    hashes = np.zeros(n, np.uint64)
    lib().xo_transport(C.byref(pr), int(first), int(n), nthreads,
                       tally.ctypes.data_as(C.POINTER(C.c_int64)),
                       _p(events, _i32), _p(status, _u8),
                       hashes.ctypes.data_as(C.POINTER(C.c_uint64)))
    return {"tally_fx": tally, "events": events, "status": status,
            "hash": hashes}


# ---- adjoint of the ray distance -------------------------------------------

def adjoint(verts, tris, org, dirs, t, face, dldt):
    """xo_adjoint: returns (dL/dverts (nv, 3), dL/dorg (n, 3), dL/ddir)."""
    verts, org, dirs, t, dldt = (_f64(verts), _f64(org), _f64(dirs), _f64(t),
                                 _f64(dldt))
    tris, face = _i32a(tris), _i32a(face)
    nv, n = verts.shape[0], org.shape[0]
    gv = np.zeros((nv, 3))
    go = np.zeros((n, 3))
    gd = np.zeros((n, 3))
    lib().xo_adjoint(nv, _p(verts, _d), tris.shape[0], _p(tris, _i32), n,
                     _p(org, _d), _p(dirs, _d), _p(t, _d), _p(face, _i32),
                     _p(dldt, _d), _p(gv, _d), _p(go, _d), _p(gd, _d))
    return gv, go, gd


# ---- overlap check (tools/overlap.cpp) ----------------------------------------

class OverlapQueries:
    """The overlap check's two batched queries composed from the oracle's
    point_in_volume and ray_fire (test infrastructure: the checker of
    xdg_hip_point_in_volumes / xdg_hip_edge_rays).  trees = {volume id:
    Tree} for every volume of the model, in model order."""

    def __init__(self, model, trees, max_k=4):
        self.vols = [v for v in model.volumes if v in trees]
        self.trees = [trees[v] for v in self.vols]
        self.max_k = max_k

    def inside_matrix(self, pts, dirs):
        """(n, nvol) booleans: point_in_volume of each location."""
        return np.stack([t.point_in_volume(pts, dirs) for t in self.trees], 1)

    def point_in_volumes(self, pts, dirs):
        m = self.inside_matrix(pts, dirs)
        cnt = m.sum(1).astype(np.int32)
        ids = np.full((pts.shape[0], self.max_k), -1, np.int32)
        full = {}
        vols = np.asarray(self.vols, np.int32)
        for i in np.nonzero(cnt)[0]:
            s = vols[m[i]]
            ids[i, :min(cnt[i], self.max_k)] = s[:self.max_k]
            if cnt[i] > self.max_k:
                full[int(i)] = s
        return cnt, ids, full

    def edge_rays(self, org, dirs, length, skip):
        n = org.shape[0]
        t = np.full(n, np.finfo(np.float64).max)
        surf = np.full(n, -1, np.int32)
        vol = np.full(n, -1, np.int32)
        todo = np.ones(n, bool)
        for v, tr in zip(self.vols, self.trees):
            sel = np.nonzero(todo & (skip[:, 0] != v) & (skip[:, 1] != v))[0]
            if sel.size == 0:
                continue
            tt, ss, _ = tr.ray_fire(org[sel], dirs[sel], tmax=length[sel],
                                    orientation=0)
            hit = ss >= 0
            t[sel[hit]] = tt[hit]
            surf[sel[hit]] = ss[hit]
            vol[sel[hit]] = v
            todo[sel[hit]] = False
        return t, surf, vol
