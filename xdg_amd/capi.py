"""ctypes binding of libxdg_hip.so (the C-ABI in include/xdg_hip.h).

torch is imported first so that this process uses a single HIP runtime
(torch's); tensors passed to the query functions must live on the context's
device.  The library is built in-tree by __graft_entry__.build()
(xdg_amd/csrc/Makefile) and is loaded from xdg_amd/lib; a missing library
raises: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch  # noqa: F401  (must precede loading libxdg_hip.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libxdg_hip.so")

OK = 0
ERR_INVALID = -1
ERR_NOT_FOUND = -2
ERR_HIP = -3
ERR_NO_DEVICE = -4
ERR_CAPPED = -5

PLAIN = 0
MIXED = 1
PERSIST_DEFAULT = 7  # persistent ray_fire for both precisions, and find_volume
ANY = -1
EXITING = 0
ENTERING = 1
MAX_EXCLUDE = 1024
NOALLOC_MAX_RAYS = 4194304  # XDG_HIP_NOALLOC_MAX_RAYS

INFTY = np.finfo(np.float64).max

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64


class XdgHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"xdg_hip error {code}: {msg}")
        self.code = code


class TreeInfo(C.Structure):
    _fields_ = [("n_prims", C.c_int64), ("n_nodes", C.c_int64),
                ("depth", C.c_int32), ("root", C.c_int32),
                ("bytes", C.c_int64), ("bump", C.c_double),
                ("o_limit", C.c_double), ("n_nodes4", C.c_int64),
                ("stack4", C.c_int32), ("reserved", C.c_int32)]


class TransportProblem(C.Structure):
    """xdg_hip_transport_problem (include/xdg_hip.h)."""
    _fields_ = [("nvol", _i32), ("vol_ids", _vp), ("vol_trees", _vp),
                ("ipc", _i32), ("nsurf_ids", _i32), ("sense_fwd", _vp),
                ("sense_rev", _vp), ("seed", C.c_uint32),
                ("stream", C.c_uint32), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("mfp", C.c_double),
                ("p_absorb", C.c_double), ("max_events", _i32),
                ("surf_bc", _vp)]


# per-history transport status codes (XDG_HIP_TR_*)
TR_ABSORBED, TR_LEAKED, TR_LOST, TR_STUCK, TR_CAPPED, TR_ERROR, TR_VACUUM = range(1, 8)
# surface boundary conditions (XDG_HIP_BC_*)
BC_TRANSMISSION, BC_VACUUM, BC_REFLECTING = 0, 1, 2

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C xdg_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "xdg_hip_create": [C.POINTER(_vp), C.c_int],
            "xdg_hip_destroy": [_vp],
            "xdg_hip_set_launch": [_vp, C.c_int, C.c_int],
            "xdg_hip_set_traversal": [_vp, C.c_int, C.c_int, C.c_int],
            "xdg_hip_set_option": [_vp, C.c_char_p, C.c_int],
            "xdg_hip_get_option": [_vp, C.c_char_p, _vp],
            "xdg_hip_get_launch": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "xdg_hip_synchronize": [_vp, _vp],
            "xdg_hip_build_surface_tree": [_vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                           _vp, C.c_double, C.POINTER(_i32)],
            "xdg_hip_build_element_tree": [_vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                           C.POINTER(_i32)],
            "xdg_hip_attach_boundary": [_vp, _i32, _i32, _i64, _vp],
            "xdg_hip_get_tree_info": [_vp, _i32, C.c_int, C.POINTER(TreeInfo)],
            "xdg_hip_num_trees": [_vp, C.POINTER(_i32), C.POINTER(_i32)],
            "xdg_hip_release_surface_tree": [_vp, _i32],
            "xdg_hip_refit_surface_tree": [_vp, _i32, _i64, _vp, C.c_double],
            "xdg_hip_ray_fire": [_vp, _i32, _vp, _i64, _vp, _vp, _vp, C.c_int,
                                 _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp],
            "xdg_hip_ray_fire_noalloc": [_vp, _i32, _vp, _i64, _vp, _vp, _vp,
                                         C.c_int, _vp, C.c_int, C.c_int, _vp,
                                         _vp, _vp, _vp],
            "xdg_hip_ray_fire_stats": [_vp, _i32, _vp, _i64, _vp, _vp, _vp,
                                       C.c_int, _vp, C.c_int, C.c_int, _vp],
            "xdg_hip_point_in_volume": [_vp, _i32, _i64, _vp, _vp, _vp, C.c_int,
                                        C.c_int, _vp, _vp],
            "xdg_hip_find_volume": [_vp, C.c_int, _vp, _vp, _i32, _i64, _vp,
                                    _vp, C.c_int, _vp, _vp],
            "xdg_hip_occluded": [_vp, _i32, _i64, _vp, _vp, C.c_int, _vp, _vp,
                                 _vp],
            "xdg_hip_point_in_volumes": [_vp, C.c_int, _vp, _vp, _i64, _vp,
                                         _vp, C.c_int, C.c_int, _vp, _vp,
                                         _vp],
            "xdg_hip_edge_rays": [_vp, C.c_int, _vp, _vp, _i64, _vp, _vp, _vp,
                                  _vp, C.c_int, _vp, _vp, _vp, _vp],
            "xdg_hip_closest": [_vp, _i32, _i64, _vp, _vp, _vp, _vp],
            "xdg_hip_find_element": [_vp, _i32, _i64, _vp, _vp, _vp],
            "xdg_hip_next_element": [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp,
                                     _vp],
            "xdg_hip_tally_segments": [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp],
            "xdg_hip_tally_segments_volume": [_vp, _i32, _i32, _i64, _vp, _vp,
                                              _vp, _vp, _vp, _vp, _vp, _vp,
                                              _vp],
            "xdg_hip_status": [_vp, C.POINTER(C.c_uint32)],
            "xdg_hip_mesh_walk": [_vp, _i32, C.c_uint64, _i64, C.c_uint32,
                                  C.c_uint32, _vp, _vp, C.c_double, _i32, _vp,
                                  _vp, _vp, _vp, _vp, _vp],
            "xdg_hip_eval_primitive": [_vp, C.c_int, _i64, _vp, _vp, _vp],
            "xdg_hip_transport_create": [_vp, C.POINTER(TransportProblem),
                                         C.POINTER(_i32)],
            "xdg_hip_transport_destroy": [_vp, _i32],
            "xdg_hip_transport_set_config": [_vp, _i32, C.c_int, C.c_int,
                                             C.c_int, C.c_int],
            "xdg_hip_transport_get_config": [_vp, _i32, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int),
                                             C.POINTER(C.c_int),
                                             C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)],
            "xdg_hip_transport_run": [_vp, _i32, C.c_uint64, _i64, C.c_int,
                                      _vp, _vp, _vp, _vp, _vp],
            "xdg_hip_transport_capture": [_vp, _i32, C.c_uint64, _i64,
                                          C.c_int, _vp, _vp, _vp, _vp,
                                          C.POINTER(_vp)],
            "xdg_hip_graph_launch": [_vp, _vp, _vp],
            "xdg_hip_graph_destroy": [_vp, _vp],
            "xdg_hip_transport_autotune": [_vp, _i32, C.c_uint64, _i64,
                                           C.c_int, C.POINTER(C.c_double)],
            "xdg_hip_ray_distance_adjoint": [_vp, _i64, _vp, _vp, _vp, _vp,
                                             _vp, _i64, _vp, _i64, _vp, _vp,
                                             _vp, _vp, _vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.xdg_hip_last_error.argtypes = [_vp]
        L.xdg_hip_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


F64 = (torch.float64,)
I32 = (torch.int32,)
I64 = (torch.int64, torch.uint64)  # 64-bit tallies and hashes (two's complement)
U8 = (torch.uint8,)


def _ptr(x):
    """Device pointer of an already checked torch tensor (or None)."""
    if x is None:
        return None
    return _vp(x.data_ptr())


def _count(x):
    """Rows of a query array (its leading dimension)."""
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"expected a torch tensor, got {type(x)}")
    return x.shape[0] if x.dim() > 0 else 0


def _ordered(fn):
    """Without an explicit stream a query runs on the context's stream:
    first let torch's current stream finish producing the inputs."""
    def wrapper(self, *args, **kwargs):
        if kwargs.get("stream") is None:
            torch.cuda.current_stream(self.dev).synchronize()
        return fn(self, *args, **kwargs)
    wrapper.__name__ = fn.__name__
    wrapper.__doc__ = fn.__doc__
    return wrapper


def _host(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_vp)


class Context:
    """Owns one xdg_hip context on one device."""

    def __init__(self, device: int = 0):
        self.device = device
        h = _vp()
        rc = lib().xdg_hip_create(C.byref(h), device)
        if rc != OK:
            raise XdgHipError(rc, "xdg_hip_create failed (no usable GPU?)")
        self.h = h
        self.dev = torch.device("cuda", device)

    def close(self):
        if getattr(self, "h", None):
            lib().xdg_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arg(self, x, name, dtypes, numel=None, required=True, min_numel=None):
        """Validate one device array before any launch: a contiguous torch
        tensor on this context's device, of one of `dtypes`, holding exactly
        `numel` (or at least `min_numel`) elements.  The C-ABI takes plain
        pointers and cannot check sizes itself, so a wrong array here would
        be read or written out of bounds."""
        if x is None:
            if required:
                raise ValueError(f"{name} is required")
            return None
        if not isinstance(x, torch.Tensor):
            raise TypeError(f"{name}: expected a torch tensor, got {type(x)}")
        if not x.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
        if x.device.index != self.device:
            raise ValueError(f"{name} is on {x.device}, the context on cuda:{self.device}")
        if x.dtype not in dtypes:
            raise TypeError(f"{name} must be {' or '.join(str(d) for d in dtypes)}, got {x.dtype}")
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if numel is not None and x.numel() != numel:
            raise ValueError(f"{name} must hold {numel} elements, has {x.numel()}")
        if min_numel is not None and x.numel() < min_numel:
            raise ValueError(f"{name} must hold at least {min_numel} elements, has {x.numel()}")
        return x

    def _prims(self, tree, element=False):
        """Triangles / elements of a tree (sizes the per-element tallies)."""
        return self.tree_info(tree, element)["n_prims"]

    def _check(self, rc):
        if rc != OK:
            msg = lib().xdg_hip_last_error(self.h)
            raise XdgHipError(rc, msg.decode() if msg else "")
        return rc

    def _fence(self, stream):
        """Without an explicit stream, wait for torch's current stream once
        the wrapper's own allocations (a zero fill, an upload) are queued:
        the context's stream is not ordered with it."""
        if stream is None:
            torch.cuda.current_stream(self.dev).synchronize()

    def _done(self, rc, stream):
        """Check a query's status; without an explicit stream the call ran on
        the context's own stream, so wait for it before torch reads."""
        self._check(rc)
        if stream is None:
            self._check(lib().xdg_hip_synchronize(self.h, None))

    # ---- configuration ---------------------------------------------------
    def set_launch(self, block, stack):
        self._check(lib().xdg_hip_set_launch(self.h, block, stack))

    def set_traversal(self, wide=True, defer=True, persistent=None):
        """persistent: None for the default (PERSIST_DEFAULT, every
        kernel), True / False for every kernel, or the raw mask (bit 0 plain
        ray_fire, bit 1 mixed ray_fire, bit 2 find_volume)."""
        if persistent is None:
            mask = PERSIST_DEFAULT
        elif persistent is True or persistent is False:
            mask = 7 if persistent else 0
        else:
            mask = int(persistent)
        self._check(lib().xdg_hip_set_traversal(self.h, int(wide), int(defer), mask))

    def set_option(self, name, value):
        """Tuning option by name (see include/xdg_hip.h)."""
        self._check(lib().xdg_hip_set_option(self.h, name.encode(), int(value)))

    def get_option(self, name):
        v = C.c_int()
        self._check(lib().xdg_hip_get_option(self.h, name.encode(), C.byref(v)))
        return v.value

    def set_ray_order(self, mode):
        self.set_option("ray_order", mode)

    def get_launch(self):
        b, s = C.c_int(), C.c_int()
        self._check(lib().xdg_hip_get_launch(self.h, C.byref(b), C.byref(s)))
        return b.value, s.value

    def synchronize(self, stream=None):
        self._check(lib().xdg_hip_synchronize(self.h, stream))

    # ---- builds ------------------------------------------------------------
    def build_surface_tree(self, verts, tris, face_ids, surf_ids, flip, bump):
        v, vp = _host(verts, np.float64)
        t, tp = _host(tris, np.int32)
        f, fp = _host(face_ids, np.int32)
        s, sp = _host(surf_ids, np.int32)
        fl, flp = _host(flip, np.uint8)
        out = _i32()
        self._check(lib().xdg_hip_build_surface_tree(
            self.h, v.size // 3, vp, t.size // 3, tp, fp, sp, flp, float(bump),
            C.byref(out)))
        return out.value

    def release_surface_tree(self, tree):
        self._check(lib().xdg_hip_release_surface_tree(self.h, tree))

    def refit_surface_tree(self, tree, verts, bump=-1.0):
        """New vertex coordinates (device float64 (nv, 3), the build's nv)
        for a tree built with the "refit" option; bump < 0 keeps the old."""
        nv = _count(verts)
        self._arg(verts, "verts", F64, 3 * nv)
        torch.cuda.current_stream(self.dev).synchronize()
        self._check(lib().xdg_hip_refit_surface_tree(self.h, tree, nv, _ptr(verts), float(bump)))

    def build_element_tree(self, verts, tets, adj, face_order):
        v, vp = _host(verts, np.float64)
        t, tp = _host(tets, np.int32)
        a, ap = _host(adj, np.int32)
        fo, fop = _host(face_order, np.int32)
        out = _i32()
        self._check(lib().xdg_hip_build_element_tree(
            self.h, v.size // 3, vp, t.size // 4, tp, ap, fop, C.byref(out)))
        return out.value

    def attach_boundary(self, element_tree, surface_tree, bface_elem):
        b, bp = _host(bface_elem, np.int32)
        self._check(lib().xdg_hip_attach_boundary(self.h, element_tree,
                                                  surface_tree, b.size, bp))

    def tree_info(self, tree, element=False):
        info = TreeInfo()
        self._check(lib().xdg_hip_get_tree_info(self.h, tree, int(element),
                                                C.byref(info)))
        return {k: getattr(info, k) for k, _ in TreeInfo._fields_}

    def num_trees(self):
        s, e = _i32(), _i32()
        self._check(lib().xdg_hip_num_trees(self.h, C.byref(s), C.byref(e)))
        return s.value, e.value

    # ---- queries (device tensors) ------------------------------------------
    def _rays(self, org, dirs, dirs_required=True, names=("org", "dirs")):
        n = _count(org)
        self._arg(org, names[0], F64, 3 * n)
        self._arg(dirs, names[1], F64, 3 * n, required=dirs_required)
        return n

    def _excl(self, excl, n):
        if excl is None:
            return 0
        k = excl.shape[1] if excl.dim() == 2 else -1
        if k < 0 or k > MAX_EXCLUDE:
            raise ValueError(f"excl must be (n, k) with k <= {MAX_EXCLUDE}")
        self._arg(excl, "excl", I32, k * n)
        return k

    @_ordered
    def ray_fire(self, tree, org, dirs, tmax=None, orientation=EXITING,
                 excl=None, precision=PLAIN, tree_per_ray=None, out=None,
                 stream=None):
        n = self._rays(org, dirs)
        self._arg(tmax, "tmax", F64, n, required=False)
        self._arg(tree_per_ray, "tree_per_ray", I32, n, required=False)
        k = self._excl(excl, n)
        if out is None:
            out = (torch.empty(n, dtype=torch.float64, device=org.device),
                   torch.empty(n, dtype=torch.int32, device=org.device),
                   torch.empty(n, dtype=torch.int32, device=org.device))
        t, s, f = out
        self._arg(t, "t_out", F64, n)
        self._arg(s, "surf_out", I32, n)
        self._arg(f, "face_out", I32, n, required=False)
        self._done(lib().xdg_hip_ray_fire(
            self.h, tree, _ptr(tree_per_ray), n, _ptr(org), _ptr(dirs),
            _ptr(tmax), orientation, _ptr(excl), k, precision, _ptr(t),
            _ptr(s), _ptr(f), stream), stream)
        return t, s, f

    @_ordered
    def ray_fire_noalloc(self, tree, org, dirs, out, tmax=None, orientation=EXITING,
                         excl=None, precision=PLAIN, tree_per_ray=None, stream=None):
        """xdg_hip_ray_fire_noalloc: ray_fire into caller-provided outputs
        (t, surf, face) with no device allocation by the library; at most
        NOALLOC_MAX_RAYS rays; calls that overlap on one GPU race."""
        n = self._rays(org, dirs)
        self._arg(tmax, "tmax", F64, n, required=False)
        self._arg(tree_per_ray, "tree_per_ray", I32, n, required=False)
        k = self._excl(excl, n)
        t, s, f = out
        self._arg(t, "t_out", F64, n)
        self._arg(s, "surf_out", I32, n)
        self._arg(f, "face_out", I32, n, required=False)
        self._done(lib().xdg_hip_ray_fire_noalloc(
            self.h, tree, _ptr(tree_per_ray), n, _ptr(org), _ptr(dirs),
            _ptr(tmax), orientation, _ptr(excl), k, precision, _ptr(t),
            _ptr(s), _ptr(f), stream), stream)
        return t, s, f

    @_ordered
    def ray_fire_stats(self, tree, org, dirs, tmax=None, orientation=EXITING,
                       excl=None, precision=PLAIN, tree_per_ray=None,
                       stream=None):
        """Work counters of the instrumented traversal (synchronous)."""
        n = self._rays(org, dirs)
        self._arg(tmax, "tmax", F64, n, required=False)
        self._arg(tree_per_ray, "tree_per_ray", I32, n, required=False)
        k = self._excl(excl, n)
        cnt = (C.c_uint64 * 12)()
        self._check(lib().xdg_hip_ray_fire_stats(
            self.h, tree, _ptr(tree_per_ray), n, _ptr(org), _ptr(dirs),
            _ptr(tmax), orientation, _ptr(excl), k, precision, cnt))
        keys = ["rays", "nodes", "tri_tests", "edge_tests", "finishes",
                "normals", "hits", "fallbacks", "wave_node_iters",
                "wave_tri_iters", "wave_max_nodes", "wave_max_tris"]
        return dict(zip(keys, [int(x) for x in cnt]))

    @_ordered
    def point_in_volume(self, tree, pts, dirs=None, excl=None, precision=PLAIN,
                        stream=None):
        n = self._rays(pts, dirs, dirs_required=False, names=("pts", "dirs"))
        k = self._excl(excl, n)
        out = torch.empty(n, dtype=torch.uint8, device=pts.device)
        self._done(lib().xdg_hip_point_in_volume(
            self.h, tree, n, _ptr(pts), _ptr(dirs), _ptr(excl), k, precision,
            _ptr(out), stream), stream)
        return out

    @_ordered
    def find_volume(self, vol_trees, vol_ids, ipc, pts, dirs=None,
                    precision=PLAIN, out=None, stream=None):
        n = self._rays(pts, dirs, dirs_required=False, names=("pts", "dirs"))
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=pts.device)
        self._arg(out, "vol_out", I32, n)
        vt, vtp = _host(vol_trees, np.int32)
        vi, vip = _host(vol_ids, np.int32)
        if vt.size != vi.size:
            raise ValueError("vol_trees and vol_ids differ in length")
        self._done(lib().xdg_hip_find_volume(
            self.h, vt.size, vtp, vip, ipc, n, _ptr(pts), _ptr(dirs), precision,
            _ptr(out), stream), stream)
        return out

    @_ordered
    def point_in_volumes(self, vol_trees, vol_ids, pts, dirs, max_k=4,
                         precision=PLAIN, stream=None):
        """point_in_volume of every point against every listed volume
        (tools/overlap.cpp:9-61).  Returns (count (n,), ids (n, max_k))."""
        n = self._rays(pts, dirs, names=("pts", "dirs"))
        cnt = torch.empty(n, dtype=torch.int32, device=pts.device)
        ids = torch.full((n, max_k), -1, dtype=torch.int32, device=pts.device)
        vt, vtp = _host(vol_trees, np.int32)
        vi, vip = _host(vol_ids, np.int32)
        if vt.size != vi.size:
            raise ValueError("vol_trees and vol_ids differ in length")
        self._done(lib().xdg_hip_point_in_volumes(
            self.h, vt.size, vtp, vip, n, _ptr(pts), _ptr(dirs), max_k,
            precision, _ptr(cnt), _ptr(ids), stream), stream)
        return cnt, ids

    @_ordered
    def edge_rays(self, vol_trees, vol_ids, org, dirs, length, skip,
                  precision=PLAIN, stream=None):
        """First EXITING hit within length over the listed volumes, skipping
        two volume ids per ray (tools/overlap.cpp:229-265).  Returns
        (t, surface, volume)."""
        n = self._rays(org, dirs)
        self._arg(length, "length", F64, n)
        self._arg(skip, "skip", I32, 2 * n)
        t = torch.empty(n, dtype=torch.float64, device=org.device)
        surf = torch.empty(n, dtype=torch.int32, device=org.device)
        vol = torch.empty(n, dtype=torch.int32, device=org.device)
        vt, vtp = _host(vol_trees, np.int32)
        vi, vip = _host(vol_ids, np.int32)
        if vt.size != vi.size:
            raise ValueError("vol_trees and vol_ids differ in length")
        self._done(lib().xdg_hip_edge_rays(
            self.h, vt.size, vtp, vip, n, _ptr(org), _ptr(dirs), _ptr(length),
            _ptr(skip), precision, _ptr(t), _ptr(surf), _ptr(vol), stream),
            stream)
        return t, surf, vol

    @_ordered
    def occluded(self, tree, org, dirs, precision=PLAIN, stream=None):
        n = self._rays(org, dirs)
        hit = torch.empty(n, dtype=torch.uint8, device=org.device)
        dist = torch.empty(n, dtype=torch.float64, device=org.device)
        self._done(lib().xdg_hip_occluded(self.h, tree, n, _ptr(org),
                                           _ptr(dirs), precision, _ptr(hit),
                                           _ptr(dist), stream), stream)
        return hit, dist

    @_ordered
    def closest(self, tree, pts, stream=None):
        n = _count(pts)
        self._arg(pts, "pts", F64, 3 * n)
        dist = torch.empty(n, dtype=torch.float64, device=pts.device)
        face = torch.empty(n, dtype=torch.int32, device=pts.device)
        self._done(lib().xdg_hip_closest(self.h, tree, n, _ptr(pts),
                                          _ptr(dist), _ptr(face), stream), stream)
        return dist, face

    @_ordered
    def find_element(self, etree, pts, stream=None):
        n = _count(pts)
        self._arg(pts, "pts", F64, 3 * n)
        out = torch.empty(n, dtype=torch.int32, device=pts.device)
        self._done(lib().xdg_hip_find_element(self.h, etree, n, _ptr(pts),
                                               _ptr(out), stream), stream)
        return out

    @_ordered
    def next_element(self, etree, elem, r, u, stream=None):
        n = _count(elem)
        self._arg(elem, "elem", I32, n)
        self._arg(r, "r", F64, 3 * n)
        self._arg(u, "u", F64, 3 * n)
        nx = torch.empty(n, dtype=torch.int32, device=r.device)
        d = torch.empty(n, dtype=torch.float64, device=r.device)
        self._done(lib().xdg_hip_next_element(self.h, etree, n, _ptr(elem),
                                               _ptr(r), _ptr(u), _ptr(nx),
                                               _ptr(d), stream), stream)
        return nx, d

    def _segment_args(self, etree, p0, p1, tally, seq_off, seq_elem, seq_len, out):
        n = self._rays(p0, p1, names=("p0", "p1"))
        self._arg(tally, "tally", I64, self._prims(etree, element=True), required=False)
        if seq_off is not None:
            self._arg(seq_off, "seq_off", (torch.int64,), n + 1)
            total = int(seq_off[-1].item()) if n > 0 else 0
            self._arg(seq_elem, "seq_elem", I32, min_numel=total)
            self._arg(seq_len, "seq_len", F64, min_numel=total)
        if out is None:
            out = (torch.empty(n, dtype=torch.int32, device=p0.device),
                   torch.empty(n, dtype=torch.int64, device=p0.device))
        self._arg(out[0], "count_out", I32, n)
        self._arg(out[1], "hash_out", I64, n)
        return n, out

    @_ordered
    def tally_segments(self, etree, p0, p1, tally=None, seq_off=None,
                       seq_elem=None, seq_len=None, out=None, stream=None):
        n, (cnt, hsh) = self._segment_args(etree, p0, p1, tally, seq_off, seq_elem, seq_len, out)
        self._done(lib().xdg_hip_tally_segments(
            self.h, etree, n, _ptr(p0), _ptr(p1), _ptr(tally), _ptr(cnt),
            _ptr(hsh), _ptr(seq_off), _ptr(seq_elem), _ptr(seq_len), stream), stream)
        return cnt, hsh

    @_ordered
    def tally_segments_volume(self, etree, stree, p0, p1, tally=None,
                              seq_off=None, seq_elem=None, seq_len=None,
                              out=None, stream=None):
        """XDG::segments(volume, start, end) (src/xdg.cpp:200-227): etree and
        stree are the volume's element and surface trees."""
        n, (cnt, hsh) = self._segment_args(etree, p0, p1, tally, seq_off, seq_elem, seq_len, out)
        self._done(lib().xdg_hip_tally_segments_volume(
            self.h, etree, stree, n, _ptr(p0), _ptr(p1), _ptr(tally),
            _ptr(cnt), _ptr(hsh), _ptr(seq_off), _ptr(seq_elem),
            _ptr(seq_len), stream), stream)
        return cnt, hsh

    @_ordered
    def mesh_walk(self, etree, first, n, seed, stream_id, lo, hi, mfp,
                  max_events=100000, tally=None, stream=None):
        """Element-by-element particle walk (tools/walk_elements.h:53-117).
        Returns (distance, events, status, hash) device tensors."""
        self._arg(tally, "tally", I64, self._prims(etree, element=True), required=False)
        dev = self.dev
        dist = torch.empty(n, dtype=torch.float64, device=dev)
        ev = torch.empty(n, dtype=torch.int32, device=dev)
        st = torch.empty(n, dtype=torch.uint8, device=dev)
        hs = torch.empty(n, dtype=torch.int64, device=dev)
        lo_a = (C.c_double * 3)(*[float(x) for x in lo])
        hi_a = (C.c_double * 3)(*[float(x) for x in hi])
        self._done(lib().xdg_hip_mesh_walk(
            self.h, etree, int(first), int(n), int(seed), int(stream_id),
            C.cast(lo_a, _vp), C.cast(hi_a, _vp), float(mfp), int(max_events),
            _ptr(dist), _ptr(ev), _ptr(st), _ptr(hs), _ptr(tally), stream),
            stream)
        return dist, ev, st, hs

    # ---- batched particle random walk ------------------------------------
    def transport_create(self, vol_ids, vol_trees, ipc, senses, seed, stream,
                         lo, hi, mfp, p_absorb, max_events=10000, bc=None):
        """Register a transport problem.  senses: {surface id: (forward
        volume, reverse volume)}; bc: {surface id: BC_*} (others transmit).
        Returns the handle."""
        ids = np.ascontiguousarray(vol_ids, dtype=np.int32)
        trees = np.ascontiguousarray(vol_trees, dtype=np.int32)
        ns = max(max(senses), max(bc) if bc else 0) + 1
        bct = None
        if bc:
            bct = np.zeros(ns, np.uint8)
            for sid, code in bc.items():
                bct[sid] = code
        fwd = np.full(ns, -1, np.int32)
        rev = np.full(ns, -1, np.int32)
        for sid, (f, r) in senses.items():
            fwd[sid] = f
            rev[sid] = r
        p = TransportProblem()
        p.nvol = ids.size
        p.vol_ids = ids.ctypes.data
        p.vol_trees = trees.ctypes.data
        p.ipc = int(ipc)
        p.nsurf_ids = ns
        p.sense_fwd = fwd.ctypes.data
        p.sense_rev = rev.ctypes.data
        p.seed = int(seed)
        p.stream = int(stream)
        p.lo = (C.c_double * 3)(*[float(x) for x in lo])
        p.hi = (C.c_double * 3)(*[float(x) for x in hi])
        p.mfp = float(mfp)
        p.p_absorb = float(p_absorb)
        p.max_events = int(max_events)
        p.surf_bc = bct.ctypes.data if bct is not None else None
        h = _i32()
        self._check(lib().xdg_hip_transport_create(self.h, C.byref(p), C.byref(h)))
        return h.value

    def transport_destroy(self, handle):
        self._check(lib().xdg_hip_transport_destroy(self.h, handle))

    def transport_set_config(self, handle, block=0, lds_stack=0, chunk=0,
                             refill=0):
        self._check(lib().xdg_hip_transport_set_config(self.h, handle, block,
                                                       lds_stack, chunk, refill))

    def transport_config(self, handle):
        v = [C.c_int() for _ in range(5)]
        self._check(lib().xdg_hip_transport_get_config(
            self.h, handle, *[C.byref(x) for x in v]))
        return dict(zip(["block", "lds_stack", "chunk", "refill", "stack_need"],
                        [x.value for x in v]))

    def transport_outputs(self, nvol, n, per_history=True):
        z = lambda *a, **k: torch.zeros(*a, device=self.dev, **k)
        out = {"tally_fx": z(nvol, dtype=torch.int64)}
        if per_history:
            out.update(events=z(n, dtype=torch.int32),
                       status=z(n, dtype=torch.uint8),
                       hash=z(n, dtype=torch.int64))
        return out

    def _transport_out(self, out, nvol, n):
        self._arg(out["tally_fx"], "tally_fx", I64, nvol)
        self._arg(out.get("events"), "events", I32, n, required=False)
        self._arg(out.get("status"), "status", U8, n, required=False)
        self._arg(out.get("hash"), "hash", I64, n, required=False)

    @_ordered
    def transport_run(self, handle, first, n, nvol, precision=MIXED,
                      out=None, per_history=True, stream=None):
        if out is None:
            out = self.transport_outputs(nvol, n, per_history)
            self._fence(stream)
        self._transport_out(out, nvol, n)
        self._done(lib().xdg_hip_transport_run(
            self.h, handle, int(first), int(n), precision,
            _ptr(out["tally_fx"]), _ptr(out.get("events")),
            _ptr(out.get("status")), _ptr(out.get("hash")), stream), stream)
        return out

    def transport_capture(self, handle, first, n, out, precision=MIXED):
        """Instantiated graph of one transport launch writing into out."""
        self._transport_out(out, out["tally_fx"].numel(), n)
        g = _vp()
        self._check(lib().xdg_hip_transport_capture(
            self.h, handle, int(first), int(n), precision,
            _ptr(out["tally_fx"]), _ptr(out.get("events")),
            _ptr(out.get("status")), _ptr(out.get("hash")), C.byref(g)))
        return g.value

    @_ordered
    def graph_launch(self, graph, stream=None):
        self._check(lib().xdg_hip_graph_launch(self.h, graph, stream))

    def graph_destroy(self, graph):
        self._check(lib().xdg_hip_graph_destroy(self.h, graph))

    def transport_autotune(self, handle, first, sample, precision=MIXED):
        ms = C.c_double()
        self._check(lib().xdg_hip_transport_autotune(
            self.h, handle, int(first), int(sample), precision, C.byref(ms)))
        return ms.value

    # ---- adjoint of the ray distance --------------------------------------
    @_ordered
    def ray_distance_adjoint(self, org, dirs, t, face, dldt, verts, tris,
                             dldv=None, dldo=None, dldd=None, stream=None):
        """dL/d(verts) accumulated into dldv; dL/d(org), dL/d(dir) written
        into dldo / dldd (all device tensors; None skips)."""
        n = self._rays(org, dirs)
        nv, ntri = _count(verts), _count(tris)
        self._arg(t, "t", F64, n)
        self._arg(face, "face", I32, n)
        self._arg(dldt, "dldt", F64, n)
        self._arg(verts, "verts", F64, 3 * nv)
        self._arg(tris, "tris", I32, 3 * ntri)
        self._arg(dldv, "dldv", F64, 3 * nv, required=False)
        self._arg(dldo, "dldo", F64, 3 * n, required=False)
        self._arg(dldd, "dldd", F64, 3 * n, required=False)
        self._done(lib().xdg_hip_ray_distance_adjoint(
            self.h, n, _ptr(org), _ptr(dirs), _ptr(t), _ptr(face), _ptr(dldt),
            nv, _ptr(verts), ntri, _ptr(tris),
            _ptr(dldv), _ptr(dldo), _ptr(dldd), stream), stream)

    def status(self):
        f = C.c_uint32()
        self._check(lib().xdg_hip_status(self.h, C.byref(f)))
        return f.value

    def eval_primitive(self, kind, inputs, width_out, stream=None):
        inp = torch.as_tensor(np.ascontiguousarray(inputs, dtype=np.float64),
                              device=self.dev).contiguous()
        n = inp.shape[0]
        out = torch.zeros((n, width_out), dtype=torch.float64, device=self.dev)
        self._fence(stream)
        self._check(lib().xdg_hip_eval_primitive(self.h, kind, n, _ptr(inp),
                                                 _ptr(out), stream))
        self.synchronize(stream)
        return out.cpu().numpy()
