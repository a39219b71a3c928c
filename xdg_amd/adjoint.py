"""Differentiable ray distance: torch.autograd.Function over the device
ray_fire (forward) and the hand-written adjoint kernel (backward).

    surf = DifferentiableSurface(ctx, tris, surf_ids, flip, bump)
    t = ray_distance(org, dirs, verts, surf)      # verts: (nv, 3) float64
    (w * t).sum().backward()                      # -> verts.grad, org.grad ...

The surface tree is built once (with the library's "refit" option) and
refitted on the device whenever the vertex values differ from those of the
last refit (compared on the device), so an optimiser stepping `verts` sees
fresh geometry without a rebuild; the connectivity is fixed.  Rays that miss
give t = DBL_MAX and zero gradients.
No counterpart in the reference (its queries are not differentiable).
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi
from . import geometry as G


class DifferentiableSurface:
    """One closed surface (a single volume's triangles) whose vertices are a
    tensor.  tris: (nt, 3) int; face ids are row indices of tris."""

    def __init__(self, ctx: capi.Context, tris, surf_ids=None, flip=None,
                 bump=None, orientation=capi.EXITING, precision=capi.MIXED):
        self.ctx = ctx
        self.tris_host = np.ascontiguousarray(tris, dtype=np.int32)
        nt = self.tris_host.shape[0]
        self.surf_ids = (np.ones(nt, np.int32) if surf_ids is None
                         else np.ascontiguousarray(surf_ids, dtype=np.int32))
        self.flip = (np.zeros(nt, np.uint8) if flip is None
                     else np.ascontiguousarray(flip, dtype=np.uint8))
        self.face_ids = np.arange(nt, dtype=np.int32)
        self.bump = bump
        self.orientation = orientation
        self.precision = precision
        self.tris = torch.from_numpy(self.tris_host).to(ctx.dev)
        self._key = None
        self.tree = None

    def _bump(self, vd):
        if self.bump is not None:
            return self.bump
        lo = torch.clamp(vd.min(dim=0).values, max=0.0).cpu().numpy()
        hi = torch.clamp(vd.max(dim=0).values, min=0.0).cpu().numpy()
        return G.box_bump(lo, hi)

    def tree_for(self, verts: torch.Tensor) -> int:
        vd = verts.detach().contiguous()
        if self.tree is not None and self._key is not None and self._key.shape == vd.shape:
            if not torch.equal(self._key, vd):
                self.ctx.refit_surface_tree(self.tree, vd, self._bump(vd))
                self._key = vd.clone()
            return self.tree
        if self.tree is not None:
            self.ctx.release_surface_tree(self.tree)
            self.tree = None
        v = vd.cpu().numpy().astype(np.float64)
        keep = self.ctx.get_option("refit")
        self.ctx.set_option("refit", 1)
        try:
            self.tree = self.ctx.build_surface_tree(
                v, self.tris_host, self.face_ids, self.surf_ids, self.flip, self._bump(vd))
        finally:
            self.ctx.set_option("refit", keep)
        self._key = vd.clone()
        return self.tree


class RayDistance(torch.autograd.Function):
    """t_i = distance along ray i to the surface (ray_fire); backward runs
    xdg_hip_ray_distance_adjoint."""

    @staticmethod
    def forward(ctx, org, dirs, verts, surf: DifferentiableSurface):
        for name, x in (("org", org), ("dirs", dirs), ("verts", verts)):
            if x.dtype != torch.float64:
                raise TypeError(f"ray_distance: {name} must be float64, got {x.dtype}")
        tree = surf.tree_for(verts)
        o = org.detach().contiguous()
        d = dirs.detach().contiguous()
        t, s, f = surf.ctx.ray_fire(tree, o, d, orientation=surf.orientation,
                                    precision=surf.precision)
        ctx.surf = surf
        ctx.save_for_backward(o, d, verts.detach().contiguous(), t, f)
        return t

    @staticmethod
    def backward(ctx, gt):
        o, d, v, t, f = ctx.saved_tensors
        surf = ctx.surf
        gt = gt.contiguous().to(torch.float64)
        gv = torch.zeros_like(v) if ctx.needs_input_grad[2] else None
        go = torch.empty_like(o) if ctx.needs_input_grad[0] else None
        gd = torch.empty_like(d) if ctx.needs_input_grad[1] else None
        surf.ctx.ray_distance_adjoint(o, d, t, f, gt, v, surf.tris, gv, go, gd)
        return go, gd, gv, None


def ray_distance(org, dirs, verts, surf: DifferentiableSurface):
    return RayDistance.apply(org, dirs, verts, surf)
