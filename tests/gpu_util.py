"""Helpers shared by the GPU test modules."""
import numpy as np


def dev(torch, a, dtype=None):
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda").contiguous()


def host(t):
    return t.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def build_tree(ctx, model, vol):
    v9, f, s, fl = model.tree_arrays(vol)
    n = v9.shape[0]
    verts = v9.reshape(-1, 3)
    tris = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    return ctx.build_surface_tree(verts, tris, f, s, fl, model.bump(vol))
