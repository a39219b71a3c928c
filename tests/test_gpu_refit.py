"""Refit of a surface tree (xdg_hip_refit_surface_tree) on the GPU.

A tree built with the "refit" option keeps its topology; new vertex
coordinates update the records, normals, both node layouts and the
mixed-precision limit on the device.  Queries on the refitted tree must
equal the oracle on the moved geometry bit for bit (the traversal is exact
for any topology, only its speed depends on how tight the boxes are), for
both precisions and both tree widths, also after a large motion and after
moving back.  The differentiable-surface harness uses the refit.
"""
import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import philox as P

from gpu_util import bits, dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_mod):
    from xdg_amd import capi
    c = capi.Context(0)
    yield c
    c.close()


def _moved(m, verts):
    return G.SurfaceModel(verts, m.tris, m.tri_surf, m.surfaces, m.senses, m.volumes, ipc=m.ipc)


def _check(ctx, torch, tree, model, org, dirs):
    ot = O.Tree.from_model(model, 1)
    for orientation in (0, -1):
        rt, rs, rf = ot.ray_fire(org, dirs, orientation=orientation)
        assert (rs >= 0).mean() > 0.9
        for wide in (True, False):
            ctx.set_traversal(wide=wide)
            try:
                for prec in (0, 1):
                    t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), orientation=orientation,
                                           precision=prec)
                    assert np.array_equal(host(s), rs) and np.array_equal(host(f), rf)
                    assert np.array_equal(bits(host(t)), bits(rt))
            finally:
                ctx.set_traversal()
    inside = ot.point_in_volume(org, dirs)
    assert np.array_equal(host(ctx.point_in_volume(tree, dev(torch, org), dev(torch, dirs), precision=1)).astype(bool),
                          inside)


def test_refit_matches_oracle_on_moved_vertices(ctx, torch_mod):
    torch = torch_mod
    m = G.single_sphere(4, 10.0)
    v0 = m.verts.copy()
    tris = m.tris.astype(np.int32)
    faces = np.arange(tris.shape[0], dtype=np.int32)
    surfs = np.ones(tris.shape[0], np.int32)
    flip = np.zeros(tris.shape[0], np.uint8)
    ctx.set_option("refit", 1)
    try:
        tree = ctx.build_surface_tree(v0, tris, faces, surfs, flip, m.bump(1))
    finally:
        ctx.set_option("refit", 0)
    info0 = ctx.tree_info(tree)
    n = 20000
    org = P.ball_points(8, P.STREAM_POS, n, 8.0)
    dirs = P.iso_dirs(8, P.STREAM_DIR, n)
    _check(ctx, torch, tree, m, org, dirs)
    # a smooth radial perturbation, then a large scaling, then back
    r = np.sqrt((v0 ** 2).sum(1))
    th = np.arccos(np.clip(v0[:, 2] / r, -1, 1))
    ph = np.arctan2(v0[:, 1], v0[:, 0])
    v1 = v0 * (1.0 + 0.05 * np.sin(3 * th) * np.cos(2 * ph))[:, None]
    m1 = _moved(m, v1)
    ctx.refit_surface_tree(tree, dev(torch, v1), G.vertex_bump(v1))
    _check(ctx, torch, tree, m1, org, dirs)
    v2 = v0 * 1.5 + np.array([3.0, -2.0, 1.0])
    m2 = _moved(m, v2)
    ctx.refit_surface_tree(tree, dev(torch, v2), G.vertex_bump(v2))
    info2 = ctx.tree_info(tree)
    assert info2["o_limit"] != info0["o_limit"] and info2["bump"] == G.vertex_bump(v2)
    _check(ctx, torch, tree, m2, org * 0.9 + np.array([3.0, -2.0, 1.0]), dirs)
    ctx.refit_surface_tree(tree, dev(torch, v0), m.bump(1))
    _check(ctx, torch, tree, m, org, dirs)
    assert ctx.tree_info(tree)["o_limit"] == info0["o_limit"]


def test_refit_needs_the_option(ctx, torch_mod):
    from xdg_amd import capi
    torch = torch_mod
    m = G.single_sphere(2, 10.0)
    tris = m.tris.astype(np.int32)
    nt = tris.shape[0]
    tree = ctx.build_surface_tree(m.verts, tris, np.arange(nt, dtype=np.int32), np.ones(nt, np.int32),
                                  np.zeros(nt, np.uint8), m.bump(1))
    with pytest.raises(capi.XdgHipError, match="refit"):
        ctx.refit_surface_tree(tree, dev(torch, m.verts))
    ctx.set_option("refit", 1)
    try:
        t2 = ctx.build_surface_tree(m.verts, tris, np.arange(nt, dtype=np.int32), np.ones(nt, np.int32),
                                    np.zeros(nt, np.uint8), m.bump(1))
    finally:
        ctx.set_option("refit", 0)
    with pytest.raises(capi.XdgHipError, match="n_verts"):
        ctx.refit_surface_tree(t2, dev(torch, m.verts[:-1]))


def test_differentiable_surface_refits(ctx, torch_mod):
    """The harness refits (same tree id) and its distances follow the
    vertices: equal to the oracle after each step."""
    torch = torch_mod
    from xdg_amd import adjoint as A
    m = G.single_sphere(3, 10.0)
    surf = A.DifferentiableSurface(ctx, m.tris)
    n = 4000
    org = dev(torch, P.ball_points(9, P.STREAM_POS, n, 8.0))
    dirs = dev(torch, P.iso_dirs(9, P.STREAM_DIR, n))
    verts = torch.from_numpy(m.verts.copy()).cuda()
    t0 = A.ray_distance(org, dirs, verts, surf)
    tree = surf.tree
    for step in range(3):
        verts = verts * 1.01 + 0.02
        t = A.ray_distance(org, dirs, verts, surf)
        assert surf.tree == tree
        ref = O.Tree.from_model(_moved(m, host(verts)), 1)
        rt, _, _ = ref.ray_fire(host(org), host(dirs))
        assert np.array_equal(bits(host(t)), bits(rt)), step
    assert not torch.equal(t, t0)
