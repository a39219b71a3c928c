"""Argument checks, graph lifetime and tally range at the C-ABI boundary.

* The ctypes layer rejects device arrays of the wrong dtype, length or
  device before anything is launched (the C-ABI takes plain pointers and
  cannot size them itself).
* A captured transport graph is refused once the memory it points at may
  have been freed or moved: a tree built or released on the context, the
  transport destroyed.  The stale graph is never launched.
* The fixed-point tally guard sets XDG_HIP_STATUS_TALLY_RANGE when a tally
  entry has reached 2^62 (forced here by pre-filling the caller's tally
  array, which the walk adds to), and leaves it clear otherwise.
"""
import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import philox as P

from gpu_util import build_tree, dev, host

pytestmark = pytest.mark.gpu

STATUS_TALLY_RANGE = 4


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


@pytest.fixture(scope="module")
def sphere(ctx):
    m = G.single_sphere(3, 10.0)
    return m, build_tree(ctx, m, 1)


def test_query_arrays_are_validated(ctx, torch_mod, sphere):
    torch = torch_mod
    m, tree = sphere
    n = 64
    org = dev(torch, P.ball_points(1, 0, n, 9.0))
    dirs = dev(torch, P.iso_dirs(1, 1, n))
    with pytest.raises(TypeError):
        ctx.ray_fire(tree, org.float(), dirs)
    with pytest.raises(ValueError):
        ctx.ray_fire(tree, org, dirs[: n - 1])
    with pytest.raises(ValueError):
        ctx.ray_fire(tree, org.cpu(), dirs)
    with pytest.raises(TypeError):
        ctx.ray_fire(tree, org, dirs, excl=torch.zeros((n, 2), dtype=torch.int64, device="cuda"))
    with pytest.raises(ValueError):
        ctx.ray_fire(tree, org, dirs, tmax=torch.zeros(n - 3, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        ctx.ray_fire(tree, org, dirs, tree_per_ray=torch.ones(n, dtype=torch.int64, device="cuda"))
    with pytest.raises(ValueError):
        ctx.ray_fire(tree, org, dirs[:, :2].contiguous())
    with pytest.raises(ValueError):
        ctx.ray_fire(tree, org, dirs.t())
    with pytest.raises(TypeError):
        ctx.point_in_volume(tree, org, dirs.float())
    with pytest.raises(ValueError):
        ctx.closest(tree, org[:, :2].contiguous())
    # the same arrays, well formed, run
    t, s, f = ctx.ray_fire(tree, org, dirs)
    assert (host(s) >= 0).all()


def test_element_arrays_are_validated(ctx, torch_mod):
    torch = torch_mod
    mesh = G.kuhn_cube(4, 0.0)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    n = 16
    p0 = dev(torch, P.box_points(4, 0, n, [0] * 3, [1] * 3))
    p1 = dev(torch, P.box_points(4, 1, n, [0] * 3, [1] * 3))
    ne = mesh.tets.shape[0]
    with pytest.raises(ValueError):
        ctx.tally_segments(et, p0, p1, tally=torch.zeros(ne - 1, dtype=torch.int64, device="cuda"))
    with pytest.raises(TypeError):
        ctx.tally_segments(et, p0, p1, tally=torch.zeros(ne, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        ctx.find_element(et, p0.float())
    with pytest.raises(ValueError):
        ctx.next_element(et, torch.zeros(n, dtype=torch.int32, device="cuda"), p0, p1[: n - 1])
    cnt, _ = ctx.tally_segments(et, p0, p1, tally=torch.zeros(ne, dtype=torch.int64, device="cuda"))
    assert (host(cnt) > 0).all()


def test_adjoint_arrays_are_validated(ctx, torch_mod):
    torch = torch_mod
    from xdg_amd import adjoint as A
    m = G.single_sphere(2, 10.0)
    verts = torch.from_numpy(m.verts.copy()).cuda()
    surf = A.DifferentiableSurface(ctx, m.tris)
    n = 32
    org = dev(torch, P.ball_points(3, 0, n, 9.0))
    dirs = dev(torch, P.iso_dirs(3, 1, n))
    with pytest.raises(TypeError):
        A.ray_distance(org, dirs, verts.float(), surf)
    t = A.ray_distance(org, dirs, verts, surf)
    z = torch.zeros(n, dtype=torch.float64, device="cuda")
    f = torch.zeros(n, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        ctx.ray_distance_adjoint(org, dirs, t, f, z, verts, surf.tris,
                                 dldv=torch.zeros((3, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        ctx.ray_distance_adjoint(org, dirs, t, f.long(), z, verts, surf.tris)


def _lattice_transport(ctx):
    lat = G.sphere_lattice(2, 2, 0.8, 2.0)
    vols = [v for v in lat.volumes if v != lat.ipc]
    trees = [build_tree(ctx, lat, v) for v in vols]
    h = ctx.transport_create(vols, trees, lat.ipc, lat.senses, 3, 3, [0.0] * 3, [4.0] * 3, 2.0, 0.3)
    return lat, vols, trees, h


def test_stale_graph_is_refused(torch_mod):
    from xdg_amd import capi
    c = capi.Context(0)
    try:
        lat, vols, trees, h = _lattice_transport(c)
        out = c.transport_outputs(len(vols), 256)
        c.transport_run(h, 0, 256, len(vols), out=out)  # sizes the spill area
        g = c.transport_capture(h, 0, 256, out)
        m = G.single_sphere(2, 5.0)
        build_tree(c, m, 1)  # may regrow the tree table the graph points at
        with pytest.raises(capi.XdgHipError) as e:
            c.graph_launch(g)
        assert e.value.code == capi.ERR_INVALID and "stale" in str(e.value)
        c.graph_destroy(g)
        g = c.transport_capture(h, 0, 256, out)
        c.release_surface_tree(build_tree(c, m, 1))
        with pytest.raises(capi.XdgHipError):
            c.graph_launch(g)
        c.graph_destroy(g)
        g = c.transport_capture(h, 0, 256, out)
        c.transport_destroy(h)
        with pytest.raises(capi.XdgHipError):
            c.graph_launch(g)
        c.graph_destroy(g)
        c.synchronize()
    finally:
        c.close()


def test_fresh_graph_replays_after_unrelated_calls(torch_mod):
    """Queries that free nothing (ray_fire, tally on other arrays) keep a
    captured graph valid; its replay equals the eager run."""
    from xdg_amd import capi
    torch = torch_mod
    c = capi.Context(0)
    try:
        lat, vols, trees, h = _lattice_transport(c)
        ref = c.transport_run(h, 0, 512, len(vols))
        out = c.transport_outputs(len(vols), 512)
        c.transport_run(h, 0, 512, len(vols), out=out)
        g = c.transport_capture(h, 0, 512, out)
        org = dev(torch, P.ball_points(1, 0, 64, 0.5) + 1.0)
        dirs = dev(torch, P.iso_dirs(1, 1, 64))
        c.ray_fire(trees[0], org, dirs)
        out["tally_fx"].zero_()
        c.graph_launch(g)
        c.synchronize()
        assert torch.equal(out["tally_fx"], ref["tally_fx"])
        assert torch.equal(out["hash"], ref["hash"])
        c.graph_destroy(g)
        c.transport_destroy(h)
    finally:
        c.close()


def test_tally_range_bit(ctx, torch_mod):
    torch = torch_mod
    mesh = G.kuhn_cube(4, 0.0)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    ne = mesh.tets.shape[0]
    n = 64
    p0 = dev(torch, P.box_points(5, 0, n, [0] * 3, [1] * 3))
    p1 = dev(torch, P.box_points(5, 1, n, [0] * 3, [1] * 3))
    ctx.status()  # clear
    tally = torch.zeros(ne, dtype=torch.int64, device="cuda")
    ctx.tally_segments(et, p0, p1, tally=tally)
    assert ctx.status() & STATUS_TALLY_RANGE == 0
    # one entry already at 2^62 - 1: the walk's additions reach the guard
    tally = torch.zeros(ne, dtype=torch.int64, device="cuda")
    om = O.TetMesh(mesh, with_ipc=False)
    hit = int(om.find_element(host(p0)[:1])[0])
    tally[hit] = (1 << 62) - 1
    ctx.tally_segments(et, p0, p1, tally=tally)
    assert ctx.status() & STATUS_TALLY_RANGE
    assert ctx.status() == 0  # reading clears it
