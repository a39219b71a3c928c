"""Edge cases through the C-ABI: empty batches, a tree without triangles,
degenerate (zero-length) segments, rays with a zero distance limit, and the
exclusion-list bound."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import bits, build_tree, dev, host
from xdg_amd import geometry as G
from xdg_amd import philox as P

pytestmark = pytest.mark.gpu

INFTY = np.finfo(np.float64).max


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
def cube(ctx):
    mesh = G.kuhn_cube(6, 0.2, seed=3)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    ipc = build_tree(ctx, bm, bm.ipc)
    vol = build_tree(ctx, bm, 1)
    ctx.attach_boundary(et, ipc, mesh.bface_elem)
    return mesh, et, ipc, vol


def test_empty_batches(ctx, torch_mod, cube):
    torch = torch_mod
    mesh, et, ipc, vol = cube
    z3 = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    z1 = torch.zeros(0, dtype=torch.float64, device="cuda")
    zi = torch.zeros(0, dtype=torch.int32, device="cuda")
    for prec in (0, 1):
        t, s, f = ctx.ray_fire(vol, z3, z3, precision=prec)
        assert t.numel() == 0 and s.numel() == 0 and f.numel() == 0
        assert ctx.point_in_volume(vol, z3, z3, precision=prec).numel() == 0
        assert ctx.find_volume([vol], [1], 2, z3, z3, precision=prec).numel() == 0
        h, d = ctx.occluded(vol, z3, z3, precision=prec)
        assert h.numel() == 0
        c, i = ctx.point_in_volumes([vol, ipc], [1, 2], z3, z3, precision=prec)
        assert c.numel() == 0
        zs = torch.zeros((0, 2), dtype=torch.int32, device="cuda")
        t, s, v = ctx.edge_rays([vol, ipc], [1, 2], z3, z3, z1, zs, precision=prec)
        assert t.numel() == 0
    d, f = ctx.closest(vol, z3)
    assert d.numel() == 0
    assert ctx.find_element(et, z3).numel() == 0
    nx, dd = ctx.next_element(et, zi, z3, z3)
    assert nx.numel() == 0
    c, h = ctx.tally_segments(et, z3, z3)
    assert c.numel() == 0
    c, h = ctx.tally_segments_volume(et, vol, z3, z3)
    assert c.numel() == 0
    assert ctx.status() == 0


def test_tree_without_triangles(ctx, torch_mod):
    """Every query misses an empty tree, as the oracle's does."""
    torch = torch_mod
    e9 = np.zeros((0, 9))
    ei = np.zeros(0, np.int32)
    tree = ctx.build_surface_tree(np.zeros((0, 3)), np.zeros((0, 3), np.int32), ei, ei,
                                  np.zeros(0, np.uint8), 1e-3)
    ot = O.Tree(e9, ei, ei, np.zeros(0, np.uint8), 1e-3)
    pts = P.box_points(3, 0, 500, [-1.0] * 3, [1.0] * 3)
    dirs = P.iso_dirs(3, 1, 500)
    for prec in (0, 1):
        t, s, f = ctx.ray_fire(tree, dev(torch, pts), dev(torch, dirs), precision=prec)
        rt, rs, rf = ot.ray_fire(pts, dirs)
        assert np.array_equal(bits(host(t)), bits(rt)) and np.array_equal(host(s), rs)
        assert np.array_equal(host(f), rf)
        ins = host(ctx.point_in_volume(tree, dev(torch, pts), dev(torch, dirs), precision=prec))
        assert not ins.any() and not ot.point_in_volume(pts, dirs).any()
        h, d = ctx.occluded(tree, dev(torch, pts), dev(torch, dirs), precision=prec)
        rh, rd = ot.occluded(pts, dirs)
        assert np.array_equal(host(h).astype(bool), rh)
        assert np.array_equal(bits(host(d)), bits(rd))
    d, f = ctx.closest(tree, dev(torch, pts))
    rd, rf = ot.closest(pts)
    assert np.array_equal(bits(host(d)), bits(rd)) and np.array_equal(host(f), rf)


def test_degenerate_segments(ctx, torch_mod, cube):
    """Zero-length segments (start == end) give no entries in both forms,
    inside and outside the mesh, as the oracle's restatements do."""
    torch = torch_mod
    mesh, et, ipc, vol = cube
    om = O.TetMesh(mesh)
    ovol = O.Tree.from_model(mesh.boundary_model(), 1)
    p = np.concatenate([P.box_points(4, 0, 300, [0.0] * 3, [1.0] * 3),
                        P.box_points(4, 1, 300, [-1.0] * 3, [2.0] * 3)])
    rc, rh = om.segments(p, p)
    c, h = ctx.tally_segments(et, dev(torch, p), dev(torch, p))
    assert np.array_equal(host(c), rc) and np.array_equal(host(h).view(np.uint64), rh)
    assert (rc == 0).all()
    rc, rh, rf = om.segments_volume(ovol, p, p)
    c, h = ctx.tally_segments_volume(et, vol, dev(torch, p), dev(torch, p))
    assert np.array_equal(host(c), rc) and np.array_equal(host(h).view(np.uint64), rh)
    assert (rc == 0).all()
    assert ctx.status() == (2 if (rf & 2).any() else 0)


def test_zero_distance_limit(ctx, torch_mod):
    """dist_limit 0 accepts only hits at t == 0 (plucker.h:158 with
    tMin = 0); origins on the surface at a vertex give t == 0."""
    torch = torch_mod
    m = G.single_sphere(3, 10.0)
    tree = build_tree(ctx, m, 1)
    ot = O.Tree.from_model(m, 1)
    v9, _, _, _ = m.tree_arrays(1)
    n = 400
    org = np.concatenate([P.ball_points(5, 0, n, 9.0), v9[:n, 0:3]])
    dirs = P.iso_dirs(5, 1, 2 * n)
    tmax = np.zeros(2 * n)
    for prec in (0, 1):
        t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), tmax=dev(torch, tmax),
                               orientation=-1, precision=prec)
        rt, rs, rf = ot.ray_fire(org, dirs, tmax=tmax, orientation=-1)
        assert np.array_equal(host(s), rs) and np.array_equal(host(f), rf)
        assert np.array_equal(bits(host(t)), bits(rt))
    assert (rs[:n] == -1).all() and (rs[n:] >= 0).sum() > 0


def test_exclusion_bound(ctx, torch_mod):
    """1024 exclusion slots per ray are read; 1025 is rejected."""
    torch = torch_mod
    m = G.single_sphere(4, 10.0)  # 5120 faces
    tree = build_tree(ctx, m, 1)
    ot = O.Tree.from_model(m, 1)
    n = 64
    org = P.ball_points(7, 0, n, 9.0)
    dirs = P.iso_dirs(7, 1, n)
    excl = np.tile(np.arange(1024, dtype=np.int32)[::-1], (n, 1))  # faces 0..1023
    t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), excl=dev(torch, excl))
    rt, rs, rf = ot.ray_fire(org, dirs, excl=excl)
    assert np.array_equal(host(f), rf) and np.array_equal(bits(host(t)), bits(rt))
    assert (host(f)[host(s) >= 0] >= 1024).all()
    with pytest.raises(Exception):
        ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs),
                     excl=dev(torch, np.zeros((n, 1025), np.int32)))


def test_lattice_degenerate_segments(ctx, torch_mod):
    """Segments along the edges and faces of an unjittered Kuhn mesh and
    through its vertices (exit-face ties, zero-length entries): counts,
    sequence hashes and tallies equal to the oracle's, in the global and
    the volume form."""
    from lattice_cases import lattice_segments
    torch = torch_mod
    mesh = G.kuhn_cube(8, 0.0, seed=3)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    ipc = build_tree(ctx, bm, bm.ipc)
    vol = build_tree(ctx, bm, 1)
    ctx.attach_boundary(et, ipc, mesh.bface_elem)
    om = O.TetMesh(mesh)
    ovol = O.Tree.from_model(bm, 1)
    p0, p1 = lattice_segments(8)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    c, h = ctx.tally_segments(et, dev(torch, p0), dev(torch, p1), tally=tally)
    rc, rh, rt = om.segments(p0, p1, tally=True)
    assert np.array_equal(host(c), rc) and np.array_equal(host(h).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    rc, rh, rf = om.segments_volume(ovol, p0, p1)
    c, h = ctx.tally_segments_volume(et, vol, dev(torch, p0), dev(torch, p1))
    assert np.array_equal(host(c), rc) and np.array_equal(host(h).view(np.uint64), rh)
    assert ctx.status() == (2 if (rf & 2).any() else 0)


def test_segment_order_is_result_neutral(ctx, torch_mod):
    """Walking a batch in Morton order of the start points (the
    "segment_order" option) changes nothing: counts, sequence hashes and
    tallies equal the caller-order walk and the oracle, with re-entry."""
    torch = torch_mod
    mesh = G.kuhn_cube(16, 0.2, seed=4)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    v9, f, s, fl = bm.tree_arrays(bm.ipc)
    st = ctx.build_surface_tree(v9.reshape(-1, 3), np.arange(3 * v9.shape[0], dtype=np.int32).reshape(-1, 3),
                                f, s, fl, bm.bump(bm.ipc))
    ctx.attach_boundary(et, st, mesh.bface_elem)
    n = 100000
    p0 = P.box_points(21, P.STREAM_POS, n, [-0.1] * 3, [1.1] * 3)
    p1 = P.box_points(21, P.STREAM_DIR, n, [-0.1] * 3, [1.1] * 3)
    out = {}
    for mode in (0, 1):
        ctx.set_option("segment_order", mode)
        tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
        cnt, hsh = ctx.tally_segments(et, dev(torch, p0), dev(torch, p1), tally=tally)
        out[mode] = (host(cnt), host(hsh), host(tally))
    ctx.set_option("segment_order", 0)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    rc, rh, rt = O.TetMesh(mesh).segments(p0, p1, tally=True)
    assert np.array_equal(out[1][0], rc) and np.array_equal(out[1][1].view(np.uint64), rh)
    assert np.array_equal(out[1][2], rt)
