"""The benchmark workloads at their full sizes, through the C-ABI, against
the oracle run on the same inputs (bit for bit):
  * one closed icosphere (subdivision 5, r 10), 1 M rays from the ball of
    radius 9, plain and mixed ray_fire;
  * three nested icospheres (subdivision 6) in a cube, 10 M points in
    [-16, 16]^3: find_volume in both precisions, and ray_fire in the found
    volume;
  * the 64^3 x 6 Kuhn cube (1.57 M tets), 4 M segments: counts, sequence
    hashes and per-element tallies;
  * the 8^3 sphere lattice, 2^20 particle histories: tallies, events,
    statuses and history hashes."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import bits, build_tree, dev, host
from xdg_amd import geometry as G
from xdg_amd import philox as P

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


def test_single_sphere_1m_rays(ctx, torch_mod):
    torch = torch_mod
    m = G.single_sphere(5, 10.0)
    assert m.tris.shape[0] == 20480
    tree, ot = build_tree(ctx, m, 1), O.Tree.from_model(m, 1)
    n = 1_000_000
    org = P.ball_points(1, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(1, P.STREAM_DIR, n)
    rt, rs, rf = ot.ray_fire(org, dirs)
    assert (rs == 1).all()
    for prec in (0, 1):
        t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), precision=prec)
        assert np.array_equal(host(s), rs) and np.array_equal(host(f), rf)
        assert np.array_equal(bits(host(t)), bits(rt))


def test_nested_spheres_10m_points(ctx, torch_mod):
    torch = torch_mod
    m = G.nested_spheres(6)
    vols = list(m.volumes)
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    n = 10_000_000
    pts = P.box_points(2, P.STREAM_POS, n, [-16.0] * 3, [16.0] * 3)
    dirs = P.iso_dirs(2, P.STREAM_DIR, n)
    ref = O.find_volume(otrees, vols, m.ipc, pts, dirs)
    dp, dd = dev(torch, pts), dev(torch, dirs)
    for prec in (0, 1):
        assert np.array_equal(host(ctx.find_volume(trees, vols, m.ipc, dp, dd, precision=prec)), ref)
    # ray_fire in the containing volume for a 2 M subsample (the oracle's
    # share of the CPU time stays small)
    sel = np.arange(0, n, 5)
    tpr = np.array([trees[vols.index(v)] for v in ref[sel]], np.int32)
    t, s, f = ctx.ray_fire(0, dev(torch, pts[sel]), dev(torch, dirs[sel]),
                           tree_per_ray=dev(torch, tpr), precision=1)
    t, s, f = host(t), host(s), host(f)
    for v in vols:
        k = ref[sel] == v
        if not k.any():
            continue
        rt, rs, rf = otrees[vols.index(v)].ray_fire(pts[sel][k], dirs[sel][k])
        assert np.array_equal(s[k], rs) and np.array_equal(f[k], rf)
        assert np.array_equal(bits(t[k]), bits(rt))


def test_kuhn64_4m_segments(ctx, torch_mod):
    torch = torch_mod
    mesh = G.kuhn_cube(64, 0.2, seed=4)
    assert mesh.tets.shape[0] == 1572864
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    ctx.attach_boundary(et, build_tree(ctx, bm, bm.ipc), mesh.bface_elem)
    om = O.TetMesh(mesh)
    n = 4_000_000
    p0 = P.box_points(4, P.STREAM_POS, n, [0.0] * 3, [1.0] * 3)
    p1 = P.box_points(4, P.STREAM_DIR, n, [0.0] * 3, [1.0] * 3)
    rc, rh, rt = om.segments(p0, p1, tally=True)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    c, h = ctx.tally_segments(et, dev(torch, p0), dev(torch, p1), tally=tally)
    assert np.array_equal(host(c), rc)
    assert np.array_equal(host(h).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    assert ctx.status() == 0
    chord = np.sqrt(((p1 - p0) ** 2).sum(1)).sum()
    assert abs(rt.sum() / 2.0 ** 40 - chord) < 1e-6 * chord


def test_lattice_2e20_histories(ctx, torch_mod):
    from xdg_amd import capi
    m = G.sphere_lattice(8, 4, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    n = 1 << 20
    r = O.transport(m, otrees, vols, 3, 3, [0.0] * 3, [16.0] * 3, 2.0, 0.3, 100000, 0, n)
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, 3, 3, [0.0] * 3, [16.0] * 3,
                             2.0, 0.3, 100000)
    try:
        out = ctx.transport_run(h, 0, n, len(vols), precision=capi.MIXED)
    finally:
        ctx.transport_destroy(h)
    assert np.array_equal(host(out["tally_fx"]), r["tally_fx"])
    assert np.array_equal(host(out["events"]), r["events"])
    assert np.array_equal(host(out["status"]), r["status"])
    assert np.array_equal(host(out["hash"]).view(np.uint64), r["hash"])
