"""Element-by-element particle walk (xdg_hip_mesh_walk) vs the oracle's
replay of the same Philox streams: distance bits, events, status, the
(element, length) hash and per-element tallies, bit for bit."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import bits, build_tree, host
from xdg_amd import geometry as G

pytestmark = pytest.mark.gpu

SEED, STREAM = 13, 4


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


def _setup(ctx, mesh):
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    st = build_tree(ctx, bm, bm.ipc)
    ctx.attach_boundary(et, st, mesh.bface_elem)
    return et


CASES = {
    "kuhn16": (lambda: G.kuhn_cube(16, 0.2, seed=4), [0.0] * 3, [1.0] * 3, 0.05),
    "two_blocks": (lambda: G.two_blocks(8, 0.5), [0.0] * 3, [2.5, 1.0, 1.0], 0.3),
    "mock": (G.mock_tetmesh, [-2.0, -3.0, -4.0], [5.0, 6.0, 7.0], 2.0),
}


@pytest.fixture
def walk_form(ctx):
    """Sets (mw_wave, mw_refill) for one test and restores the defaults."""
    saved = ctx.get_option("mw_wave"), ctx.get_option("mw_refill")

    def set_form(wave, refill):
        ctx.set_option("mw_wave", wave)
        ctx.set_option("mw_refill", refill)
    yield set_form
    set_form(*saved)


# (mw_wave, mw_refill): parking rounds one particle per lane, inline
# re-entry, and parking rounds on a resident grid refilled in batches of
# 32 lanes (the default), 16 lanes and single lanes
FORMS = [(1, 0), (0, 0), (1, 32), (1, 16), (1, 1)]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", FORMS)
def test_mesh_walk_matches_oracle(ctx, torch_mod, walk_form, name, form):
    torch = torch_mod
    make, lo, hi, mfp = CASES[name]
    mesh = make()
    et = _setup(ctx, mesh)
    om = O.TetMesh(mesh)
    n, first = 20000, 321
    rd, re, rs, rh, rt = om.mesh_walk(first, n, SEED, STREAM, lo, hi, mfp, tally=True)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    walk_form(*form)
    d, e, s, h = ctx.mesh_walk(et, first, n, SEED, STREAM, lo, hi, mfp, tally=tally)
    assert np.array_equal(host(s), rs)
    assert np.array_equal(host(e), re)
    assert np.array_equal(bits(host(d)), bits(rd))
    assert np.array_equal(host(h).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    assert set(np.unique(rs)) == {O.TR_LEAKED}
    assert re.mean() > 3


@pytest.mark.parametrize("refill", [0, 40])
def test_mesh_walk_many_reentries(ctx, torch_mod, walk_form, refill):
    """Two blocks with a narrow gap and a long mean free path: particles
    cross the gap several times, more often than the batched re-entry
    rounds, so the last walk re-enters inline; small batches (below one
    wave) too.  Equal to the oracle."""
    mesh = G.two_blocks(6, 0.05)
    et = _setup(ctx, mesh)
    om = O.TetMesh(mesh)
    walk_form(1, refill)
    for n in (37, 6000):
        rd, re, rs, rh = om.mesh_walk(5, n, SEED, STREAM, [0.0] * 3, [2.05, 1.0, 1.0], 0.4)
        d, e, s, h = ctx.mesh_walk(et, 5, n, SEED, STREAM, [0.0] * 3, [2.05, 1.0, 1.0], 0.4)
        assert np.array_equal(host(e), re) and np.array_equal(bits(host(d)), bits(rd))
        assert np.array_equal(host(h).view(np.uint64), rh) and np.array_equal(host(s), rs)


@pytest.mark.parametrize("refill", [1, 16, 40])
def test_mesh_walk_resident_many_per_lane(ctx, torch_mod, walk_form, refill):
    """2^19 particles, more than the resident grid has lanes, so lanes of
    the resident walk claim several particles each (while other lanes of
    their wave are still stepping, and in every re-entry round): equal, bit
    for bit and tally for tally, to the walk with one particle per lane,
    which the oracle pins above."""
    torch = torch_mod
    mesh = G.kuhn_cube(16, 0.2, seed=4)
    et = _setup(ctx, mesh)
    n = 1 << 19
    out = []
    for form in ((1, 0), (1, refill)):
        walk_form(*form)
        tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
        d, e, s, h = ctx.mesh_walk(et, 7, n, SEED, STREAM, [0.0] * 3, [1.0] * 3, 0.05, tally=tally)
        out.append((bits(host(d)), host(e), host(s), host(h), host(tally)))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert out[0][1].mean() > 3


@pytest.mark.parametrize("form", [(1, 0), (1, 32), (0, 0)])
def test_mesh_walk_table_face_order(ctx, torch_mod, walk_form, form):
    """A face order that is neither MOAB's nor MeshMock's (MOAB's faces in
    the order 1, 2, 3, 0, each with its outward winding), so the walk reads
    the order from the tree: every walk form equals the oracle, which walks
    the same order."""
    torch = torch_mod
    base = G.kuhn_cube(12, 0.2, seed=4)
    fo = np.ascontiguousarray(base.face_order[[1, 2, 3, 0]], dtype=np.int32)
    adj, bf, be = G.tet_adjacency(base.tets, fo, base.verts.shape[0])
    mesh = G.TetMesh(base.verts, base.tets, adj, fo, bf, be)
    et = _setup(ctx, mesh)
    om = O.TetMesh(mesh)
    n, first = 12000, 77
    rd, re, rs, rh, rt = om.mesh_walk(first, n, SEED, STREAM, [0.0] * 3, [1.0] * 3, 0.05, tally=True)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    walk_form(*form)
    d, e, s, h = ctx.mesh_walk(et, first, n, SEED, STREAM, [0.0] * 3, [1.0] * 3, 0.05, tally=tally)
    assert np.array_equal(host(s), rs) and np.array_equal(host(e), re)
    assert np.array_equal(bits(host(d)), bits(rd))
    assert np.array_equal(host(h).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    assert re.mean() > 3


def test_mesh_walk_event_cap(ctx, torch_mod):
    mesh = G.kuhn_cube(8, 0.2, seed=4)
    et = _setup(ctx, mesh)
    om = O.TetMesh(mesh)
    rd, re, rs, rh = om.mesh_walk(0, 4000, SEED, STREAM, [0.0] * 3, [1.0] * 3, 0.01, max_events=20)
    d, e, s, h = ctx.mesh_walk(et, 0, 4000, SEED, STREAM, [0.0] * 3, [1.0] * 3, 0.01, max_events=20)
    assert (rs == O.TR_CAPPED).sum() > 1000
    assert np.array_equal(host(s), rs) and np.array_equal(host(e), re)
    assert np.array_equal(bits(host(d)), bits(rd))
    assert np.array_equal(host(h).view(np.uint64), rh)
