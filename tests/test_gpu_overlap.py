"""Overlap-check kernels (xdg_hip_point_in_volumes, xdg_hip_edge_rays) vs
the oracle's point_in_volume / ray_fire over the same volume lists, and the
overlap maps built on both (tools/overlap.cpp)."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import bits, build_tree, dev, host
from xdg_amd import geometry as G
from xdg_amd import overlap as OV
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


def _shifted_lattice():
    """3x3x3 spheres with one moved into its neighbour: a real overlap."""
    m = G.sphere_lattice(3, 2, 0.8, 2.0)
    f = m.surface_faces(1)
    vi = np.unique(m.tris[f])
    m.verts = m.verts.copy()
    m.verts[vi, 0] += 0.5
    return m


MODELS = {
    "overlapping_boxes": G.overlapping_boxes,
    "enclosed_boxes": G.enclosed_boxes,
    "crossed_bars": G.crossed_bars,
    "nested_spheres": lambda: G.nested_spheres(3),
    "shifted_lattice": _shifted_lattice,
}


def _setup(ctx, m):
    trees = {v: build_tree(ctx, m, v) for v in m.volumes}
    otrees = {v: O.Tree.from_model(m, v) for v in m.volumes}
    return OV.GpuQueries(ctx, m, trees, precision=1), O.OverlapQueries(m, otrees)


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("prec", [0, 1])
def test_point_in_volumes_and_edge_rays(ctx, torch_mod, name, prec):
    m = MODELS[name]()
    gq, oq = _setup(ctx, m)
    gq.precision = prec
    locs, dirs = OV.vertex_locations(m)
    lo, hi = m.verts.min(0) - 0.5, m.verts.max(0) + 0.5
    n = 4000
    pts = np.concatenate([locs, P.box_points(6, 0, n, lo, hi)])
    dd = np.concatenate([dirs, P.iso_dirs(6, 1, n)])
    gc, gi, gf = gq.point_in_volumes(pts, dd)
    oc, oi, of = oq.point_in_volumes(pts, dd)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gi, oi)
    assert gf.keys() == of.keys()
    for k in gf:
        assert np.array_equal(gf[k], of[k])
    org, d, ln, es = OV.edge_queries(m)
    skip = np.array([m.senses[int(s)] for s in es], np.int32).reshape(-1, 2)
    gt, gs, gv = gq.edge_rays(org, d, ln, skip)
    ot, os_, ov = oq.edge_rays(org, d, ln, skip)
    assert np.array_equal(gs, os_)
    assert np.array_equal(gv, ov)
    assert np.array_equal(bits(gt), bits(ot))


@pytest.mark.parametrize("name", sorted(MODELS))
def test_overlap_maps_match(ctx, name):
    m = MODELS[name]()
    gq, oq = _setup(ctx, m)
    for edges in (False, True):
        got = OV.check_instance_for_overlaps(gq, m, check_edges=edges)
        want = OV.check_instance_for_overlaps(oq, m, check_edges=edges)
        assert got == want
    if name == "shifted_lattice":
        assert frozenset({1, 2}) in got
    if name == "nested_spheres":
        assert got == {}


def test_overflowing_volume_lists(ctx, torch_mod):
    """Points inside more volumes than max_k: counts are complete and the
    harness fetches the full lists."""
    m = G.enclosed_boxes()
    gq, oq = _setup(ctx, m)
    gq.max_k = 1
    oq.max_k = 1
    pts = P.box_points(8, 0, 2000, [1.05, 1.2, 0.9], [2.0, 2.13, 1.61])
    dd = P.iso_dirs(8, 1, 2000)
    gc, gi, gf = gq.point_in_volumes(pts, dd)
    oc, oi, of = oq.point_in_volumes(pts, dd)
    assert (gc >= 2).all() and np.array_equal(gc, oc) and np.array_equal(gi, oi)
    assert gf.keys() == of.keys() and len(gf) == 2000
    for k in gf:
        assert np.array_equal(gf[k], of[k])
