"""Overlap check (tools/overlap.cpp) on the CPU: the map logic of
xdg_amd/overlap.py driven by the oracle's point_in_volume and ray_fire,
against the known answers of synthetic analogues of the reference's
tests/test_overlap_check.cpp models (its .h5m files are not available):
overlapping boxes, imprinted and enclosed boxes, crossed bars whose
overlap only the edge rays find, and nested spheres (no overlap)."""
import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import overlap as OV


def _queries(m):
    return O.OverlapQueries(m, {v: O.Tree.from_model(m, v) for v in m.volumes})


def test_overlap_direction_and_locations():
    d = OV.overlap_direction()
    assert np.allclose(d, np.full(3, 1 / np.sqrt(3.0)), rtol=0, atol=3e-16)
    m = G.overlapping_boxes()
    locs, dirs = OV.vertex_locations(m)
    v = OV.model_vertices(m)
    assert locs.shape == (2 * 36 * 2, 3) and v.shape == (72, 3)
    assert np.allclose(locs[0::2] - v, 1e-9 * d, atol=1e-15)
    assert np.allclose(locs[1::2] - v, -1e-9 * d, atol=1e-15)
    assert np.array_equal(dirs[1::2], -dirs[0::2])


def test_edge_queries_cover_every_edge():
    m = G.crossed_bars()
    org, d, ln, s = OV.edge_queries(m)
    assert org.shape == (3 * 24, 3)
    ends = org + d * ln[:, None]
    tri = m.verts[m.tris[np.concatenate([m.surface_faces(x) for x in m.surfaces])]]
    assert np.allclose(ends, tri[:, [1, 2, 0]].reshape(-1, 3), atol=1e-14)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)


def _brute_inside(model, vol, pts, dirs):
    """point_in_volume by brute force with the Moeller-Trumbore ray-triangle
    test (an independent formula): sign of d . n (sense flipped in the
    reverse volume) at the nearest hit of the volume's triangles."""
    v9, _, _, flip = model.tree_arrays(vol)
    v0, v1, v2 = v9[:, 0:3], v9[:, 3:6], v9[:, 6:9]
    e1, e2 = v1 - v0, v2 - v0
    n = np.cross(e1, e2)
    n = np.where(flip[:, None] == 1, -n, n)
    out = np.zeros(pts.shape[0], bool)
    for i in range(pts.shape[0]):
        o, d = pts[i], dirs[i]
        p = np.cross(d, e2)
        det = (e1 * p).sum(1)
        ok = np.abs(det) > 1e-300
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        s = o - v0
        u = (s * p).sum(1) * inv
        q = np.cross(s, e1)
        w = (q * d).sum(1) * inv
        t = (q * e2).sum(1) * inv
        hit = ok & (u >= 0) & (w >= 0) & (u + w <= 1) & (t >= 0)
        if hit.any():
            j = np.nonzero(hit)[0][np.argmin(t[hit])]
            out[i] = (d * n[j]).sum() > 0
    return out


BOX_MODELS = [G.overlapping_boxes, G.imprinted_boxes, G.enclosed_boxes, G.crossed_bars]


@pytest.mark.parametrize("make", BOX_MODELS)
def test_oracle_point_in_volumes_brute_force(make):
    """The oracle's point_in_volume at every vertex location equals an
    independent brute-force classification."""
    m = make()
    locs, dirs = OV.vertex_locations(m)
    got = _queries(m).inside_matrix(locs, dirs)
    want = np.stack([_brute_inside(m, v, locs, dirs) for v in m.volumes], 1)
    assert np.array_equal(got, want)


def _expected_keys(m):
    locs, dirs = OV.vertex_locations(m)
    inside = np.stack([_brute_inside(m, v, locs, dirs) for v in m.volumes], 1)
    vols = np.asarray(m.volumes)
    return {frozenset(int(x) for x in vols[r]) for r in inside if r.sum() > 1}


def test_overlapping_boxes():
    """Each corner inside the other box gives {1, 2}; points of the overlap
    also see the other box's surface from the implicit complement's side,
    which the check reports as {1, 3} and {2, 3}, as the reference's
    point_in_volume over every volume does."""
    m = G.overlapping_boxes()
    got = OV.check_instance_for_overlaps(_queries(m), m)
    assert set(got) == _expected_keys(m)
    assert frozenset({1, 2}) in got
    # the last {1, 2} location in check order: B's corner (1, 1.3, 0.7) + d 1e-9
    assert np.allclose(got[frozenset({1, 2})], [1.0, 1.3, 0.7], atol=1e-8)


def test_no_overlap():
    """Imprinted boxes with the vertex check (as the reference's test runs
    them: their edges run along the neighbour's faces, where edge rays meet
    it at t = 0), nested spheres with both checks."""
    m = G.imprinted_boxes()
    assert OV.check_instance_for_overlaps(_queries(m), m) == {}
    m = G.nested_spheres(2)
    assert OV.check_instance_for_overlaps(_queries(m), m, check_edges=True) == {}


def test_enclosed_box():
    m = G.enclosed_boxes()
    got = OV.check_instance_for_overlaps(_queries(m), m)
    assert set(got) == _expected_keys(m)
    assert frozenset({1, 2}) in got


def test_crossed_bars_need_the_edge_rays():
    m = G.crossed_bars()
    q = _queries(m)
    assert _expected_keys(m) == set()
    assert OV.check_instance_for_overlaps(q, m) == {}
    got = OV.check_instance_for_overlaps(q, m, check_edges=True)
    assert list(got) == [frozenset({1, 2})]
    x, y, z = got[frozenset({1, 2})]
    # the edge hit lies on B's face |x| = 0.5 inside A, or on A's boundary
    # inside B
    on_b = np.isclose(abs(x), 0.5) and abs(y) <= 0.5 and abs(z) <= 0.1
    on_a = (np.isclose(abs(y), 0.5) or np.isclose(abs(z), 0.1)) and abs(x) <= 0.5
    assert on_a or on_b
