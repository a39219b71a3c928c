"""Pins of the oracle's find_element and tet containment.

find_element (ray_tracer.cpp:226-257 with tetrahedron_contain.cpp:47-114)
returns the lowest-index element whose containment test passes.  It is
pinned here without reference to its own code:

  * on the unjittered Kuhn cube the containing element follows from the
    point alone (tests/kuhn_cases.py): uniform points away from every face
    have one element; dyadic points on shared faces, edges and vertices have
    several, and the lowest index is the answer;
  * on a jittered mesh the element-tree search equals brute force over all
    elements (lowest containing index), also for points a few ulps off the
    mesh boundary, which only the per-element box dilation keeps;
  * planted mistakes fail those checks: returning the first element the tree
    visits instead of the lowest index, and building the element boxes
    without their dilation.

tet containment (tetrahedron_contain.cpp:11-43) is pinned by barycentric
identities: lambda(sum w_i v_i) = w for weights on the simplex, vertices map
to unit vectors, a vertex permutation permutes lambda; a transposed inverse
(planted) fails them.  Its lambda bits at exact face ties depend on the
absent linalg library and stay unpinned.
"""
import itertools

import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G

import kuhn_cases as K


@pytest.fixture(scope="module")
def cube16():
    m = G.kuhn_cube(16, 0.0)
    return m, O.TetMesh(m, with_ipc=False)


@pytest.fixture(scope="module")
def jit16():
    m = G.kuhn_cube(16, 0.2, seed=9)
    return m, O.TetMesh(m, with_ipc=False)


def test_analytic_interior_points(cube16):
    _, om = cube16
    pts, exp = K.interior_points(16, 4000, seed=1)
    got = om.find_element(pts)
    assert np.array_equal(got, exp)
    assert np.array_equal(om.find_element(pts, brute=True), exp)


def test_analytic_shared_faces_lowest_index(cube16):
    _, om = cube16
    pts = K.dyadic_points(16, 3, 3000, seed=2)
    exp = np.array([K.lowest_containing(16, p) for p in pts], dtype=np.int32)
    assert (exp >= 0).all()
    got = om.find_element(pts)
    assert np.array_equal(got, exp)
    # the ties are real: many of these points lie in more than one element
    multi = sum(1 for p in pts[:400] if _n_containing(p) > 1)
    assert multi > 100


def _n_containing(p):
    s = [v * 16 for v in p]
    cnt = 0
    cand = []
    for a in range(3):
        c = int(np.floor(s[a]))
        opts = {c, c - 1} if s[a] == c else {c}
        cand.append([x for x in opts if 0 <= x < 16])
    for cx, cy, cz in itertools.product(*cand):
        cnt += len(K._contained_exact((s[0] - cx, s[1] - cy, s[2] - cz)))
    return cnt


def test_outside_points_find_nothing(cube16):
    _, om = cube16
    pts = np.array([[-0.01, 0.5, 0.5], [0.5, 1.01, 0.5], [0.5, 0.5, 2.0], [-1.0, -1.0, -1.0]])
    assert (om.find_element(pts) == -1).all()
    assert (om.find_element(pts, brute=True) == -1).all()


def _near_boundary_points(count, seed):
    """Points just outside a face of the unit cube, closer than the
    containment tolerance (20 eps in barycentric units, 16 * 20 eps here):
    k * 2^-56 below 0, or one ulp above 1."""
    rng = np.random.default_rng(seed)
    p = rng.random((count, 3)) * 0.98 + 0.01
    ax = rng.integers(0, 3, count)
    side = rng.integers(0, 2, count)
    k = rng.integers(1, 5, count)
    for i in range(count):
        p[i, ax[i]] = np.nextafter(1.0, 2.0) if side[i] else -float(k[i]) * 2.0 ** -56
    return p


def test_tree_equals_brute_force_jittered(jit16):
    m, om = jit16
    rng = np.random.default_rng(3)
    pts = [rng.random((3000, 3))]
    # vertices, edge midpoints and face centroids of random elements, and
    # points a few ulps around them: where neighbouring elements tie
    e = rng.integers(0, m.tets.shape[0], 1500)
    v = m.verts[m.tets[e]]
    pts.append(v[:, rng.integers(0, 4)])
    pts.append(0.5 * (v[:, 0] + v[:, 1]))
    pts.append((v[:, 0] + v[:, 1] + v[:, 2]) / 3.0)
    q = (v[:, 1] + v[:, 2] + v[:, 3]) / 3.0
    pts.append(np.nextafter(q, q + rng.choice([-1.0, 1.0], q.shape)))
    pts.append(_near_boundary_points(1500, 4))
    pts = np.concatenate(pts)
    tree = om.find_element(pts)
    brute = om.find_element(pts, brute=True)
    assert np.array_equal(tree, brute)
    # the boundary points are inside by the containment tolerance only
    assert (brute[-1500:] >= 0).mean() > 0.99


def test_planted_first_visited_fails(cube16):
    _, om = cube16
    pts = K.dyadic_points(16, 3, 3000, seed=2)
    exp = np.array([K.lowest_containing(16, p) for p in pts], dtype=np.int32)
    with O.planted_fault(O.FAULT_FIRST_VISITED):
        got = om.find_element(pts)
    assert not np.array_equal(got, exp)
    assert np.array_equal(om.find_element(pts), exp)


def test_planted_missing_dilation_fails():
    m = G.kuhn_cube(16, 0.2, seed=9)
    pts = _near_boundary_points(1500, 4)
    with O.planted_fault(O.FAULT_NO_DILATION):
        bad = O.TetMesh(m, with_ipc=False)
    good = O.TetMesh(m, with_ipc=False)
    brute = good.find_element(pts, brute=True)
    assert np.array_equal(good.find_element(pts), brute)
    assert not np.array_equal(bad.find_element(pts), brute)


def _random_tets(rng, count):
    out = []
    while len(out) < count:
        v = rng.normal(size=(4, 3)) * rng.choice([1e-3, 1.0, 1e3]) + rng.normal(size=3) * 10
        vol = np.dot(np.cross(v[1] - v[0], v[2] - v[0]), v[3] - v[0]) / 6.0
        scale = np.max(np.abs(v - v[0])) ** 3
        if abs(vol) > 1e-2 * scale:
            out.append(v)
    return out


def _lambda(p, v):
    lam = np.zeros(4)
    r = O.lib().xo_tet_contains(O._p(O._f64(p), O._d), O._p(O._f64(v.reshape(12)), O._d),
                                O._p(lam, O._d))
    return r, lam


def _identity_errors(tets, rng):
    worst = 0.0
    for v in tets:
        w = rng.dirichlet([1.0] * 4)
        p = w @ v
        ok, lam = _lambda(p, v)
        assert ok == 1
        worst = max(worst, np.max(np.abs(lam - w)))
    return worst


def test_barycentric_identities():
    rng = np.random.default_rng(5)
    tets = _random_tets(rng, 300)
    assert _identity_errors(tets, np.random.default_rng(6)) < 1e-9
    for v in tets[:100]:
        for i in range(4):
            ok, lam = _lambda(v[i], v)
            assert ok == 1
            assert np.allclose(lam, np.eye(4)[i], atol=1e-12)
        perm = rng.permutation(4)
        w = rng.dirichlet([1.0] * 4)
        _, lam = _lambda(w @ v, v)
        _, lamp = _lambda(w @ v, v[perm])
        assert np.allclose(lamp, lam[perm], atol=1e-9)
        # a weight just below zero leaves the tet
        w2 = w.copy()
        w2[rng.integers(4)] = -1e-3
        w2 /= w2.sum()
        ok, _ = _lambda(w2 @ v, v)
        assert ok == 0


def test_planted_transposed_inverse_fails():
    tets = _random_tets(np.random.default_rng(5), 50)
    w = np.array([0.1, 0.2, 0.3, 0.4])
    with O.planted_fault(O.FAULT_TET_TRANSPOSE):
        err = max(np.max(np.abs(_lambda(w @ v, v)[1] - w)) for v in tets)
    assert err > 1e-3
    assert max(np.max(np.abs(_lambda(w @ v, v)[1] - w)) for v in tets) < 1e-9
