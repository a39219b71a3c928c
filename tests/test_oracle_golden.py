"""The CPU oracle against the golden vectors produced by the reference's own
compiled sources (tests/golden/make_golden.py) -- bit-exact."""
import os

import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def prim(golden_dir):
    return np.load(os.path.join(golden_dir, "golden_primitives.npz"))


@pytest.fixture(scope="module")
def walk(golden_dir):
    return np.load(os.path.join(golden_dir, "golden_mock_walk.npz"))


def test_plucker_bits(prim):
    cin, cout = prim["plucker_in"], prim["plucker_out"]
    hits = np.zeros(len(cin))
    ts = np.zeros(len(cin))
    for i, c in enumerate(cin):
        h, t = O.plucker(c[:9], c[9:12], c[12:15], c[15], c[16], c[17] != 0,
                         int(c[18]))
        hits[i] = h
        ts[i] = t
    assert np.array_equal(hits, cout[:, 0])
    assert np.array_equal(bits(ts), bits(cout[:, 1]))


def test_edge_bits(prim):
    cin, cout = prim["edge_in"], prim["edge_out"]
    got = np.array([O.edge_test(c[0:3], c[3:6], c[6:9], c[9:12]) for c in cin])
    assert np.array_equal(bits(got), bits(cout))


def test_normal_bits(prim):
    cin, cout = prim["normal_in"], prim["normal_out"]
    got = np.stack([O.face_normal(c) for c in cin])
    assert np.array_equal(bits(got), bits(cout))


def test_closest_bits(prim):
    cin, cout = prim["closest_in"], prim["closest_out"]
    reg = np.array([O.region(c[:9], c[9:]) for c in cin])
    pts = np.stack([O.closest_point(c[:9], c[9:]) for c in cin])
    assert np.array_equal(reg, cout[:, 0].astype(int))
    assert np.array_equal(bits(pts), bits(cout[:, 1:]))


def test_mock_next_element(walk):
    tm = O.TetMesh(G.mock_tetmesh())
    cin, cout = walk["next_in"], walk["next_out"]
    for c, ref in zip(cin, cout):
        nx, d = O.TetMesh.next_element(tm, int(c[0]), c[1:4], c[4:7])
        if ref[1] == np.finfo(np.float64).max:
            # reference: adjacent_element(elem, -1) is undefined; the build
            # returns ID_NONE with INFTY
            assert d == ref[1]
            continue
        assert nx == int(ref[0])
        assert bits([d])[0] == bits([ref[1]])[0]


def test_mock_walk_sequences(walk):
    tm = O.TetMesh(G.mock_tetmesh())
    cin, cout = walk["walk_in"], walk["walk_out"]
    for c, ref in zip(cin, cout):
        e, ln = tm.walk_elements(int(c[0]), c[1:4], c[4:7], c[7])
        n = int(ref[0])
        assert len(e) == n
        pairs = ref[1:1 + 2 * n].reshape(-1, 2)
        assert np.array_equal(e, pairs[:, 0].astype(np.int32))
        assert np.array_equal(bits(ln), bits(pairs[:, 1]))


def test_mock_face_normals(walk):
    m = G.mesh_mock_model()
    v9 = m.verts[m.tris].reshape(-1, 9)
    got = np.stack([O.face_normal(v) for v in v9])
    assert np.array_equal(bits(got), bits(walk["face_normal"]))


def _volume_tree(mesh):
    bm = mesh.boundary_model()
    return O.Tree.from_model(bm, 1)


def test_mock_segments_volume_golden(golden_dir):
    """XDG::segments(volume, start, end) (src/xdg.cpp:200-227) on MeshMock:
    the oracle's sequences equal walk_elements(elem, start, end) of the
    compiled reference, for starts inside the mesh and for starts outside it
    (entered through the first ENTERING boundary hit)."""
    g = np.load(os.path.join(golden_dir, "golden_mock_walk_se.npz"))
    tracks = np.load(os.path.join(golden_dir, "golden_mock_tracks.npz"))["tracks"]
    sin, sout = g["walk_se_in"], g["walk_se_out"]
    nin = tracks.shape[0]
    p0 = np.concatenate([tracks[:, :3], g["outside_start"]])
    p1 = np.concatenate([tracks[:, 3:], sin[nin:, 4:7]])
    mesh = G.mock_tetmesh()
    tm = O.TetMesh(mesh)
    cap = np.full(p0.shape[0], 64, np.int64)
    cnt, hsh, flg, (off, se, sl) = tm.segments_volume(_volume_tree(mesh), p0, p1,
                                                      sequences=cap)
    assert not flg.any()
    assert np.array_equal(cnt, sout[:, 0].astype(np.int32))
    for i in range(p0.shape[0]):
        n = cnt[i]
        pairs = sout[i, 1:1 + 2 * n].reshape(-1, 2)
        assert np.array_equal(se[off[i]:off[i] + n], pairs[:, 0].astype(np.int32))
        assert np.array_equal(bits(sl[off[i]:off[i] + n]), bits(pairs[:, 1]))
    # the per-segment hash is the fold of the sequence
    for i in range(0, p0.shape[0], 97):
        h = np.uint64(0x6a09e667f3bcc909)
        for e, ln in zip(se[off[i]:off[i] + cnt[i]], sl[off[i]:off[i] + cnt[i]]):
            h = np.uint64(O.mix64(int(h) ^ int(np.uint32(e))))
            h = np.uint64(O.mix64(int(h) ^ int(bits([ln])[0])))
        assert h == hsh[i]


def _slab(o, u, lo=0.0, hi=1.0):
    """Entry and exit distances of the ray o + t u through [lo, hi]^3."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - o) / u
        t1 = (hi - o) / u
    tn = np.minimum(t0, t1).max(1)
    tf = np.maximum(t0, t1).min(1)
    return tn, tf


def test_segments_volume_entry_closed_form():
    """Starts outside the unit cube (whose boundary the Kuhn mesh keeps
    unjittered): the walk begins at the slab entry point, so the lengths sum
    to |end - entry| when the end is inside, to the slab chord when the end
    lies beyond the cube, and rays that miss the cube give no entries."""
    mesh = G.kuhn_cube(8, 0.2)
    tm = O.TetMesh(mesh)
    vol = _volume_tree(mesh)
    rng = np.random.default_rng(7)
    n = 600
    o = rng.uniform(-1.0, 2.0, (n, 3))
    o = o[~((o > 0) & (o < 1)).all(1)]
    e_in = rng.uniform(0.02, 0.98, (o.shape[0], 3))
    cnt, hsh, flg, tal, (off, se, sl) = tm.segments_volume(
        vol, o, e_in, tally=True, sequences=np.full(o.shape[0], 256))
    assert not flg.any()
    u = e_in - o
    u = u / np.sqrt((u * u).sum(1))[:, None]
    tn, tf = _slab(o, u)
    entry = o + u * tn[:, None]
    want = np.sqrt(((e_in - entry) ** 2).sum(1))
    got = np.array([sl[off[i]:off[i] + cnt[i]].sum() for i in range(o.shape[0])])
    assert (cnt > 0).all()
    assert np.abs(got - want).max() < 1e-12
    assert abs(tal.sum() / 2.0 ** 40 - want.sum()) < 1e-6
    # ends beyond the cube along the same lines: the walk leaves at the exit
    e_out = o + u * (tf + 0.5)[:, None]
    cnt2, _, flg2, (off2, se2, sl2) = tm.segments_volume(
        vol, o, e_out, sequences=np.full(o.shape[0], 256))
    assert not flg2.any()
    got2 = np.array([sl2[off2[i]:off2[i] + cnt2[i]].sum() for i in range(o.shape[0])])
    assert np.abs(got2 - (tf - tn)).max() < 1e-12
    # rays pointing away from the cube never enter it
    away = o + (o - 0.5) * 3.0
    cnt3, _, flg3 = tm.segments_volume(vol, o, away)
    tn3, tf3 = _slab(o, (away - o) / np.sqrt(((away - o) ** 2).sum(1))[:, None])
    miss = ~((tn3 <= tf3) & (tf3 >= 0))
    assert miss.sum() > 0
    assert (cnt3[miss] == 0).all()
    # without a surface tree an outside start gives nothing
    cnt4, _, _ = tm.segments_volume(None, o, e_in)
    assert (cnt4 == 0).all()


def test_segments_volume_matches_global_inside():
    """For segments inside a convex mesh the volume form and the global form
    (src/xdg.cpp:142-198) walk the same elements: the volume form's sequence
    is the global one, or its prefix where the global form re-enters to cover
    a rounding residual."""
    mesh = G.kuhn_cube(8, 0.2)
    tm = O.TetMesh(mesh)
    vol = _volume_tree(mesh)
    rng = np.random.default_rng(11)
    n = 400
    p0 = rng.uniform(0.01, 0.99, (n, 3))
    p1 = rng.uniform(0.01, 0.99, (n, 3))
    cap = np.full(n, 256)
    c1, h1, f1, (o1, e1, l1) = tm.segments_volume(vol, p0, p1, sequences=cap)
    c2, h2, (o2, e2, l2) = tm.segments(p0, p1, sequences=cap)
    assert not f1.any()
    same = h1 == h2
    assert same.mean() > 0.95
    for i in np.nonzero(~same)[0]:
        a = e1[o1[i]:o1[i] + c1[i]]
        b = e2[o2[i]:o2[i] + c2[i]]
        assert c1[i] < c2[i] and np.array_equal(a, b[:c1[i]])


def test_watertight_sphere_golden(golden_dir):
    """Rays from the centre of the subdivision-5 icosphere through vertices
    and edge midpoints.  The reference's own plucker_ray_tri_intersect
    leaks two of the 40 962 (no sharing triangle accepts: the shared edge's
    coordinate is zeroed by the 20-epsilon tolerance and the barycentric
    guard u + v > 1 then fails by rounding in both triangles).  The oracle
    must give the same per-triangle bits, and its ray_fire over the whole
    sphere must leak exactly those rays and hit every other one at the
    nearest accepting triangle's distance."""
    g = np.load(os.path.join(golden_dir, "golden_watertight.npz"))
    cin, cout = g["plucker_in"], g["plucker_out"]
    for c, want in zip(cin, cout):
        h, t = O.plucker(c[:9], c[9:12], c[12:15], c[15], c[16], c[17] != 0, int(c[18]))
        assert (float(h), bits(t)) == (want[0], bits(want[1]))
    leak = g["leak"]
    assert 0 < leak.sum() < leak.size
    m = G.single_sphere(5, 10.0)
    tree = O.Tree.from_model(m, 1)
    d = g["dir"]
    t, s, _ = tree.ray_fire(np.zeros_like(d), d)
    assert np.array_equal(s < 0, leak)
    case_ray = g["case_ray"]
    for j, r in enumerate(g["ray"]):
        acc = (case_ray == r) & (cout[:, 0] != 0.0)
        if leak[j]:
            assert not acc.any()
            continue
        assert s[j] == 1
        assert t[j] == cout[acc, 1].min()
