"""HIP kernels vs the CPU oracle and the reference's golden vectors.

Bit-exact bar: hit distances compared as uint64 bit patterns, ids and flags
exactly.  Every case calls through the C-ABI (libxdg_hip.so).
"""
import math
import os

import numpy as np
import pytest

import pyoracle as O
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


# ---------------------------------------------------------------------------
# primitives vs the reference's compiled functions
# ---------------------------------------------------------------------------

def test_primitives_golden(ctx, golden_dir):
    g = np.load(os.path.join(golden_dir, "golden_primitives.npz"))
    out = ctx.eval_primitive(0, g["plucker_in"], 2)
    ref = g["plucker_out"]
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert np.array_equal(bits(out[:, 1]), bits(ref[:, 1]))
    out = ctx.eval_primitive(1, g["normal_in"], 3)
    assert np.array_equal(bits(out), bits(g["normal_out"]))
    out = ctx.eval_primitive(2, g["closest_in"], 3)
    assert np.array_equal(bits(out), bits(g["closest_out"][:, 1:]))


def test_sqrt_div_correctly_rounded(ctx):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0, 1, 200000),
                        np.exp(rng.uniform(-700, 700, 200000)),
                        rng.integers(1, 1 << 40, 100000).astype(np.float64),
                        [0.0, 1.0, 2.0, 4.0, 1e-310, 5e-324, INFTY]])
    out = ctx.eval_primitive(4, x.reshape(-1, 1), 1)[:, 0]
    assert np.array_equal(bits(out), bits(np.sqrt(x)))
    a = np.exp(rng.uniform(-300, 300, 300000)) * rng.choice([-1, 1], 300000)
    b = np.exp(rng.uniform(-300, 300, 300000)) * rng.choice([-1, 1], 300000)
    out = ctx.eval_primitive(5, np.stack([a, b], 1), 1)[:, 0]
    assert np.array_equal(bits(out), bits(a / b))


def test_tet_containment_reference_cases(ctx):
    # tests/test_tet_containment.cpp:14-95
    t1 = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]
    t2 = [1, 1, 1, 2, 1, 1, 1, 2, 1, 1, 1, 2]
    cases = [([0.1, 0.1, 0.1], t1, 1), ([2, 2, 2], t1, 0), ([0, 0, 0.5], t1, 1),
             ([-0.5, -0.5, 0], t1, 0), ([0.3, 0.3, 0.1], t1, 1),
             ([0.3, 0.1, 0.3], t1, 1), ([0.1, 0.3, 0.3], t1, 1),
             ([0.3, 0.3, 0.3], t1, 1), ([0.3, 0.3, -0.1], t1, 0),
             ([0.3, -0.1, 0.3], t1, 0), ([-0.1, 0.3, 0.3], t1, 0),
             ([0.4, 0.4, 0.4], t1, 0), ([1.2, 1.2, 1.2], t2, 1),
             ([0, 0, 0], t2, 0), ([1, 1, 1.5], t2, 1), ([1.3, 1.3, 1.1], t2, 1),
             ([1.3, 1.3, 0.9], t2, 0), ([0.1, 0.1, 0.1], t2, 0),
             ([1.2, 1.2, 1.2], t1, 0)]
    inp = np.array([p + t for p, t, _ in cases], dtype=np.float64)
    out = ctx.eval_primitive(3, inp, 1)[:, 0]
    assert list(out.astype(int)) == [e for _, _, e in cases]
    # and the oracle agrees case by case
    for (p, t, e) in cases:
        assert O.tet_contains(p, t)[0] == bool(e)


# ---------------------------------------------------------------------------
# MeshMock known answers (tests/test_ray_fire.cpp, test_point_in_volume.cpp,
# test_occluded.cpp, test_closest.cpp, test_find_element.cpp)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", [0, 1])
def test_mock_ray_fire(ctx, torch_mod, prec):
    torch = torch_mod
    m = G.mesh_mock_model()
    tree = build_tree(ctx, m, 0)
    o = np.zeros((6, 3))
    d = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                  [0, 0, -1]], float)
    t, s, f = ctx.ray_fire(tree, dev(torch, o), dev(torch, d), precision=prec)
    assert np.allclose(host(t), [5, 2, 6, 3, 7, 4], atol=1e-6)
    t, s, f = ctx.ray_fire(tree, dev(torch, [[-10.0, 0, 0], [10.0, 0, 0]]),
                           dev(torch, [[1.0, 0, 0], [-1.0, 0, 0]]),
                           precision=prec)
    assert np.allclose(host(t), [15, 12], atol=1e-6)
    t, s, f = ctx.ray_fire(tree, dev(torch, [[-10.0, 0, 0], [10.0, 0, 0]]),
                           dev(torch, [[1.0, 0, 0], [-1.0, 0, 0]]),
                           orientation=1, precision=prec)
    assert np.allclose(host(t), [8, 5], atol=1e-6)
    t, s, f = ctx.ray_fire(tree, dev(torch, np.zeros((2, 3))),
                           dev(torch, [[1.0, 0, 0], [1.0, 0, 0]]),
                           tmax=dev(torch, [4.5, 5.1]), precision=prec)
    assert host(s)[0] == -1 and host(s)[1] != -1
    assert abs(host(t)[1] - 5.0) < 1e-6
    t, s, f = ctx.ray_fire(tree, dev(torch, np.zeros((1, 3))),
                           dev(torch, [[1.0, 0, 0]]), precision=prec)
    ex = dev(torch, host(f).reshape(1, 1).astype(np.int32))
    t2, s2, f2 = ctx.ray_fire(tree, dev(torch, np.zeros((1, 3))),
                              dev(torch, [[1.0, 0, 0]]), excl=ex,
                              precision=prec)
    assert host(s2)[0] == -1


def test_mock_piv_occluded_closest(ctx, torch_mod):
    torch = torch_mod
    m = G.mesh_mock_model()
    tree = build_tree(ctx, m, 0)
    pts = [[0, 0, 0], [0, 0, 1000], [4 - 1e-6, 0, 0], [5.001, 0, 0]]
    r = host(ctx.point_in_volume(tree, dev(torch, np.array(pts, float))))
    assert list(r) == [1, 0, 1, 0]
    r = host(ctx.point_in_volume(tree, dev(torch, [[5.0, 0, 0], [5.1, 0, 0], [5.1, 0, 0]]),
                                 dirs=dev(torch, [[1.0, 0, 0], [1.0, 0, 0], [-1.0, 0, 0]])))
    assert list(r) == [1, 0, 0]
    hit, dist = ctx.occluded(tree, dev(torch, [[-100.0, 0, 0]] * 2),
                             dev(torch, [[1.0, 0, 0], [-1.0, 0, 0]]))
    assert list(host(hit)) == [1, 0]
    assert host(dist)[0] == -INFTY and host(dist)[1] == INFTY
    d, f = ctx.closest(tree, dev(torch, [[0.0, 0, 0], [4.0, 0, 0], [10.0, 0, 0]]))
    assert np.allclose(host(d), [2, 1, 5], atol=1e-6)
    # test_closest.cpp:46-56: 1000 points along x through the box centre
    rng = np.random.default_rng(3)
    p = np.tile([1.5, 1.5, 1.5], (1000, 1))
    p[:, 0] = rng.uniform(-10, 10, 1000)
    d, f = ctx.closest(tree, dev(torch, p))
    exp = np.where(p[:, 0] < 1.5, np.abs(p[:, 0] + 2.0), np.abs(p[:, 0] - 5.0))
    assert np.allclose(host(d), exp, atol=1e-6)


def test_mock_elements(ctx, torch_mod, golden_dir):
    torch = torch_mod
    tm = G.mock_tetmesh()
    et = ctx.build_element_tree(tm.verts, tm.tets, tm.adj, tm.face_order)
    e = host(ctx.find_element(et, dev(torch, [[0.0, 0, 0], [10.0, 10, 10]])))
    assert e[0] != -1 and e[1] == -1
    g = np.load(os.path.join(golden_dir, "golden_mock_walk.npz"))
    cin, cout = g["next_in"], g["next_out"]
    nx, d = ctx.next_element(et, dev(torch, cin[:, 0].astype(np.int32)),
                             dev(torch, cin[:, 1:4]), dev(torch, cin[:, 4:7]))
    nx, d = host(nx), host(d)
    ok = cout[:, 1] != INFTY
    assert np.array_equal(nx[ok], cout[ok, 0].astype(np.int32))
    assert np.array_equal(bits(d), bits(cout[:, 1]))


# ---------------------------------------------------------------------------
# seeded synthetic workloads vs the oracle
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def sphere5(ctx):
    m = G.single_sphere(5, 10.0)
    return m, build_tree(ctx, m, 1), O.Tree.from_model(m, 1)


# traversal variants: (wide, persistent, ray order mode, speculative);
# batches of 65536 rays and more are reordered when the mode is non-zero
VARIANTS = [(True, True, 0, 0), (True, False, 0, 0), (False, False, 0, 0),
            (True, True, 1, 0), (True, False, 2, 0), (True, False, 0, 1),
            (True, False, 1, 1)]


@pytest.mark.parametrize("wide,persist,order,spec", VARIANTS)
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("orientation", [0, 1, -1])
def test_sphere_ray_fire_bitexact(ctx, torch_mod, sphere5, prec, orientation, wide, persist, order, spec):
    torch = torch_mod
    m, tree, ot = sphere5
    ctx.set_traversal(wide=wide, defer=True, persistent=persist)
    ctx.set_ray_order(order)
    ctx.set_option("speculative", spec)
    n = 65536
    org = P.ball_points(1, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(1, P.STREAM_DIR, n)
    if orientation == 1:  # entering hits need origins outside
        org = org * 2.5
    try:
        t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs),
                               orientation=orientation, precision=prec)
    finally:
        ctx.set_traversal()
        ctx.set_ray_order(0)
        ctx.set_option("speculative", 1)  # the default
    rt, rs, rf = ot.ray_fire(org, dirs, orientation=orientation)
    assert np.array_equal(host(s), rs)
    assert np.array_equal(host(f), rf)
    assert np.array_equal(bits(host(t)), bits(rt))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("wide", [True, False])
def test_watertight_vertex_and_edge_rays(ctx, torch_mod, sphere5, golden_dir, prec, wide):
    """Rays through every vertex and every edge midpoint of the icosphere,
    from its centre and three other interior points: hits, misses and
    distance bits equal to the oracle's.  From the centre, the two rays the
    reference's own Plücker test leaks (golden_watertight.npz) miss."""
    torch = torch_mod
    m, tree, ot = sphere5
    v, t3 = m.verts, m.tris
    e = np.concatenate([t3[:, [0, 1]], t3[:, [1, 2]], t3[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    targets = np.concatenate([v, 0.5 * (v[e[:, 0]] + v[e[:, 1]])])
    origins = np.concatenate([np.zeros((1, 3)), P.ball_points(7, P.STREAM_POS, 3, 4.0)])
    org = np.repeat(origins, targets.shape[0], axis=0)
    d = np.tile(targets, (origins.shape[0], 1)) - org
    d = d / np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])[:, None]
    g = np.load(os.path.join(golden_dir, "golden_watertight.npz"))
    assert np.array_equal(bits(d[g["ray"]]), bits(g["dir"]))
    ctx.set_traversal(wide=wide, defer=True)
    try:
        t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, d), precision=prec)
    finally:
        ctx.set_traversal()
    rt, rs, rf = ot.ray_fire(org, d)
    assert np.array_equal(host(s), rs)
    assert np.array_equal(host(f), rf)
    assert np.array_equal(bits(host(t)), bits(rt))
    leaks = g["ray"][g["leak"]]
    assert leaks.size and (host(s)[leaks] == -1).all()
    assert (rs >= 0).mean() > 0.999


@pytest.mark.parametrize("refill", [1, 17, 64])
def test_persistent_refill_thresholds(ctx, torch_mod, sphere5, refill):
    """Persistent ray_fire with lanes refilled one at a time, in groups, or
    only once the whole wave is idle: identical to the oracle."""
    torch = torch_mod
    m, tree, ot = sphere5
    n = 70000
    org = P.ball_points(31, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(31, P.STREAM_DIR, n)
    rt, rs, rf = ot.ray_fire(org, dirs)
    keep = ctx.get_option("refill")
    ctx.set_traversal(persistent=True)
    ctx.set_option("refill", refill)
    try:
        for prec in (0, 1):
            t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), precision=prec)
            assert np.array_equal(host(s), rs)
            assert np.array_equal(host(f), rf)
            assert np.array_equal(bits(host(t)), bits(rt))
    finally:
        ctx.set_traversal()
        ctx.set_option("refill", keep)


@pytest.mark.parametrize("refill", [1, 44])
def test_persistent_ray_fire_many_per_lane(ctx, torch_mod, sphere5, refill):
    """2^21 rays, several for each lane of the resident grid, so lanes claim
    new rays while others of their wave are still tracing: persistent
    ray_fire in both precisions equals the per-ray kernel bit for bit, and
    its first 20000 rays equal the oracle."""
    torch = torch_mod
    m, tree, ot = sphere5
    n, k = 1 << 21, 20000
    org = P.ball_points(37, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(37, P.STREAM_DIR, n)
    o_d, d_d = dev(torch, org), dev(torch, dirs)
    rt, rs, rf = ot.ray_fire(org[:k], dirs[:k])
    keep = ctx.get_option("refill")
    try:
        for prec in (0, 1):
            ctx.set_traversal(persistent=False)
            ref = [host(x) for x in ctx.ray_fire(tree, o_d, d_d, precision=prec)]
            ctx.set_traversal(persistent=True)
            ctx.set_option("refill", refill)
            got = [host(x) for x in ctx.ray_fire(tree, o_d, d_d, precision=prec)]
            assert np.array_equal(bits(got[0]), bits(ref[0]))
            assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
            assert np.array_equal(bits(got[0][:k]), bits(rt))
            assert np.array_equal(got[1][:k], rs) and np.array_equal(got[2][:k], rf)
    finally:
        ctx.set_traversal()
        ctx.set_option("refill", keep)


@pytest.mark.parametrize("fv_refill", [0, 1])
def test_persistent_find_volume_many_per_lane(ctx, torch_mod, fv_refill):
    """2^21 points, several for each lane of the resident grid: persistent
    find_volume (lanes moving on to the next volume and claiming new points
    while others trace) equals the per-point kernel in both precisions, and
    its first 20000 points equal the oracle."""
    torch = torch_mod
    m = G.nested_spheres(2)
    vols = list(m.volumes)
    trees = [build_tree(ctx, m, v) for v in vols]
    n, k = 1 << 21, 20000
    pts = P.box_points(43, P.STREAM_POS, n, [-19] * 3, [19] * 3)
    dirs = P.iso_dirs(43, P.STREAM_DIR, n)
    ref_k = O.find_volume([O.Tree.from_model(m, v) for v in vols], vols, m.ipc, pts[:k], dirs[:k])
    p_d, d_d = dev(torch, pts), dev(torch, dirs)
    try:
        ctx.set_option("fv_refill", fv_refill)
        for prec in (0, 1):
            ctx.set_traversal(persistent=False)
            ref = host(ctx.find_volume(trees, vols, m.ipc, p_d, d_d, precision=prec))
            ctx.set_traversal(persistent=True)
            got = host(ctx.find_volume(trees, vols, m.ipc, p_d, d_d, precision=prec))
            assert np.array_equal(got, ref)
            assert np.array_equal(got[:k], ref_k)
    finally:
        ctx.set_traversal()
        ctx.set_option("fv_refill", 0)


def test_tree_build_modes(ctx, torch_mod):
    """Trees built by the binned-SAH hierarchy and by the Karras radix tree
    answer ray_fire (both precisions, both widths) and point_in_volume
    identically, and both equal the oracle; both trees stay within the 64
    levels the traversal stacks allow."""
    torch = torch_mod
    m = G.nested_spheres(3)
    vols = [v for v in m.volumes if v != m.ipc]
    n = 30000
    org = P.box_points(53, P.STREAM_POS, n, [-16] * 3, [16] * 3)
    dirs = P.iso_dirs(53, P.STREAM_DIR, n)
    o_d, d_d = dev(torch, org), dev(torch, dirs)
    keep = ctx.get_option("build")
    try:
        for v in vols:
            ot = O.Tree.from_model(m, v)
            rt, rs, rf = ot.ray_fire(org, dirs)
            rin = ot.point_in_volume(org, dirs)
            got = {}
            for mode in (1, 0):
                ctx.set_option("build", mode)
                tree = build_tree(ctx, m, v)
                got[mode] = ctx.tree_info(tree)["depth"]
                for wide in (True, False):
                    ctx.set_traversal(wide=wide)
                    for prec in (0, 1):
                        t, s_, f = ctx.ray_fire(tree, o_d, d_d, precision=prec)
                        assert np.array_equal(bits(host(t)), bits(rt))
                        assert np.array_equal(host(s_), rs) and np.array_equal(host(f), rf)
                        pin = host(ctx.point_in_volume(tree, o_d, d_d, precision=prec))
                        assert np.array_equal(pin.astype(bool), rin.astype(bool))
                ctx.set_traversal()
                ctx.release_surface_tree(tree)
            assert got[1] <= 64 and got[0] <= 64
    finally:
        ctx.set_option("build", keep)
        ctx.set_traversal()


@pytest.mark.parametrize("mode", [1, 0])
def test_tree_build_degenerate_and_tiny(ctx, torch_mod, mode):
    """Trees whose centroids no SAH bin boundary separates (600 copies of
    one triangle beside a sphere; their order then splits at the median) and
    trees of 1, 2 and 3 triangles, in both build modes: ray_fire (every
    orientation, both precisions) equals the oracle's brute force, duplicate
    faces resolved to the smallest face id."""
    torch = torch_mod
    m = G.single_sphere(2, 10.0)
    v9, f, s_, fl = m.tree_arrays(1)
    dup = 600
    v9d = np.concatenate([v9, np.repeat(v9[:1], dup, axis=0)])
    fd = np.concatenate([f, np.arange(dup, dtype=np.int32) + int(f.max()) + 1])
    cases = [(v9d, fd, np.concatenate([s_, np.repeat(s_[:1], dup)]), np.concatenate([fl, np.repeat(fl[:1], dup)]))]
    for k in (1, 2, 3):
        cases.append((v9[:k], f[:k], s_[:k], fl[:k]))
    n = 4000
    org = P.ball_points(61, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(61, P.STREAM_DIR, n)
    # half of the rays aimed at the duplicated triangle's centroid
    c = v9[0].reshape(3, 3).mean(0)
    aim = c - org[: n // 2]
    dirs[: n // 2] = aim / np.linalg.norm(aim, axis=1, keepdims=True)
    o_d, d_d = dev(torch, org), dev(torch, dirs)
    keep = ctx.get_option("build")
    try:
        ctx.set_option("build", mode)
        for cv9, cf, cs, cfl in cases:
            nt = cv9.shape[0]
            tree = ctx.build_surface_tree(cv9.reshape(-1, 3), np.arange(3 * nt, dtype=np.int32).reshape(-1, 3),
                                          cf, cs, cfl, m.bump(1))
            ot = O.Tree(cv9, cf, cs, cfl, m.bump(1))
            for orient in (0, 1, -1):
                rt, rs, rf = ot.ray_fire(org, dirs, orientation=orient, brute=True)
                for prec in (0, 1):
                    t, s2, f2 = ctx.ray_fire(tree, o_d, d_d, orientation=orient, precision=prec)
                    assert np.array_equal(bits(host(t)), bits(rt))
                    assert np.array_equal(host(s2), rs) and np.array_equal(host(f2), rf)
            ctx.release_surface_tree(tree)
    finally:
        ctx.set_option("build", keep)


@pytest.mark.parametrize("wide", [True, False])
def test_nested_piv_occluded_variants(ctx, torch_mod, wide):
    """Binary and 4-wide traversal, plain and mixed, give identical
    point_in_volume / occluded answers (nested spheres, subdivision 3)."""
    torch = torch_mod
    m = G.nested_spheres(3)
    ctx.set_traversal(wide=wide, defer=True)
    try:
        n = 20000
        pts = P.box_points(21, P.STREAM_POS, n, [-16] * 3, [16] * 3)
        dirs = P.iso_dirs(21, P.STREAM_DIR, n)
        for v in m.volumes:
            tree = build_tree(ctx, m, v)
            ot = O.Tree.from_model(m, v)
            ref_in = ot.point_in_volume(pts, dirs)
            ref_hit, _ = ot.occluded(pts, dirs)
            for prec in (0, 1):
                got = host(ctx.point_in_volume(tree, dev(torch, pts), dev(torch, dirs), precision=prec))
                assert np.array_equal(got.astype(bool), ref_in)
                hit, _ = ctx.occluded(tree, dev(torch, pts), dev(torch, dirs), precision=prec)
                assert np.array_equal(host(hit).astype(bool), ref_hit)
    finally:
        ctx.set_traversal(wide=True, defer=True)


@pytest.mark.parametrize("persist", [True, False])
def test_mixed_far_origins_use_fp64_pass(ctx, torch_mod, sphere5, persist):
    """Origins beyond the FP32-rounding limit of the tree (|o| >> o_limit)
    go through the deferred FP64 pass of the mixed kernel; results must be
    identical to the oracle, mixed with near origins in one batch."""
    torch = torch_mod
    m, tree, ot = sphere5
    ctx.set_traversal(persistent=persist)
    lim = ctx.tree_info(tree)["o_limit"]
    assert 100.0 < lim < 1e7
    n = 8192
    dirs = P.iso_dirs(13, P.STREAM_DIR, n)
    org = P.ball_points(13, P.STREAM_POS, n, 9.0)
    far = np.arange(n) % 3 == 0
    org[far] = -dirs[far] * (4.0 * lim) + org[far] * 0.01
    for orientation in (-1, 1):
        rt, rs, rf = ot.ray_fire(org, dirs, orientation=orientation)
        assert (rs[far] >= 0).mean() > 0.9
        for prec in (0, 1):
            t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs),
                                   orientation=orientation, precision=prec)
            assert np.array_equal(host(s), rs)
            assert np.array_equal(host(f), rf)
            assert np.array_equal(bits(host(t)), bits(rt))
        st = ctx.ray_fire_stats(tree, dev(torch, org), dev(torch, dirs),
                                orientation=orientation, precision=1)
        assert st["fallbacks"] == int(far.sum())
    ctx.set_traversal()


def test_sphere_limits_and_exclusions(ctx, torch_mod, sphere5):
    torch = torch_mod
    m, tree, ot = sphere5
    n = 20000
    org = P.ball_points(11, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(11, P.STREAM_DIR, n)
    rt, rs, rf = ot.ray_fire(org, dirs)
    # dist_limit exactly at / just below the hit distance
    tm = rt.copy()
    tm[::2] = np.nextafter(rt[::2], 0)
    excl = np.full((n, 2), -1, np.int32)
    excl[::3, 0] = rf[::3]
    et, es, ef = ot.ray_fire(org, dirs, tmax=tm, excl=excl)
    for prec, persist in [(0, True), (1, True), (0, False), (1, False)]:
        ctx.set_traversal(persistent=persist)
        try:
            t, s, f = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs),
                                   tmax=dev(torch, tm), excl=dev(torch, excl),
                                   precision=prec)
        finally:
            ctx.set_traversal()
        assert np.array_equal(host(s), es)
        assert np.array_equal(host(f), ef)
        assert np.array_equal(bits(host(t)), bits(et))


def test_surface_id_range(ctx, torch_mod):
    """Surface ids up to 2^27 - 1 come back intact beside the edge-order
    bits packed into the triangle record; larger ids are rejected."""
    from xdg_amd import capi
    torch = torch_mod
    m = G.single_sphere(3, 10.0)
    v9, f, s, fl = m.tree_arrays(1)
    nt = v9.shape[0]
    verts = v9.reshape(-1, 3)
    tris = np.arange(3 * nt, dtype=np.int32).reshape(-1, 3)
    big = np.full(nt, (1 << 27) - 1, np.int32)
    tree = ctx.build_surface_tree(verts, tris, f, big, fl, m.bump(1))
    org = P.ball_points(12, P.STREAM_POS, 4096, 9.0)
    dirs = P.iso_dirs(12, P.STREAM_DIR, 4096)
    ot = O.Tree.from_model(m, 1)
    rt, rs, rf = ot.ray_fire(org, dirs)
    for prec in (0, 1):
        t, s_, f_ = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), precision=prec)
        hit = rs >= 0
        assert np.array_equal(host(s_)[hit], np.full(hit.sum(), (1 << 27) - 1, np.int32))
        assert np.array_equal(host(f_), rf)
        assert np.array_equal(bits(host(t)), bits(rt))
    with pytest.raises(capi.XdgHipError):
        ctx.build_surface_tree(verts, tris, f, np.full(nt, 1 << 27, np.int32), fl, m.bump(1))


def test_sphere_occluded_closest_piv(ctx, torch_mod, sphere5):
    torch = torch_mod
    m, tree, ot = sphere5
    n = 20000
    org = P.box_points(12, P.STREAM_POS, n, [-15] * 3, [15] * 3)
    dirs = P.iso_dirs(12, P.STREAM_DIR, n)
    for prec in (0, 1):
        hit, dist = ctx.occluded(tree, dev(torch, org), dev(torch, dirs),
                                 precision=prec)
        rh, rd = ot.occluded(org, dirs)
        assert np.array_equal(host(hit).astype(bool), rh)
        assert np.array_equal(bits(host(dist)), bits(rd))
        inside = ctx.point_in_volume(tree, dev(torch, org), dev(torch, dirs),
                                     precision=prec)
        assert np.array_equal(host(inside).astype(bool),
                              ot.point_in_volume(org, dirs))
    d, f = ctx.closest(tree, dev(torch, org[:4096]))
    rd, rf = ot.closest(org[:4096])
    assert np.array_equal(host(f), rf)
    assert np.array_equal(bits(host(d)), bits(rd))


def test_nested_find_volume_and_fire(ctx, torch_mod):
    torch = torch_mod
    m = G.nested_spheres(4)
    vols = [v for v in m.volumes]
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    n = 100000
    pts = P.box_points(2, P.STREAM_POS, n, [-16] * 3, [16] * 3)
    dirs = P.iso_dirs(2, P.STREAM_DIR, n)
    ref_vol = O.find_volume(otrees, vols, m.ipc, pts, dirs)
    # analytic classification away from the surfaces
    r = np.sqrt((pts ** 2).sum(1))
    far = (np.abs(r - 5) > 0.1) & (np.abs(r - 10) > 0.1) & (np.abs(r - 15) > 0.2)
    analytic = np.select([r < 5, r < 10, r < 15], [1, 2, 3], 4)
    assert np.array_equal(ref_vol[far], analytic[far])
    for prec, order, spec in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 2, 0),
                              (0, 0, 1), (1, 1, 1)]:
        vol = host(ctx.find_volume(trees, vols, m.ipc, dev(torch, pts),
                                   dev(torch, dirs), precision=prec))
        assert np.array_equal(vol, ref_vol)
        tpr = np.array([trees[vols.index(v)] for v in vol], np.int32)
        ctx.set_ray_order(order)
        ctx.set_option("speculative", spec)
        try:
            t, s, f = ctx.ray_fire(0, dev(torch, pts), dev(torch, dirs),
                                   tree_per_ray=dev(torch, tpr), precision=prec)
        finally:
            ctx.set_ray_order(0)
            ctx.set_option("speculative", 1)  # the default
        t, s, f = host(t), host(s), host(f)
        for v in vols:
            sel = vol == v
            if not sel.any():
                continue
            rt, rs, rf = otrees[vols.index(v)].ray_fire(pts[sel], dirs[sel])
            assert np.array_equal(s[sel], rs)
            assert np.array_equal(f[sel], rf)
            assert np.array_equal(bits(t[sel]), bits(rt))


@pytest.mark.parametrize("persist,fv_refill", [(True, 0), (False, 0), (True, 1), (True, 23), (True, 64)])
def test_find_volume_far_points_and_variants(ctx, torch_mod, persist, fv_refill):
    """Points beyond the FP32 limit of the trees (deferred to the FP64 pass
    by the persistent mixed kernel) mixed with near ones; persistent (at the
    default and at other refill thresholds) and per-point kernels, both
    precisions, two volume orders, equal the oracle."""
    torch = torch_mod
    m = G.nested_spheres(2)
    vols = list(m.volumes)
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    lim = min(ctx.tree_info(t)["o_limit"] for t in trees)
    n = 30000
    pts = P.box_points(41, P.STREAM_POS, n, [-19] * 3, [19] * 3)
    dirs = P.iso_dirs(41, P.STREAM_DIR, n)
    far = np.arange(n) % 5 == 0
    pts[far] = pts[far] / 19.0 * (3.0 * lim)
    ref = O.find_volume(otrees, vols, m.ipc, pts, dirs)
    ctx.set_traversal(persistent=persist)
    ctx.set_option("fv_refill", fv_refill)
    try:
        for prec in (0, 1):
            got = host(ctx.find_volume(trees, vols, m.ipc, dev(torch, pts), dev(torch, dirs), precision=prec))
            assert np.array_equal(got, ref)
            # a volume order with the implicit complement first and the
            # outermost volume before the inner ones
            order = [vols[-1], vols[-2]] + vols[:-2]
            tr = [trees[vols.index(v)] for v in order]
            ref2 = O.find_volume([otrees[vols.index(v)] for v in order], order, m.ipc, pts, dirs)
            got2 = host(ctx.find_volume(tr, order, m.ipc, dev(torch, pts), dev(torch, dirs), precision=prec))
            assert np.array_equal(got2, ref2)
    finally:
        ctx.set_traversal()
        ctx.set_option("fv_refill", 0)


def test_global_surface_tree(ctx, torch_mod):
    """A tree over every surface, never flipped (create_global_surface_tree,
    ray_tracer.cpp:190-206): ray_fire in each orientation and
    point_in_volume equal the oracle on the same tree."""
    torch = torch_mod
    m = G.nested_spheres(3)
    v9, f, s, fl, bump = m.global_tree_arrays()
    n = v9.shape[0]
    gt = ctx.build_surface_tree(v9.reshape(-1, 3), np.arange(3 * n, dtype=np.int32).reshape(-1, 3),
                                f, s, fl, bump)
    ot = O.Tree(v9, f, s, fl, bump)
    pts = P.box_points(12, 0, 20000, [-18.0] * 3, [18.0] * 3)
    dirs = P.iso_dirs(12, 1, 20000)
    for orient in (-1, 0, 1):
        for prec in (0, 1):
            t, su, fa = ctx.ray_fire(gt, dev(torch, pts), dev(torch, dirs), orientation=orient,
                                     precision=prec)
            rt, rs, rf = ot.ray_fire(pts, dirs, orientation=orient)
            assert np.array_equal(host(su), rs) and np.array_equal(host(fa), rf)
            assert np.array_equal(bits(host(t)), bits(rt))
    ins = host(ctx.point_in_volume(gt, dev(torch, pts), dev(torch, dirs))).astype(bool)
    assert np.array_equal(ins, ot.point_in_volume(pts, dirs))


def _segments_setup(ctx, mesh):
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    st = build_tree(ctx, bm, bm.ipc)
    ctx.attach_boundary(et, st, mesh.bface_elem)
    return et


@pytest.mark.parametrize("table_order", [False, True])
def test_kuhn_find_element_and_segments(ctx, torch_mod, table_order):
    """MOAB's face order, and the same faces in the order 1, 2, 3, 0 (read
    from the tree by the walk): counts, hashes and tallies equal the
    oracle's."""
    torch = torch_mod
    mesh = G.kuhn_cube(16, 0.2, seed=4)
    if table_order:
        fo = np.ascontiguousarray(mesh.face_order[[1, 2, 3, 0]], dtype=np.int32)
        adj, bf, be = G.tet_adjacency(mesh.tets, fo, mesh.verts.shape[0])
        mesh = G.TetMesh(mesh.verts, mesh.tets, adj, fo, bf, be)
    et = _segments_setup(ctx, mesh)
    om = O.TetMesh(mesh)
    n = 20000
    p0 = P.box_points(4, P.STREAM_POS, n, [0] * 3, [1] * 3)
    p1 = P.box_points(4, P.STREAM_DIR, n, [0] * 3, [1] * 3)
    e = host(ctx.find_element(et, dev(torch, p0)))
    assert np.array_equal(e, om.find_element(p0))
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    cnt, hsh = ctx.tally_segments(et, dev(torch, p0), dev(torch, p1), tally=tally)
    rc, rh, rt = om.segments(p0, p1, tally=True)
    assert np.array_equal(host(cnt), rc)
    assert np.array_equal(host(hsh).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    assert ctx.status() == 0
    # track length conservation (tests/test_tracks.cpp:48,121)
    chord = np.sqrt(((p1 - p0) ** 2).sum(1))
    total = host(tally).sum() / 2.0 ** 40
    assert abs(total - chord.sum()) < 1e-6 * chord.sum()


def test_mock_segments_sequences(ctx, torch_mod, golden_dir):
    torch = torch_mod
    mesh = G.mock_tetmesh()
    et = _segments_setup(ctx, mesh)
    om = O.TetMesh(mesh)
    tracks = np.load(os.path.join(golden_dir, "golden_mock_tracks.npz"))["tracks"]
    p0, p1 = tracks[:, :3], tracks[:, 3:]
    cap = np.full(len(p0), 64, np.int64)
    rc, rh, (off, se, sl) = om.segments(p0, p1, sequences=cap)
    seq_off = dev(torch, off)
    seq_e = torch.zeros(int(off[-1]), dtype=torch.int32, device="cuda")
    seq_l = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
    cnt, hsh = ctx.tally_segments(et, dev(torch, p0), dev(torch, p1),
                                  seq_off=seq_off, seq_elem=seq_e, seq_len=seq_l)
    assert np.array_equal(host(cnt), rc)
    assert np.array_equal(host(hsh).view(np.uint64), rh)
    ge, gl = host(seq_e), host(seq_l)
    for i in range(len(p0)):
        a, b = off[i], off[i] + rc[i]
        assert np.array_equal(ge[a:b], se[a:b])
        assert np.array_equal(bits(gl[a:b]), bits(sl[a:b]))


def _segments_volume_cases(n, seed):
    """Segments for XDG::segments(volume, start, end) on the unit cube:
    inside -> inside, outside -> inside, outside -> beyond the cube, outside
    -> away from it (no entry), and outside -> short of the cube (the end
    lies before the entry, so the reference walks back from the entry)."""
    rng = np.random.default_rng(seed)
    k = n // 5
    p0 = [rng.uniform(0.01, 0.99, (k, 3))]
    p1 = [rng.uniform(0.01, 0.99, (k, 3))]
    o = rng.uniform(-1.0, 2.0, (8 * k, 3))
    o = o[~((o > 0) & (o < 1)).all(1)][:4 * k]
    oa, ob, oc, od = np.split(o, 4)
    e_in = rng.uniform(0.02, 0.98, (k, 3))
    p0 += [oa, ob, oc, od]
    p1 += [e_in[:oa.shape[0]], 0.5 + (0.5 - ob) * 1.5, oc + (oc - 0.5) * 3.0,
           od + (0.5 - od) * 0.05]
    return np.concatenate(p0), np.concatenate(p1)


def test_kuhn_segments_volume(ctx, torch_mod):
    torch = torch_mod
    mesh = G.kuhn_cube(16, 0.2, seed=4)
    et = _segments_setup(ctx, mesh)
    bm = mesh.boundary_model()
    st = build_tree(ctx, bm, 1)
    om = O.TetMesh(mesh)
    vol = O.Tree.from_model(bm, 1)
    p0, p1 = _segments_volume_cases(20000, 5)
    n = p0.shape[0]
    rc, rh, rf, rt = om.segments_volume(vol, p0, p1, tally=True)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device="cuda")
    cnt, hsh = ctx.tally_segments_volume(et, st, dev(torch, p0), dev(torch, p1),
                                         tally=tally)
    assert np.array_equal(host(cnt), rc)
    assert np.array_equal(host(hsh).view(np.uint64), rh)
    assert np.array_equal(host(tally), rt)
    want = (2 if (rf & 2).any() else 0) | (1 if (rf & 1).any() else 0)
    assert ctx.status() == want
    # each case kind produced entries where the closed form says it must
    k = n // 5
    assert (rc[:2 * k] > 0).all() and (rc[3 * k:4 * k] == 0).sum() > 0


def test_mock_segments_volume_sequences(ctx, torch_mod, golden_dir):
    torch = torch_mod
    mesh = G.mock_tetmesh()
    et = _segments_setup(ctx, mesh)
    bm = mesh.boundary_model()
    st = build_tree(ctx, bm, 1)
    g = np.load(os.path.join(golden_dir, "golden_mock_walk_se.npz"))
    tracks = np.load(os.path.join(golden_dir, "golden_mock_tracks.npz"))["tracks"]
    nin = tracks.shape[0]
    p0 = np.concatenate([tracks[:, :3], g["outside_start"]])
    p1 = np.concatenate([tracks[:, 3:], g["walk_se_in"][nin:, 4:7]])
    sout = g["walk_se_out"]
    cap = np.full(len(p0), 64, np.int64)
    off = np.zeros(len(p0) + 1, np.int64)
    off[1:] = np.cumsum(cap)
    seq_e = torch.zeros(int(off[-1]), dtype=torch.int32, device="cuda")
    seq_l = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
    cnt, hsh = ctx.tally_segments_volume(et, st, dev(torch, p0), dev(torch, p1),
                                         seq_off=dev(torch, off), seq_elem=seq_e,
                                         seq_len=seq_l)
    cnt = host(cnt)
    assert np.array_equal(cnt, sout[:, 0].astype(np.int32))
    ge, gl = host(seq_e), host(seq_l)
    for i in range(len(p0)):
        pairs = sout[i, 1:1 + 2 * cnt[i]].reshape(-1, 2)
        a, b = off[i], off[i] + cnt[i]
        assert np.array_equal(ge[a:b], pairs[:, 0].astype(np.int32))
        assert np.array_equal(bits(gl[a:b]), bits(pairs[:, 1]))
    assert ctx.status() == 0


def test_fresh_context_status_is_clear(torch_mod):
    """The sticky status word starts at zero on every new context, also when
    its memory is recycled from contexts that were destroyed."""
    from xdg_amd import capi
    for _ in range(3):
        c = capi.Context(0)
        junk = torch_mod.full((1 << 20,), -1, dtype=torch_mod.int32, device="cuda")
        del junk
        assert c.status() == 0
        c.close()
