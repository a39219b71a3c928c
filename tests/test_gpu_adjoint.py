"""Ray-distance adjoint on the device: against the oracle's closed form, and
through the torch.autograd.Function against central finite differences
(h = 1e-6, relative 1e-6) of the device's own distances."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import bits, dev, host
from xdg_amd import geometry as G
from xdg_amd import philox as P

pytestmark = pytest.mark.gpu

H = 1e-6


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
def problem():
    m = G.perturbed_sphere(3, 10.0, 0.05)
    n = 100000
    org = P.ball_points(5, 0, n, 9.5)
    dirs = P.iso_dirs(5, 1, n)
    w = P.normals(5, 2, n)
    return m, org, dirs, w


@pytest.mark.parametrize("adj_sort", [1, 0])
def test_adjoint_matches_oracle(ctx, torch_mod, problem, adj_sort):
    """Both forms of the kernel (rays ordered by hit face with per-wave
    vertex sums, and one atomic set per ray), also with rays that missed
    mixed in: ray gradients bit for bit, vertex gradients to 1e-12."""
    torch = torch_mod
    from xdg_amd import adjoint as A
    m, org, dirs, w = problem
    ctx.set_option("adjoint_sort", adj_sort)
    surf = A.DifferentiableSurface(ctx, m.tris)
    verts = dev(torch, m.verts)
    tree = surf.tree_for(verts)
    o, d = dev(torch, org), dev(torch, dirs)
    t, s, f = ctx.ray_fire(tree, o, d, precision=1)
    rt, rs, rf = O.Tree.from_model(m, 1).ray_fire(org, dirs)
    assert np.array_equal(host(f), rf) and np.array_equal(bits(host(t)), bits(rt))
    gv = torch.zeros_like(verts)
    go = torch.empty_like(o)
    gd = torch.empty_like(d)
    ctx.ray_distance_adjoint(o, d, t, f, dev(torch, w), verts, surf.tris, gv, go, gd)
    ov, oo, od = O.adjoint(m.verts, m.tris, org, dirs, rt, rf, w)
    assert np.array_equal(bits(host(go)), bits(oo))
    assert np.array_equal(bits(host(gd)), bits(od))
    g = host(gv)
    assert np.allclose(g, ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    # every seventh ray reads as a miss
    tm, fm = rt.copy(), rf.copy()
    tm[::7], fm[::7] = np.inf, -1
    gv.zero_()
    ctx.ray_distance_adjoint(o, d, dev(torch, tm), dev(torch, fm), dev(torch, w), verts, surf.tris, gv, go, gd)
    ov, oo, od = O.adjoint(m.verts, m.tris, org, dirs, tm, fm, w)
    assert np.array_equal(bits(host(go)), bits(oo)) and np.array_equal(bits(host(gd)), bits(od))
    assert not host(go)[::7].any()
    assert np.allclose(host(gv), ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    ctx.set_option("adjoint_sort", 1)


def test_autograd_matches_central_differences(ctx, torch_mod, problem):
    torch = torch_mod
    from xdg_amd import adjoint as A
    m, org, dirs, w = problem
    surf = A.DifferentiableSurface(ctx, m.tris)
    verts = dev(torch, m.verts).requires_grad_(True)
    o = dev(torch, org).requires_grad_(True)
    d = dev(torch, dirs).requires_grad_(True)
    wt = dev(torch, w)
    t = A.ray_distance(o, d, verts, surf)
    (wt * t).sum().backward()
    gv = host(verts.grad)
    rng = np.random.default_rng(3)

    def fire(v):
        vt = dev(torch, v)
        tr = surf.tree_for(vt)
        tt, _, _ = ctx.ray_fire(tr, o.detach(), d.detach(), precision=1)
        return host(tt)

    checked = 0
    for vi in rng.choice(m.verts.shape[0], 16, replace=False):
        for k in range(3):
            vp = m.verts.copy()
            vm = m.verts.copy()
            vp[vi, k] += H
            vm[vi, k] -= H
            fd = (w * (fire(vp) - fire(vm))).sum() / (2 * H)
            assert abs(fd - gv[vi, k]) <= 1e-6 * max(1.0, abs(gv[vi, k])), (vi, k, fd, gv[vi, k])
            checked += 1
    assert checked == 48
    # ray gradients: per-ray central differences on a subset
    sel = slice(0, 2000)
    for k in range(3):
        e = np.zeros(3)
        e[k] = H
        tree = surf.tree_for(dev(torch, m.verts))
        for base, grad in ((org, o.grad), (dirs, d.grad)):
            plus = base.copy()
            minus = base.copy()
            plus[sel] += e
            minus[sel] -= e
            if base is org:
                tp, _, _ = ctx.ray_fire(tree, dev(torch, plus), d.detach(), precision=1)
                tm, _, _ = ctx.ray_fire(tree, dev(torch, minus), d.detach(), precision=1)
            else:
                tp, _, _ = ctx.ray_fire(tree, o.detach(), dev(torch, plus), precision=1)
                tm, _, _ = ctx.ray_fire(tree, o.detach(), dev(torch, minus), precision=1)
            fd = w[sel] * (host(tp)[sel] - host(tm)[sel]) / (2 * H)
            g = host(grad)[sel, k]
            assert np.allclose(fd, g, rtol=1e-6, atol=1e-6)


def test_released_tree_is_gone(ctx, problem):
    from xdg_amd import capi
    m, _, _, _ = problem
    v9, f, s, fl = m.tree_arrays(1)
    tree = ctx.build_surface_tree(v9.reshape(-1, 3), np.arange(v9.shape[0] * 3, dtype=np.int32).reshape(-1, 3),
                                  f, s, fl, m.bump(1))
    ctx.release_surface_tree(tree)
    with pytest.raises(capi.XdgHipError):
        ctx.tree_info(tree)


def test_full_size_central_differences(ctx, torch_mod):
    """The adjoint workload at its full size: icosphere subdivision 6 (81 920
    triangles) radially perturbed, 4 M rays from points with r < 9.5,
    L = sum w_i t_i with w_i ~ N(0, 1).  The full dL/dV from the autograd
    harness equals the oracle's closed form (relative 1e-12).  On 64 random
    vertices the adjoint of the rays that hit a triangle of the vertex
    matches central differences (h = 1e-6, fourth order) of the same rays'
    distances, relative 1e-6.  Rays whose hit lies within 1e-4 (barycentric)
    of an edge are left out of that comparison: a perturbation of h moves
    them across the edge onto another triangle, where t has another slope,
    so no finite difference of width 4h is the derivative there."""
    torch = torch_mod
    from xdg_amd import adjoint as A
    m = G.perturbed_sphere(6, 10.0, 0.05)
    assert m.tris.shape[0] == 81920
    n = 4_000_000
    org = P.ball_points(5, 0, n, 9.5)
    dirs = P.iso_dirs(5, 1, n)
    w = P.normals(5, 2, n)
    surf = A.DifferentiableSurface(ctx, m.tris)
    verts = dev(torch, m.verts).requires_grad_(True)
    o, d = dev(torch, org), dev(torch, dirs)
    t = A.ray_distance(o, d, verts, surf)
    (dev(torch, w) * t).sum().backward()
    gv = host(verts.grad)
    tree = surf.tree_for(dev(torch, m.verts))
    tt, _, face = ctx.ray_fire(tree, o, d, precision=1)
    tt, face = host(tt), host(face)
    assert (face >= 0).all()
    ov, _, _ = O.adjoint(m.verts, m.tris, org, dirs, tt, face, w)
    assert np.allclose(gv, ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    # barycentric coordinates of every hit
    q = org + tt[:, None] * dirs
    a, b, c = (m.verts[m.tris[face, j]] for j in range(3))
    e1, e2, qa = b - a, c - a, q - a
    d11, d12, d22 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    q1, q2 = (qa * e1).sum(1), (qa * e2).sum(1)
    den = d11 * d22 - d12 * d12
    l1 = (d22 * q1 - d12 * q2) / den
    l2 = (d11 * q2 - d12 * q1) / den
    interior = np.minimum(np.minimum(l1, l2), 1.0 - l1 - l2) > 1e-4
    rng = np.random.default_rng(11)
    for vi in rng.choice(m.verts.shape[0], 64, replace=False):
        star = np.nonzero((m.tris == vi).any(1))[0]
        sel = np.nonzero(np.isin(face, star) & interior)[0]
        assert sel.size > 0
        so, sd, sw = dev(torch, org[sel]), dev(torch, dirs[sel]), w[sel]
        g = torch.zeros_like(verts.detach())
        ctx.ray_distance_adjoint(so, sd, dev(torch, tt[sel]), dev(torch, face[sel]), dev(torch, sw),
                                 verts.detach(), surf.tris, g, None, None)
        g = host(g)
        for k in range(3):
            def fire(step):
                v = m.verts.copy()
                v[vi, k] += step
                tk, _, _ = ctx.ray_fire(surf.tree_for(dev(torch, v)), so, sd, precision=1)
                return host(tk)

            fd = (sw * (8.0 * (fire(H) - fire(-H)) - (fire(2 * H) - fire(-2 * H)))).sum() / (12 * H)
            assert abs(fd - g[vi, k]) <= 1e-6 * max(1.0, abs(g[vi, k])), (vi, k, fd, g[vi, k])
