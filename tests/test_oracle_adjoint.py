"""The oracle's closed-form ray-distance adjoint against central finite
differences of the oracle's own ray_fire (h = 1e-6), and a planted sign
mistake that the check must catch.  CPU only."""
import numpy as np

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import philox as P

H = 1e-6


def _fire(model, verts, org, dirs):
    m = G.SurfaceModel(verts, model.tris, model.tri_surf, model.surfaces,
                       model.senses, model.volumes, model.ipc)
    t, s, f = O.Tree.from_model(m, 1).ray_fire(org, dirs)
    return t, f


def _setup(n=3000):
    m = G.perturbed_sphere(2, 10.0, 0.05)
    org = P.ball_points(5, 0, n, 9.5)
    dirs = P.iso_dirs(5, 1, n)
    w = P.normals(5, 2, n)
    t, f = _fire(m, m.verts, org, dirs)
    assert (f >= 0).all()
    return m, org, dirs, w, t, f


def test_vertex_gradient_matches_central_differences():
    m, org, dirs, w, t, f = _setup()
    gv, go, gd = O.adjoint(m.verts, m.tris, org, dirs, t, f, w)
    rng = np.random.default_rng(1)
    for vi in rng.choice(m.verts.shape[0], 12, replace=False):
        for k in range(3):
            vp = m.verts.copy()
            vm = m.verts.copy()
            vp[vi, k] += H
            vm[vi, k] -= H
            tp, _ = _fire(m, vp, org, dirs)
            tm, _ = _fire(m, vm, org, dirs)
            fd = (w * (tp - tm)).sum() / (2 * H)
            assert abs(fd - gv[vi, k]) <= 1e-6 * max(1.0, abs(gv[vi, k])), (vi, k, fd, gv[vi, k])


def test_ray_gradients_match_central_differences():
    m, org, dirs, w, t, f = _setup(200)
    gv, go, gd = O.adjoint(m.verts, m.tris, org, dirs, t, f, w)
    for k in range(3):
        e = np.zeros(3)
        e[k] = H
        tp, _ = _fire(m, m.verts, org + e, dirs)
        tm, _ = _fire(m, m.verts, org - e, dirs)
        fd = w * (tp - tm) / (2 * H)
        assert np.allclose(fd, go[:, k], rtol=1e-6, atol=1e-6)
        tp, _ = _fire(m, m.verts, org, dirs + e)
        tm, _ = _fire(m, m.verts, org, dirs - e)
        fd = w * (tp - tm) / (2 * H)
        assert np.allclose(fd, gd[:, k], rtol=1e-6, atol=1e-6)


def test_planted_sign_mistake_is_caught():
    m, org, dirs, w, t, f = _setup(500)
    gv, go, gd = O.adjoint(m.verts, m.tris, org, dirs, t, f, w)
    # the FD of dt/do with the sign of n flipped would be +n/D: must disagree
    e = np.array([H, 0.0, 0.0])
    tp, _ = _fire(m, m.verts, org + e, dirs)
    tm, _ = _fire(m, m.verts, org - e, dirs)
    fd = w * (tp - tm) / (2 * H)
    assert not np.allclose(fd, -go[:, 0], rtol=1e-3, atol=1e-6)
