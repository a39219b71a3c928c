"""The FP32 box test of the mixed traversal can never miss a hit the FP64
Pluecker test would accept, for origins within the tree's o_limit.

A triangle's box is its FP64 box dilated by `bump` (triangle_intersect.cpp:
36-50, ray_tracing_interface.cpp:20-24) and rounded outward to FP32, so a
hit point P = o + t d lies at least `bump` inside every face of every
ancestor box: in ray-parameter units the true entry distance of axis i is at
most t - bump/|d_i| and the exit at least t + bump/|d_i|.

The kernel evaluates entry_i = fmaf(lo_i, ix, nx) with ix = 1/fl32(d_i) and
nx = fl32(-fl32(o_i) * ix).  With u = 2^-24:
  ix = (1/d_i)(1 + eta), |eta| <= 2u            (two roundings)
  fl32(o_i) = o_i + e1, |e1| <= u|o_i|;  nx = -fl32(o_i) ix (1 + e2)
  entry_i = ix (lo_i - o_i - e1 - fl32(o_i) e2)(1 + e3)     (fma: one rounding)
so |entry_i - (lo_i - o_i)/d_i| <= (3u|lo_i| + 5u|o_i|) / |d_i| to first
order.  The test cannot miss while 3u max|x| + 5u |o|_inf < bump.  The
library sets o_limit = bump 2^20 - max|x| (capped at 2^26, which keeps
|fl32(o) * ix| below FLT_MAX even for the clamped reciprocal 1e30 of a
near-axis direction), where the error is at most 5/16 of the margin.
Origins beyond o_limit are traced with FP64 boxes.

This file checks the derivation by emulating the kernel's FP32 arithmetic in
numpy on the tightest boxes the bound allows: the error stays within 5/16 of
the margin up to o_limit and no hit is lost, near-axis directions (components
1e-20 and 1e-35, through the clamped reciprocal) included, while far beyond
the limit the same emulation does lose hits (so the check can fail).
"""
import numpy as np
import pytest

U = 2.0 ** -24
F32 = np.float32


def f32_up(x):
    f = x.astype(np.float32)
    m = f.astype(np.float64) < x
    f[m] = np.nextafter(f[m], F32(np.inf))
    return f


def f32_down(x):
    f = x.astype(np.float32)
    m = f.astype(np.float64) > x
    f[m] = np.nextafter(f[m], F32(-np.inf))
    return f


def safe_inv_f(d):
    """safe_inv_f of surface_traverse.h: 1 / d in FP32 with |d| clamped to
    at least 1e-30."""
    df = d.astype(np.float32)
    c = np.where(np.abs(df) > F32(1e-30), df, np.copysign(F32(1e-30), df)).astype(np.float32)
    return (F32(1.0) / c).astype(np.float32)


def fmaf(a, b, c):
    # the product of two floats is exact in long double (48 of 64 bits); the
    # sum is rounded there first, which can differ from a true fma only at
    # ties of the final rounding, far below the margins measured here
    return (a.astype(np.longdouble) * b.astype(np.longdouble) + c.astype(np.longdouble)).astype(np.float32)


def emulate(o, d, P, t, bump):
    """Per-axis error of entry/exit over the margin, and whether the box
    [f32_down(P - bump), f32_up(P + bump)] is missed by the FP32 test."""
    lo = f32_down(P - bump)
    hi = f32_up(P + bump)
    ix = safe_inv_f(d)
    nx = (-o.astype(np.float32) * ix).astype(np.float32)
    neg = ix < 0
    near = np.where(neg, hi, lo)
    far = np.where(neg, lo, hi)
    a = fmaf(near, ix, nx).astype(np.float64)
    b = fmaf(far, ix, nx).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = (near.astype(np.float64) - o) / d
        tb = (far.astype(np.float64) - o) / d
        err = np.maximum(np.abs(a - ta), np.abs(b - tb)) / (bump / np.abs(d))
    # axes with a clamped reciprocal do not follow the relative bound; they
    # only need the right sign, which the miss test below covers
    err = np.where(np.abs(d) > 1e-30, err, 0.0)
    entry = np.maximum(a.max(axis=1), 0.0)
    exit_ = np.minimum(b.min(axis=1), f32_up(t).astype(np.float64))
    return err.max(), int((entry > exit_).sum())


def rays(rng, n, scale, radius, near_axis):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    P = u * rng.uniform(0.3, 1.0, size=(n, 1)) * radius
    if near_axis:
        # origin on an axis at |o|_inf = scale, direction (-+1, 1e-20, 1e-35)
        ax = rng.integers(0, 3, n)
        sgn = rng.choice([-1.0, 1.0], n)
        d = np.zeros((n, 3))
        d[np.arange(n), ax] = -sgn
        d[np.arange(n), (ax + 1) % 3] = rng.choice([-1.0, 1.0], n) * 1e-20
        d[np.arange(n), (ax + 2) % 3] = rng.choice([-1.0, 1.0], n) * 1e-35
        o = P.copy()
        o[np.arange(n), ax] = sgn * scale
        t = np.abs(o[np.arange(n), ax] - P[np.arange(n), ax])
        P = o + t[:, None] * d
        return o, d, P, t
    o = rng.normal(size=(n, 3))
    o = o / np.abs(o).max(axis=1)[:, None] * scale
    d = P - o
    t = np.linalg.norm(d, axis=1)
    d /= t[:, None]
    return o, d, P, t


@pytest.mark.parametrize("radius", [10.0, 15.0])
def test_no_miss_up_to_o_limit(radius):
    bump = 1e-3
    max_abs = radius + bump
    o_limit = bump * 2.0 ** 20 - max_abs
    bound = (3 * U * max_abs + 5 * U * o_limit) / bump
    assert bound < 0.3126
    rng = np.random.default_rng(int(radius))
    for s in (0.25, 0.5, 0.9, 0.999, 1.0):
        for near_axis in (False, True):
            o, d, P, t = rays(rng, 400_000, s * o_limit, radius, near_axis)
            err, miss = emulate(o, d, P, t, bump)
            assert miss == 0
            assert err <= bound * (1 + 1e-6)
    # the errors are real (the emulation is not trivially exact) ...
    o, d, P, t = rays(rng, 400_000, o_limit, radius, False)
    assert emulate(o, d, P, t, bump)[0] > 0.5 * (5 * U * o_limit / bump)


def test_misses_far_beyond_the_limit():
    # ... and far past the limit the FP32 test loses hits: the check can fail
    bump, radius = 1e-3, 15.0
    o_limit = bump * 2.0 ** 20 - radius - bump
    rng = np.random.default_rng(3)
    o, d, P, t = rays(rng, 400_000, 16 * o_limit, radius, False)
    err, miss = emulate(o, d, P, t, bump)
    assert err > 1.0 and miss > 0


def test_o_limit_cap_keeps_products_finite():
    # the 2^26 cap: fl32(o) * 1e30 stays finite
    o = np.float32(2.0 ** 26)
    assert np.isfinite(np.float32(o * np.float32(1e30)))
    assert float(o) * 1e30 < float(np.finfo(np.float32).max)
