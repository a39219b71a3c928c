"""Analytic point location on the unjittered Kuhn cube.

kuhn_cube(n, 0) splits hex (cx, cy, cz) into the six tetrahedra
{1 >= f[p0] >= f[p1] >= f[p2] >= 0} of the local coordinates f = n*p - c,
one per permutation p of the axes, stored at element index
6 * (cx + n*cy + n*n*cz) + j with j the permutation's position in
PERMS.  The element that contains a point therefore follows from the
point alone, without any containment test: this is what pins the
oracle's find_element (ray_tracer.cpp:226-257, lowest containing element)
independently of its own code.

Points on a dyadic grid (k / (n * 2^m)) have exact local coordinates and
exact barycentric coordinates in every tet, so a point on a shared face,
edge or vertex lies in several tets, and the lowest index among them is the
answer the tie rule asks for.
"""
import itertools

import numpy as np

PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _contained_exact(f):
    """Indices j of PERMS whose tet holds local coordinates f (exact)."""
    out = []
    for j, p in enumerate(PERMS):
        if 1.0 >= f[p[0]] >= f[p[1]] >= f[p[2]] >= 0.0:
            out.append(j)
    return out


def lowest_containing(n, p):
    """Lowest element index of kuhn_cube(n, 0) that contains p exactly
    (ties on shared faces, edges and vertices included), or -1."""
    s = [v * n for v in p]
    cand = []
    for a in range(3):
        c = int(np.floor(s[a]))
        opts = {c}
        if s[a] == c:
            opts.add(c - 1)
        cand.append(sorted(x for x in opts if 0 <= x < n))
    best = -1
    for cx, cy, cz in itertools.product(*cand):
        f = (s[0] - cx, s[1] - cy, s[2] - cz)
        for j in _contained_exact(f):
            e = 6 * (cx + n * cy + n * n * cz) + j
            if best < 0 or e < best:
                best = e
    return best


def dyadic_points(n, m, count, seed):
    """count points on the grid k / (n 2^m) inside [0, 1]^3, half of them
    snapped onto lattice planes, edges or vertices (shared faces)."""
    rng = np.random.default_rng(seed)
    g = n * (1 << m)
    k = rng.integers(0, g + 1, size=(count, 3))
    snap = rng.integers(0, 4, size=count)  # 0: none, 1: one plane, 2: edge, 3: vertex
    for i in range(count):
        axes = rng.permutation(3)[:snap[i]]
        for a in axes:
            k[i, a] = (k[i, a] >> m) << m
    # diagonal ties inside a hex: f[a] == f[b]
    d = rng.integers(0, 2, size=count).astype(bool)
    k[d, 1] = (k[d, 1] >> m << m) + (k[d, 0] & ((1 << m) - 1))
    k = np.minimum(k, g)
    return k.astype(np.float64) / g


def interior_points(n, count, seed, gap=1e-6):
    """Uniform points whose local coordinates are at least `gap` apart and
    from 0 and 1, with their unique containing element."""
    rng = np.random.default_rng(seed)
    pts, exp = [], []
    while len(pts) < count:
        p = rng.random(3)
        s = p * n
        c = np.floor(s)
        f = s - c
        srt = np.sort(f)
        gaps = np.concatenate([[srt[0]], np.diff(srt), [1.0 - srt[-1]]])
        if gaps.min() < gap or np.any(c >= n):
            continue
        order = tuple(int(x) for x in np.argsort(-f, kind="stable"))
        j = PERMS.index(order)
        cx, cy, cz = (int(x) for x in c)
        pts.append(p)
        exp.append(6 * (cx + n * cy + n * n * cz) + j)
    return np.array(pts), np.array(exp, dtype=np.int32)
