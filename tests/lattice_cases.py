"""Segments that run along the faces, edges and vertices of an unjittered
Kuhn mesh (kuhn_cube(n, 0.0)): the walk's ties, zero-length entries and
exit-face choices at their most degenerate."""
import numpy as np

from xdg_amd import philox as P


def lattice_segments(n):
    """(start, end) arrays, in this order:
    * axis-parallel lines through interior lattice vertices (along Kuhn
      edges), in x, y and z;
    * lines parallel to the cube diagonal (the Kuhn diagonal edges);
    * generic lines inside the lattice planes z = k/n and x = k/n;
    * segments that start at an interior lattice vertex, then segments
      that end at one (the other end Philox-uniform in the cube)."""
    h = 1.0 / n
    s, e = [], []
    for j in range(1, n, 2):
        for k in range(1, n, 2):
            for ax in range(3):
                a, b = [x for x in range(3) if x != ax]
                p, q = np.zeros(3), np.zeros(3)
                p[ax], q[ax] = 0.05, 0.95
                p[a] = q[a] = j * h
                p[b] = q[b] = k * h
                s.append(p)
                e.append(q)
    for a in range(4):
        for b in range(3):
            p = np.array([a * h, b * h, 0.0]) + 0.01
            s.append(p)
            e.append(p + 0.6)
    for k in range(1, n):
        s.append(np.array([0.1, 0.13, k * h]))
        e.append(np.array([0.9, 0.77, k * h]))
        s.append(np.array([k * h, 0.1, 0.13]))
        e.append(np.array([k * h, 0.9, 0.77]))
    g = np.arange(1, n) * h
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    r = P.box_points(9, 0, 2 * lat.shape[0], [0.0] * 3, [1.0] * 3)
    start = np.concatenate([np.array(s), lat, r[lat.shape[0]:]])
    end = np.concatenate([np.array(e), r[:lat.shape[0]], lat])
    return start, end
