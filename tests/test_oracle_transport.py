"""The oracle's particle random walk against what the mathematics fixes.

The reference driver (tools/particle_sim.h) has no absorption and draws
from drand48, so no reference output pins these histories; the checks here
pin the sampling and the bookkeeping instead:
  * a homogeneous medium with absorption probability p per collision and
    exponential flights of mean mfp gives an exponentially distributed total
    track length of mean mfp / p and a geometric collision count of mean 1/p
    (compound-geometric sum of exponentials);
  * very short single flights from uniform starts put the track-length tally
    in proportion to the region volumes (sphere polyhedra vs the box);
  * fixed-point tallies do not depend on the thread count;
  * a planted mistake in the flight sampling (mfp doubled) fails the first
    check.
"""
import numpy as np

import pyoracle as O
from xdg_amd import geometry as G


def _sphere_problem(mfp, p_absorb, n, first=0, nthreads=0):
    m = G.single_sphere(3, 10.0)
    tree = O.Tree.from_model(m, 1)
    return O.transport(m, [tree], [1], 5, 9, [-2.0] * 3, [2.0] * 3, mfp,
                       p_absorb, 100000, first, n, nthreads)


def test_homogeneous_medium_track_length_and_collisions():
    n, mfp, p = 20000, 0.5, 0.3
    r = _sphere_problem(mfp, p, n)
    assert (r["status"] == 2).mean() < 1e-3  # leaks: only the rare long walk
    assert ((r["status"] == 1) | (r["status"] == 2)).all()
    total = r["tally_fx"].sum() / 2.0 ** 32 / n
    mean = mfp / p
    assert abs(total - mean) < 4.0 * mean / np.sqrt(n)
    ev = r["events"].mean()
    sd = np.sqrt((1 - p) / p ** 2) / np.sqrt(n)
    assert abs(ev - 1.0 / p) < 4.0 * sd


def test_planted_flight_mistake_is_caught():
    n, mfp, p = 20000, 0.5, 0.3
    wrong = _sphere_problem(2.0 * mfp, p, n)  # as if -log(1-u) were doubled
    total = wrong["tally_fx"].sum() / 2.0 ** 32 / n
    assert abs(total - mfp / p) > 4.0 * (mfp / p) / np.sqrt(n)


def _poly_volume(v, t):
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def test_short_flights_tally_by_region_volume():
    m = G.sphere_lattice(3, 2, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    trees = [O.Tree.from_model(m, v) for v in vols]
    n = 20000
    r = O.transport(m, trees, vols, 11, 2, [0.0] * 3, [6.0] * 3, 1e-4, 1.0,
                    1000, 0, n)
    tal = r["tally_fx"].astype(np.float64)
    sph = tal[:-1].sum() / tal.sum()
    sv, st = G.icosphere(2, 0.8)
    frac = 27 * _poly_volume(sv, st) / 6.0 ** 3
    assert abs(sph - frac) < 5.0 * np.sqrt(frac * (1 - frac) / n)


def test_tallies_do_not_depend_on_threads():
    a = _sphere_problem(0.5, 0.3, 4000, first=123, nthreads=1)
    b = _sphere_problem(0.5, 0.3, 4000, first=123, nthreads=8)
    for k in a:
        assert np.array_equal(a[k], b[k])
