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
from xdg_amd import philox as P


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


# ---- surface boundary conditions (particle_sim.h:85-130) -------------------

def _unit_box_model():
    v, f = G.box_surface([0.0] * 3, [1.0] * 3)
    ts = np.ones(f.shape[0], np.int32)
    return G.SurfaceModel(v, f, ts, [1], {1: (1, 2)}, [1, 2], ipc=2)


def test_reflecting_box_unfolds_to_straight_lines():
    """Specular reflection in the unit cube unfolds every history into one
    straight line from the start along the start direction (mirror images
    of the cube tile space).  With absorption at the first collision the
    history's events are its wall hits plus one, and the wall hits are the
    integer planes that the unfolded line of length T (the history's tally)
    crosses: sum_k |floor(r0_k + T u0_k) - floor(r0_k)|.  A wrong mirror
    formula bends the unfolded line and changes the counts."""
    m = _unit_box_model()
    tree = O.Tree.from_model(m, 1)
    seed, stream, mfp = 21, 3, 2.5
    r0 = P.box_points(seed, stream, 200, [0.0] * 3, [1.0] * 3)
    u0 = P.iso_dirs(seed, stream, 200, base=2)
    walls = []
    for i in range(200):
        r = O.transport(m, [tree], [1], seed, stream, [0.0] * 3, [1.0] * 3, mfp,
                        1.0, 100000, i, 1, bc={1: O.BC_REFLECTING})
        assert r["status"][0] == O.TR_ABSORBED
        T = r["tally_fx"][0] / 2.0 ** 32
        end = r0[i] + T * u0[i]
        want = int(np.abs(np.floor(end) - np.floor(r0[i])).sum())
        assert r["events"][0] == want + 1
        walls.append(want)
    assert np.mean(walls) > 2.0  # the check exercised many reflections


def test_reflecting_sphere_keeps_every_particle():
    """A reflecting outer surface keeps the medium infinite in effect: no
    history leaks and the mean track per history is mfp / p."""
    n, mfp, p = 20000, 4.0, 0.3
    m = G.single_sphere(3, 10.0)
    tree = O.Tree.from_model(m, 1)
    r = O.transport(m, [tree], [1], 5, 9, [-2.0] * 3, [2.0] * 3, mfp, p,
                    100000, 0, n, bc={1: O.BC_REFLECTING})
    assert (r["status"] == O.TR_ABSORBED).all()
    total = r["tally_fx"].sum() / 2.0 ** 32 / n
    mean = mfp / p
    assert abs(total - mean) < 4.0 * mean / np.sqrt(n)


def test_vacuum_spheres_in_reflecting_box():
    """Vacuum sphere surfaces end a history on entry, so a history that
    starts outside the spheres never tallies inside one; the reflecting box
    lets nothing leak."""
    m = G.sphere_lattice(3, 2, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    trees = [O.Tree.from_model(m, v) for v in vols]
    box_sid = max(m.surfaces)
    bc = {s: O.BC_VACUUM for s in m.surfaces if s != box_sid}
    bc[box_sid] = O.BC_REFLECTING
    n = 5000
    # starts in the slab 1.85 < x < 2.15 between the sphere layers (spheres
    # span [0.2, 1.8] and [2.2, 3.8] in x): every history starts in the box
    # region
    r = O.transport(m, trees, vols, 11, 2, [1.85, 0.0, 0.0], [2.15, 6.0, 6.0],
                    1.5, 0.2, 1000, 0, n, bc=bc)
    st = r["status"]
    assert set(np.unique(st)) <= {O.TR_ABSORBED, O.TR_VACUUM}
    assert (st == O.TR_VACUUM).sum() > 0.1 * n
    assert (r["tally_fx"][:-1] == 0).all() and r["tally_fx"][-1] > 0
