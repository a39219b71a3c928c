"""Persistent transport kernel vs the oracle's replay of the same histories.

Every history is compared bit for bit: event count, final status, the hash
of its (volume, flight length) sequence, and the fixed-point per-volume
track-length tallies (sums of llrint(len * 2^32), order independent).
"""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import build_tree, host
from xdg_amd import geometry as G

pytestmark = pytest.mark.gpu

SEED, STREAM = 3, 3


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
def lattice(ctx):
    """3x3x3 spheres (subdivision 2, r 0.8, pitch 2) in the box [0, 6]^3."""
    m = G.sphere_lattice(3, 2, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    return m, vols, trees, otrees


def _problem(ctx, lattice, mfp=2.0, p_absorb=0.3, max_events=10000):
    m, vols, trees, otrees = lattice
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, SEED, STREAM,
                             [0.0] * 3, [6.0] * 3, mfp, p_absorb, max_events)
    ref = lambda first, n: O.transport(m, otrees, vols, SEED, STREAM,
                                       [0.0] * 3, [6.0] * 3, mfp, p_absorb,
                                       max_events, first, n)
    return h, ref


def _check(out, ref):
    assert np.array_equal(host(out["tally_fx"]), ref["tally_fx"])
    assert np.array_equal(host(out["events"]), ref["events"])
    assert np.array_equal(host(out["status"]), ref["status"])
    assert np.array_equal(host(out["hash"]).view(np.uint64), ref["hash"])


@pytest.mark.parametrize("prec", [0, 1])
def test_transport_matches_oracle(ctx, lattice, prec):
    h, ref = _problem(ctx, lattice)
    try:
        n, first = 20000, 1000
        r = ref(first, n)
        assert set(np.unique(r["status"])) <= {1, 2}
        assert (r["status"] == 1).sum() > 1000 and (r["status"] == 2).sum() > 1000
        out = ctx.transport_run(h, first, n, len(lattice[1]), precision=prec)
        _check(out, r)
    finally:
        ctx.transport_destroy(h)


@pytest.mark.parametrize("prec", [0, 1])
def test_transport_boundary_conditions(ctx, lattice, prec):
    """Reflecting box and a mix of vacuum, reflecting and transmitting
    sphere surfaces (particle_sim.h:85-130): every history and the tallies
    equal the oracle's."""
    m, vols, trees, otrees = lattice
    box_sid = max(m.surfaces)
    bc = {box_sid: 2}
    for k, s in enumerate(sorted(x for x in m.surfaces if x != box_sid)):
        bc[s] = (0, 1, 2)[k % 3]
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, SEED, STREAM,
                             [0.0] * 3, [6.0] * 3, 2.0, 0.3, 10000, bc=bc)
    try:
        n, first = 20000, 77
        r = O.transport(m, otrees, vols, SEED, STREAM, [0.0] * 3, [6.0] * 3,
                        2.0, 0.3, 10000, first, n, bc=bc)
        st = set(np.unique(r["status"]))
        assert st <= {O.TR_ABSORBED, O.TR_VACUUM}
        assert (r["status"] == O.TR_VACUUM).sum() > 1000
        out = ctx.transport_run(h, first, n, len(vols), precision=prec)
        _check(out, r)
    finally:
        ctx.transport_destroy(h)


def test_transport_configs_and_autotune(ctx, lattice):
    """Block size, LDS stack depth (down to one entry, the rest spilling to
    global memory), claim chunk and refill threshold change nothing in the
    results."""
    h, ref = _problem(ctx, lattice, mfp=0.7, p_absorb=0.2)
    try:
        n = 8000
        r = ref(0, n)
        need = ctx.transport_config(h)["stack_need"]
        assert need >= 2
        for block, ls, chunk, refill in [(64, 1, 1, 1), (128, need, 4, 32), (256, 2, 16, 64),
                                         (192, need - 1, 3, 7)]:
            ctx.transport_set_config(h, block, ls, chunk, refill)
            cfg = ctx.transport_config(h)
            assert (cfg["block"], cfg["lds_stack"], cfg["chunk"], cfg["refill"]) == (block, ls, chunk, refill)
            _check(ctx.transport_run(h, 0, n, len(lattice[1])), r)
        ms = ctx.transport_autotune(h, 0, 4096)
        assert ms > 0.0
        _check(ctx.transport_run(h, 0, n, len(lattice[1])), r)
    finally:
        ctx.transport_destroy(h)


def test_transport_graph_replays(ctx, torch_mod, lattice):
    """A captured launch replayed several times, with eager work and
    synchronisation between replays, gives the eager result every time."""
    torch = torch_mod
    h, ref = _problem(ctx, lattice)
    try:
        n, first = 6000, 77
        nvol = len(lattice[1])
        eager = ctx.transport_run(h, first, n, nvol)
        out = ctx.transport_outputs(nvol, n)
        g = ctx.transport_capture(h, first, n, out)
        try:
            for rep in range(3):
                for v in out.values():
                    v.fill_(7)
                torch.cuda.synchronize()
                ctx.graph_launch(g)
                ctx.synchronize()
                for k in eager:
                    assert torch.equal(out[k], eager[k]), (rep, k)
                x = torch.randn(1 << 16, device="cuda")
                (x * 2).sum().item()
        finally:
            ctx.graph_destroy(g)
        _check(eager, ref(first, n))
    finally:
        ctx.transport_destroy(h)


@pytest.fixture
def start_form(ctx):
    """Restores the transport start options after a test changes them."""
    names = ("transport_start", "transport_slice", "transport_items", "transport_rounds", "transport_events")
    keep = [ctx.get_option(k) for k in names]
    yield
    for k, v in zip(names, keep):
        ctx.set_option(k, v)


@pytest.mark.parametrize("prec", [0, 1])
def test_transport_start_forms(ctx, lattice, prec, start_form):
    """Starts sampled inside the event kernel, or by the separate start
    kernels (in index order, or sorted by the start's cell and direction)
    over slices of 2^6, 2^10 and 2^22 histories (one slice, many slices, a
    ragged last slice), events by the whole-event or the traversal-level
    kernel: every history and the tallies equal the oracle's."""
    h, ref = _problem(ctx, lattice)
    try:
        n, first = 5000, 123
        r = ref(first, n)
        for staged, slice_log2, events in [(0, 22, 0), (1, 6, 1), (1, 22, 0), (2, 6, 1), (2, 10, 0), (2, 22, 1),
                                           (2, 22, 0)]:
            ctx.set_option("transport_start", staged)
            ctx.set_option("transport_slice", slice_log2)
            ctx.set_option("transport_events", events)
            out = ctx.transport_run(h, first, n, len(lattice[1]), precision=prec)
            _check(out, r)
    finally:
        ctx.transport_destroy(h)


def test_transport_forms_many_per_lane(ctx, lattice, start_form):
    """2^20 histories, several for each lane of the resident event grids, so
    lanes claim new histories (and start-stage items) while others of their
    wave are still tracing: the traversal-level event kernel, the
    whole-event kernel and starts inside the event kernel give the same
    histories and tallies, and the first 5000 histories equal the oracle."""
    h, ref = _problem(ctx, lattice)
    try:
        n, first, k = 1 << 20, 11, 5000
        r = ref(first, k)
        outs = []
        for staged, events in [(2, 1), (2, 0), (0, 0)]:
            ctx.set_option("transport_start", staged)
            ctx.set_option("transport_events", events)
            out = ctx.transport_run(h, first, n, len(lattice[1]))
            outs.append({key: host(v) for key, v in out.items()})
        for o in outs[1:]:
            for key in outs[0]:
                assert np.array_equal(o[key], outs[0][key]), key
        o = outs[0]
        assert set(np.unique(o["status"])) <= {1, 2}
        assert np.array_equal(o["events"][:k], r["events"])
        assert np.array_equal(o["status"][:k], r["status"])
        assert np.array_equal(o["hash"][:k].view(np.uint64), r["hash"])
    finally:
        ctx.transport_destroy(h)


def test_transport_boundary_conditions_many_per_lane(ctx, lattice, start_form):
    """Reflecting box and mixed vacuum / reflecting / transmitting sphere
    surfaces with 2^19 histories, several per lane of the resident grids:
    the traversal-level and whole-event kernels agree history by history
    and on the tallies, and the first 3000 histories equal the oracle."""
    m, vols, trees, otrees = lattice
    box_sid = max(m.surfaces)
    bc = {box_sid: 2}
    for k, s_ in enumerate(sorted(x for x in m.surfaces if x != box_sid)):
        bc[s_] = (0, 1, 2)[k % 3]
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, SEED, STREAM,
                             [0.0] * 3, [6.0] * 3, 2.0, 0.3, 10000, bc=bc)
    try:
        n, first, k = 1 << 19, 5, 3000
        r = O.transport(m, otrees, vols, SEED, STREAM, [0.0] * 3, [6.0] * 3,
                        2.0, 0.3, 10000, first, k, bc=bc)
        outs = []
        for events in (1, 0):
            ctx.set_option("transport_events", events)
            out = ctx.transport_run(h, first, n, len(vols))
            outs.append({key: host(v) for key, v in out.items()})
        for key in outs[0]:
            assert np.array_equal(outs[0][key], outs[1][key]), key
        o = outs[0]
        assert set(np.unique(o["status"])) <= {O.TR_ABSORBED, O.TR_VACUUM}
        assert np.array_equal(o["events"][:k], r["events"])
        assert np.array_equal(o["status"][:k], r["status"])
        assert np.array_equal(o["hash"][:k].view(np.uint64), r["hash"])
    finally:
        ctx.transport_destroy(h)


@pytest.mark.parametrize("prec", [0, 1])
def test_transport_event_rounds(ctx, lattice, prec, start_form):
    """Sorted one-event rounds before the event kernel (none, one, a few, and
    more rounds than any history has events), also over several slices:
    every history and the tallies equal the oracle's."""
    h, ref = _problem(ctx, lattice)
    try:
        n, first = 5000, 31
        r = ref(first, n)
        ctx.set_option("transport_start", 2)
        for rounds, slice_log2, events in [(0, 22, 1), (1, 22, 1), (3, 22, 0), (64, 22, 1), (3, 7, 1), (3, 7, 0)]:
            ctx.set_option("transport_rounds", rounds)
            ctx.set_option("transport_slice", slice_log2)
            ctx.set_option("transport_events", events)
            _check(ctx.transport_run(h, first, n, len(lattice[1]), precision=prec), r)
    finally:
        ctx.transport_destroy(h)


def test_transport_start_item_overflow(ctx, lattice, start_form):
    """With one start-stage item per history the point_in_volume items of a
    slice overflow the buffer; the histories that lost items find their
    volume in the resolve step instead, and every history still equals the
    oracle's."""
    h, ref = _problem(ctx, lattice)
    try:
        n, first = 4000, 9
        r = ref(first, n)
        for items in (1, 2):
            ctx.set_option("transport_start", 2)
            ctx.set_option("transport_items", items)
            _check(ctx.transport_run(h, first, n, len(lattice[1])), r)
    finally:
        ctx.transport_destroy(h)


def test_transport_sliced_graph_replays(ctx, torch_mod, lattice, start_form):
    """A captured staged launch over 12 slices replays to the eager result;
    a larger slice afterwards moves the start buffer, which makes the graph
    stale (refused before any launch)."""
    from xdg_amd import capi
    torch = torch_mod
    h, ref = _problem(ctx, lattice)
    try:
        ctx.set_option("transport_start", 2)
        ctx.set_option("transport_slice", 9)
        n, first = 6000, 5
        nvol = len(lattice[1])
        eager = ctx.transport_run(h, first, n, nvol)
        out = ctx.transport_outputs(nvol, n)
        g = ctx.transport_capture(h, first, n, out)
        try:
            for rep in range(3):
                for v in out.values():
                    v.fill_(3)
                torch.cuda.synchronize()
                ctx.graph_launch(g)
                ctx.synchronize()
                for k in eager:
                    assert torch.equal(out[k], eager[k]), (rep, k)
            ctx.set_option("transport_slice", 12)
            ctx.transport_run(h, first, n, nvol)
            with pytest.raises(capi.XdgHipError, match="stale"):
                ctx.graph_launch(g)
        finally:
            ctx.graph_destroy(g)
        _check(eager, ref(first, n))
    finally:
        ctx.transport_destroy(h)


def test_transport_over_1024_volumes(ctx, start_form):
    """11^3 spheres (subdivision 1) and their box: more volume slots than
    the start stage keeps candidate masks and LDS counters for, so its
    candidate passes test the boxes twice and reserve items with global
    atomics per wave; every history and the tallies equal the oracle's."""
    m = G.sphere_lattice(11, 1, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    assert len(vols) > 1024
    trees = [build_tree(ctx, m, v) for v in vols]
    otrees = [O.Tree.from_model(m, v) for v in vols]
    hi = [22.0] * 3
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, SEED, STREAM, [0.0] * 3, hi, 2.0, 0.3, 10000)
    try:
        n, first = 3000, 11
        r = O.transport(m, otrees, vols, SEED, STREAM, [0.0] * 3, hi, 2.0, 0.3, 10000, first, n)
        assert (r["status"] == O.TR_LEAKED).sum() > 100
        ctx.set_option("transport_start", 2)
        for events in (1, 0):
            ctx.set_option("transport_events", events)
            _check(ctx.transport_run(h, first, n, len(vols)), r)
    finally:
        ctx.transport_destroy(h)
        for t in trees:
            ctx.release_surface_tree(t)


def test_transport_prefilter_far_from_origin(ctx, start_form):
    """A sphere 3e5 from the origin, its trees built with a 1e-3 bump (the
    C-ABI takes any bump), and every start sampled 0.005 inside the
    sphere's extreme vertex.  At that offset the vertex's x rounds up by
    0.015 to the nearest FP32 value, past the start and the start stage's
    slack; the trees' FP32 bounds are rounded outward, so the prefilter
    keeps the sphere and every history starts in it, as in the oracle."""
    base = G.sphere_lattice(1, 2, 0.8, 2.0)
    off = 300000.0035  # float32(off + 0.2) - (off + 0.2) = 0.01525
    verts = base.verts + off
    m = G.SurfaceModel(verts, base.tris, base.tri_surf, base.surfaces, base.senses, base.volumes, ipc=base.ipc)
    xmin = verts[np.unique(m.tris[m.surface_faces(1)]), 0].min()
    assert float(np.float32(xmin)) - xmin > 0.01
    vols = [v for v in m.volumes if v != m.ipc]
    trees = []
    for v in vols:
        v9, f, sid, fl = m.tree_arrays(v)
        nt = v9.shape[0]
        trees.append(ctx.build_surface_tree(v9.reshape(-1, 3), np.arange(3 * nt, dtype=np.int32).reshape(-1, 3),
                                            f, sid, fl, 1e-3))
    otrees = [O.Tree.from_model(m, v) for v in vols]
    p = np.array([xmin + 0.005, off + 1.0, off + 1.0])
    lo, hi = list(p - 1e-7), list(p + 1e-7)
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, SEED, STREAM, lo, hi, 2.0, 0.3, 10000)
    try:
        n = 4000
        r = O.transport(m, otrees, vols, SEED, STREAM, lo, hi, 2.0, 0.3, 10000, 0, n)
        for staged in (2, 1, 0):
            ctx.set_option("transport_start", staged)
            out = ctx.transport_run(h, 0, n, len(vols))
            _check(out, r)
    finally:
        ctx.transport_destroy(h)
        for t in trees:
            ctx.release_surface_tree(t)


def test_transport_event_cap(ctx, lattice):
    h, ref = _problem(ctx, lattice, mfp=0.05, p_absorb=0.0, max_events=7)
    try:
        n = 3000
        r = ref(0, n)
        assert (r["status"] == 5).mean() > 0.9
        _check(ctx.transport_run(h, 0, n, len(lattice[1])), r)
    finally:
        ctx.transport_destroy(h)
