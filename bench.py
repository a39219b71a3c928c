"""Benchmark of the geometry-query hot path on one MI355X.

Headline workload (the metric's configuration): three nested icospheres
r = 5/10/15 (subdivision 6, 81 920 triangles each) inside the cube
[-20, 20]^3; 10 M query points uniform in [-16, 16]^3 with isotropic
directions (Philox seed 2).  Every point is located with find_volume, then
one step = ray_fire of all points against the surface tree of the volume
that contains them (EXITING, distance and surface/face ids), with inputs
resident in device memory.

Reported on the same JSON line: the mixed FP32-box / FP64-Pluecker kernel's
Mrays/s (value) and the plain FP64 kernel's, whether their hits are
identical, find_volume Mqueries/s, the tet walk (64^3 x 6 jittered Kuhn
tets, 4 M segments, Msegs/s), the particle transport (8^3 spheres of 5 120
triangles in [0, 16]^3, 2^20 histories, one captured graph launch), the
ray-distance adjoint (perturbed icosphere, 81 920 triangles, 4 M rays),
a roofline object for the dominant kernel and the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ray_fire Mrays/s: mixed vs FP64 kernel, identical hits; tet-walk Msegs/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# VALU issue roof: 256 CUs x 4 SIMD16 = 1024 SIMDs at the 2.4 GHz peak
# engine clock; a wave64 VALU instruction holds its SIMD for 4 cycles, so
# the chip issues at most 1024 x 2.4e9 / 4 = 614.4 G wave-instructions/s
SIMDS = 1024
CLOCK_HZ = 2.4e9
VALU_PEAK_GINST = SIMDS * CLOCK_HZ / 4 / 1e9


def load_profile(name):
    """Per-launch counters committed under profiles/ (rocprofv3 runs of this
    bench; each summary next to the JSON gives its command)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--points", type=int, default=10_000_000)
    p.add_argument("--segments", type=int, default=4_000_000)
    p.add_argument("--tet-n", type=int, default=64)
    p.add_argument("--histories", type=int, default=1 << 20)
    p.add_argument("--adjoint-rays", type=int, default=4_000_000)
    p.add_argument("--cpu-sample", type=int, default=10_000_000)
    p.add_argument("--cpu-tet-sample", type=int, default=500_000)
    p.add_argument("--sphere-rays", type=int, default=1_000_000)
    p.add_argument("--no-sphere", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tet", action="store_true")
    p.add_argument("--no-transport", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--legs", default="",
                   help="comma list of legs to run alone (ray_fire, sphere, tet, transport, adjoint), "
                        "for profiling one workload at a time; prints a short JSON line")
    p.add_argument("--segment-order", type=int, default=0,
                   help="segment_order option for the tet walk (0 caller's order, 1 Morton of the start)")
    p.add_argument("--transport-config", default="",
                   help="block,lds_stack,chunk,refill for the transport leg instead of the autotuner")
    p.add_argument("--copy-probe", action="store_true",
                   help="also run a 256 MiB device copy (scale for PMC byte counters)")
    return p.parse_args()


def time_launches(torch, stream, fn, steps, warmup):
    """Per-launch durations from HIP events recorded on the launch stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k].record(stream)
        fn()
    evs[steps].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[k].elapsed_time(evs[k + 1]) * 1e-3 for k in range(steps)]
    return per, wall


def build_trees(ctx, model, vols):
    trees = []
    for v in vols:
        v9, f, s, fl = model.tree_arrays(v)
        nt = v9.shape[0]
        trees.append(ctx.build_surface_tree(
            v9.reshape(-1, 3), np.arange(3 * nt, dtype=np.int32).reshape(-1, 3),
            f, s, fl, model.bump(v)))
    return trees


# ---------------------------------------------------------------------------
# headline: find_volume + ray_fire on the nested spheres
# ---------------------------------------------------------------------------

def bench_ray_fire(args, ctx, torch, stream, sh):
    from xdg_amd import capi
    from xdg_amd import geometry as G
    from xdg_amd import philox as P
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    model = G.nested_spheres(6)
    vols = list(model.volumes)
    trees = build_trees(ctx, model, vols)
    log(f"built {len(trees)} surface trees in {time.perf_counter() - t0:.2f}s: "
        + ", ".join(f"v{v}:{ctx.tree_info(t)['n_prims']}" for v, t in zip(vols, trees)))

    n = args.points
    t0 = time.perf_counter()
    pts = P.box_points(2, P.STREAM_POS, n, [-16.0] * 3, [16.0] * 3)
    dirs = P.iso_dirs(2, P.STREAM_DIR, n)
    d_pts = torch.from_numpy(pts).to(dev)
    d_dirs = torch.from_numpy(dirs).to(dev)
    log(f"generated {n} queries in {time.perf_counter() - t0:.2f}s")

    # find_volume (mixed and plain must agree)
    d_vol = torch.empty(n, dtype=torch.int32, device=dev)
    fv = lambda prec: ctx.find_volume(trees, vols, model.ipc, d_pts, d_dirs,  # noqa: E731
                                      precision=prec, out=d_vol, stream=sh)
    fv_per, _ = time_launches(torch, stream, lambda: fv(capi.MIXED), 3, 1)
    vol_mixed = d_vol.clone()
    fv(capi.PLAIN)
    torch.cuda.synchronize()
    vol_identical = bool(torch.equal(vol_mixed, d_vol))
    fv_mqs = n / (sum(fv_per) / len(fv_per)) / 1e6
    lut = torch.zeros(max(vols) + 1, dtype=torch.int32, device=dev)
    for v, t in zip(vols, trees):
        lut[v] = t
    d_tpr = lut[d_vol.long()].contiguous()
    counts = torch.bincount(d_vol.long(), minlength=max(vols) + 1).cpu().numpy()
    log(f"find_volume {fv_mqs:.1f} Mqueries/s; volumes {counts.tolist()}; "
        f"mixed==plain {vol_identical}")

    outs = {}
    for prec in (capi.MIXED, capi.PLAIN):
        outs[prec] = (torch.empty(n, dtype=torch.float64, device=dev),
                      torch.empty(n, dtype=torch.int32, device=dev),
                      torch.empty(n, dtype=torch.int32, device=dev))

    def fire(prec):
        ctx.ray_fire(0, d_pts, d_dirs, tree_per_ray=d_tpr, precision=prec,
                     out=outs[prec], stream=sh)

    per_m, _ = time_launches(torch, stream, lambda: fire(capi.MIXED), args.steps, args.warmup)
    per_p, _ = time_launches(torch, stream, lambda: fire(capi.PLAIN), args.steps, args.warmup)
    avg_m = sum(per_m) / len(per_m)
    avg_p = sum(per_p) / len(per_p)
    mixed_mrays = n / avg_m / 1e6
    plain_mrays = n / avg_p / 1e6
    tm, sm, fm = outs[capi.MIXED]
    tp, sp_, fp = outs[capi.PLAIN]
    identical = bool(torch.equal(sm, sp_) and torch.equal(fm, fp)
                     and torch.equal(tm.view(torch.int64), tp.view(torch.int64)))
    hit = sp_ >= 0
    rel = ((tm[hit] - tp[hit]).abs() / tp[hit].abs()).max().item() if hit.any() else 0.0
    log(f"ray_fire mixed {mixed_mrays:.1f} Mrays/s ({avg_m * 1e3:.3f} ms), plain "
        f"{plain_mrays:.1f} Mrays/s ({avg_p * 1e3:.3f} ms), ratio "
        f"{mixed_mrays / plain_mrays:.3f}, identical {identical}, max rel dt {rel:.3e}")

    # Algorithmic work per ray, counted on a 2^16 subsample with the plain
    # FP64 binary layout (2-child node 128 B, triangle record 80 B):
    #   bytes = 48 ray in + 16 hit out + 128 V + 80 F
    #   fp64 flops = 40 V + 9 F + 23 E + 26 C + 29 N
    ns = min(n, 1 << 16)
    sub = (d_pts[:ns].contiguous(), d_dirs[:ns].contiguous(), d_tpr[:ns].contiguous())
    ctx.set_traversal(wide=False)
    try:
        sb = ctx.ray_fire_stats(0, sub[0], sub[1], tree_per_ray=sub[2], precision=capi.PLAIN)
    finally:
        ctx.set_traversal()
    r_ = sb["rays"]
    V, F, E = sb["nodes"] / r_, sb["tri_tests"] / r_, sb["edge_tests"] / r_
    Cf, N = sb["finishes"] / r_, sb["normals"] / r_
    bytes_ray = 48 + 16 + 128 * V + 80 * F
    flops_ray = 40 * V + 9 * F + 23 * E + 26 * Cf + 29 * N
    achieved = bytes_ray * n / avg_m / 1e9
    fp64_tflops = flops_ray * n / avg_m / 1e12
    # the BVH4 the kernel actually walks, and its divergence
    s4 = ctx.ray_fire_stats(0, sub[0], sub[1], tree_per_ray=sub[2], precision=capi.MIXED)
    V4 = s4["nodes"] / s4["rays"]
    simd_node = s4["nodes"] / (64 * max(s4["wave_node_iters"], 1))
    simd_tri = s4["tri_tests"] / (64 * max(s4["wave_tri_iters"], 1))
    log(f"per ray (binary FP64 layout): V {V:.2f}, F {F:.2f}, E {E:.2f}, C {Cf:.2f}, "
        f"N {N:.2f} -> {bytes_ray:.0f} B/ray, {achieved:.0f} GB/s, FP64 "
        f"{fp64_tflops:.2f} TFLOP/s; BVH4 nodes/ray {V4:.2f}, SIMD efficiency "
        f"node loop {simd_node:.3f}, triangle loop {simd_tri:.3f}")
    roof = {"bound": "valu", "unit": "Ginst/s", "peak": round(VALU_PEAK_GINST, 1),
            "achieved": None, "frac": None, "traffic": None,
            "kernel": "k_ray_fire_pt_mixed (persistent BVH4, FP32 boxes; + deferred FP64 pass)"}
    prof = load_profile("ray_fire_mixed_nested_spheres_pmc.json")
    if prof and prof.get("queries") == n:
        # issue roof: wave-level VALU instructions per launch (SQ_INSTS_VALU,
        # committed profile of this workload) over this run's launch time
        ginst = prof["sq_insts_valu_per_launch"] / avg_m / 1e9
        traffic = prof["hbm_bytes_per_launch"]
        roof.update(achieved=round(ginst, 2), frac=round(ginst / VALU_PEAK_GINST, 4),
                    traffic=traffic, traffic_source="profiles/ray_fire_mixed_nested_spheres_pmc.json "
                    "(rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-probe scaled)",
                    formula="frac = SQ_INSTS_VALU x 4 / (1024 SIMDs x 2.4 GHz x launch time)",
                    hbm={"achieved": round(traffic / avg_m / 1e9, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(traffic / avg_m / 1e9 / HBM_PEAK_GBS, 4),
                         "note": "measured L2-to-fabric bytes per launch over this run's launch time"})
    roof.update(effective_gbs=round(achieved, 2),
                effective_note="algorithmic bytes per ray (48 + 16 + 128 V + 80 F, binary FP64 "
                               "layout) x rays / launch time; node reads are served by L2 and the "
                               "Infinity Cache, so this is not a fraction of the HBM roof",
                bytes_per_ray=round(bytes_ray, 2),
                work_per_ray={"V": round(V, 3), "F": round(F, 3), "E": round(E, 3),
                              "C": round(Cf, 3), "N": round(N, 3)},
                fp64={"achieved_tflops": round(fp64_tflops, 4), "peak_tflops": FP64_PEAK_TFLOPS,
                      "frac": round(fp64_tflops / FP64_PEAK_TFLOPS, 4)},
                bvh4_nodes_per_ray=round(V4, 3),
                simd_efficiency_node_loop=round(simd_node, 4),
                simd_efficiency_triangle_loop=round(simd_tri, 4))
    res = {
        "n": n, "avg_m": avg_m, "avg_p": avg_p, "mixed_mrays": mixed_mrays,
        "plain_mrays": plain_mrays, "identical": identical, "rel": rel,
        "fv_mqs": fv_mqs, "vol_identical": vol_identical,
        "roofline": roof,
        "model": model, "vols": vols, "pts": pts, "dirs": dirs,
        "d_vol": d_vol, "t_mixed": tm,
    }
    return res


# ---------------------------------------------------------------------------
# tet walk (tally_segments)
# ---------------------------------------------------------------------------

def bench_tet(args, ctx, torch, stream, sh):
    from xdg_amd import geometry as G
    from xdg_amd import philox as P
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    mesh = G.kuhn_cube(args.tet_n, 0.2, seed=4)
    et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
    bm = mesh.boundary_model()
    v9, f, s, fl = bm.tree_arrays(bm.ipc)
    stree = ctx.build_surface_tree(
        v9.reshape(-1, 3), np.arange(3 * v9.shape[0], dtype=np.int32).reshape(-1, 3),
        f, s, fl, bm.bump(bm.ipc))
    ctx.attach_boundary(et, stree, mesh.bface_elem)
    ctx.set_option("segment_order", args.segment_order)
    ns_ = args.segments
    p0 = torch.from_numpy(P.box_points(4, P.STREAM_POS, ns_, [0.0] * 3, [1.0] * 3)).to(dev)
    p1 = torch.from_numpy(P.box_points(4, P.STREAM_DIR, ns_, [0.0] * 3, [1.0] * 3)).to(dev)
    tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device=dev)
    cnt = torch.empty(ns_, dtype=torch.int32, device=dev)
    hsh = torch.empty(ns_, dtype=torch.int64, device=dev)
    log(f"tet mesh {mesh.tets.shape[0]} tets built in {time.perf_counter() - t0:.2f}s")

    def walk():
        ctx.tally_segments(et, p0, p1, tally=tally, out=(cnt, hsh), stream=sh)

    per_w, _ = time_launches(torch, stream, walk, max(2, args.steps // 5), 1)
    avg_w = sum(per_w) / len(per_w)
    steps_per_seg = cnt.double().mean().item()
    status = ctx.status()
    # bytes per segment = 48 + 16 + S (16 conn + 16 adj + 96 verts + 16 tally RMW)
    seg_bytes = 64 + steps_per_seg * 144
    alg_gbs = seg_bytes * ns_ / avg_w / 1e9
    tet = {"msegs_s": round(ns_ / avg_w / 1e6, 3), "ms_per_launch": round(avg_w * 1e3, 4),
           "segments": ns_, "tets": int(mesh.tets.shape[0]),
           "elements_per_segment": round(steps_per_seg, 3),
           "element_steps_per_s": ns_ * steps_per_seg / avg_w,
           "capped": bool(status & 1)}
    roof = {"bound": "hbm", "achieved": round(alg_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg_gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_segments (tallies)",
            "bytes_per_segment": round(seg_bytes, 2),
            "model": "64 + S x (128 element record + 16 tally read-modify-write) bytes per segment"}
    prof = load_profile("tet_walk_kuhn64_counters.json")
    if prof and prof.get("segments") == ns_:
        tr = prof["fetch_bytes_per_launch"] + prof["write_bytes_per_launch"]
        roof.update(traffic=tr, traffic_source="profiles/tet_walk_kuhn64_counters.json "
                    "(rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-probe scaled)",
                    traffic_over_algorithmic=round(tr / (seg_bytes * ns_), 3))
    # The walk's binding ceiling is the rate of random-address int64 atomics
    # (one tally update per element step), which no published figure gives:
    # measured here on an array of the tally's size with torch's index_add_
    # (one no-return 64-bit atomic per index, uniform random indices).
    na = 1 << 27
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    aidx = torch.randint(0, tally.numel(), (na,), device=dev, generator=gen)
    asrc = torch.ones(na, dtype=torch.int64, device=dev)
    atally = torch.zeros_like(tally)

    def atomics():
        with torch.cuda.stream(stream):
            atally.index_add_(0, aidx, asrc)

    per_a, _ = time_launches(torch, stream, atomics, 3, 1)
    ceil = na / (sum(per_a) / len(per_a)) / 1e9
    steps_s = ns_ * steps_per_seg / avg_w / 1e9
    roof["atomics"] = {"achieved": round(steps_s, 2), "ceiling": round(ceil, 2), "unit": "G int64 atomics/s",
                       "frac": round(steps_s / ceil, 4),
                       "ceiling_source": "measured in this run: index_add_ of 2^27 ones at uniform random "
                                         "indices into an int64 array of the tally's size"}
    roof["binding"] = "atomics"
    del aidx, asrc, atally
    tet["roofline"] = roof
    tet["_p"] = (p0, p1, cnt, hsh, mesh)
    log(f"tet walk {tet['msegs_s']:.2f} Msegs/s, {steps_per_seg:.1f} elements/segment")

    # element-by-element particle walk (tools/walk_elements.h) on the same
    # mesh: 2^20 particles, mean free path 0.1, isotropic scattering
    npart, mfp = 1 << 20, 0.1
    mw_tally = torch.zeros(mesh.tets.shape[0], dtype=torch.int64, device=dev)
    mw_out = {}

    def mwalk():
        mw_out["r"] = ctx.mesh_walk(et, 0, npart, 6, 0, [0.0] * 3, [1.0] * 3, mfp, tally=mw_tally,
                                    stream=sh)

    per_m, _ = time_launches(torch, stream, mwalk, max(2, args.steps // 5), 1)
    avg_m = sum(per_m) / len(per_m)
    dist, ev, st, _ = mw_out["r"]
    events = ev.double().sum().item()
    tet["mesh_walk"] = {"mparticles_s": round(npart / avg_m / 1e6, 3),
                        "ms_per_launch": round(avg_m * 1e3, 4), "particles": npart,
                        "mfp": mfp, "events_per_particle": round(events / npart, 3),
                        "events_per_s": events / avg_m,
                        "left_mesh": int((st == 2).sum().item()),
                        "mean_distance": round(dist.mean().item(), 6)}
    log(f"mesh walk {tet['mesh_walk']['mparticles_s']:.2f} Mparticles/s, "
        f"{events / npart:.1f} events/particle")
    return tet


# ---------------------------------------------------------------------------
# single sphere: the reference's own CPU-runnable ray_fire case
# ---------------------------------------------------------------------------

def bench_sphere(args, ctx, torch, stream, sh):
    """Icosphere r = 10, subdivision 5 (20 480 triangles); 1 M rays with
    origins uniform in the ball r = 9 and isotropic directions (Philox seed
    1), EXITING: ray_fire mixed and plain, as tools/ray_benchmark.cpp:205-225
    times its trace phase."""
    from xdg_amd import capi
    from xdg_amd import geometry as G
    from xdg_amd import philox as P
    dev = torch.device("cuda", 0)
    m = G.single_sphere(5, 10.0)
    tree = build_trees(ctx, m, [1])[0]
    n = args.sphere_rays
    org = P.ball_points(1, P.STREAM_POS, n, 9.0)
    dirs = P.iso_dirs(1, P.STREAM_DIR, n)
    d_org = torch.from_numpy(org).to(dev)
    d_dir = torch.from_numpy(dirs).to(dev)
    outs = {p: (torch.empty(n, dtype=torch.float64, device=dev),
                torch.empty(n, dtype=torch.int32, device=dev),
                torch.empty(n, dtype=torch.int32, device=dev)) for p in (capi.MIXED, capi.PLAIN)}
    per = {}
    for prec in (capi.MIXED, capi.PLAIN):
        fn = lambda: ctx.ray_fire(tree, d_org, d_dir, precision=prec, out=outs[prec], stream=sh)  # noqa: E731
        per[prec], _ = time_launches(torch, stream, fn, args.steps, args.warmup)
    avg = {p: sum(v) / len(v) for p, v in per.items()}
    tm, sm, fm = outs[capi.MIXED]
    tp, sp_, fp = outs[capi.PLAIN]
    identical = bool(torch.equal(sm, sp_) and torch.equal(fm, fp)
                     and torch.equal(tm.view(torch.int64), tp.view(torch.int64)))
    res = {"workload": "single_sphere_ray_fire", "geometry": "icosphere r=10 subdiv 5 (20480 tris)",
           "rays": n, "seed": 1, "mixed_mrays_s": round(n / avg[capi.MIXED] / 1e6, 3),
           "plain_fp64_mrays_s": round(n / avg[capi.PLAIN] / 1e6, 3),
           "mixed_ms": round(avg[capi.MIXED] * 1e3, 4), "plain_ms": round(avg[capi.PLAIN] * 1e3, 4),
           "mixed_over_plain": round(avg[capi.PLAIN] / avg[capi.MIXED], 4),
           "identical_hits": identical, "hits": int((sp_ >= 0).sum().item())}
    log(f"single sphere: mixed {res['mixed_mrays_s']:.1f} Mrays/s, plain "
        f"{res['plain_fp64_mrays_s']:.1f} Mrays/s, identical {identical}")
    res["_p"] = (m, org, dirs, tm)
    return res


# ---------------------------------------------------------------------------
# particle transport: one captured graph launch over all histories
# ---------------------------------------------------------------------------

def bench_transport(args, ctx, torch, stream, sh):
    from xdg_amd import capi
    from xdg_amd import geometry as G
    t0 = time.perf_counter()
    m = G.sphere_lattice(8, 4, 0.8, 2.0)
    vols = [v for v in m.volumes if v != m.ipc]
    trees = build_trees(ctx, m, vols)
    log(f"lattice: {len(trees)} volume trees built in {time.perf_counter() - t0:.2f}s")
    h = ctx.transport_create(vols, trees, m.ipc, m.senses, 3, 3, [0.0] * 3, [16.0] * 3,
                             2.0, 0.3, 100000)
    n = args.histories
    nvol = len(vols)
    t0 = time.perf_counter()
    ns = min(n, 1 << 18)
    if args.transport_config:
        ctx.transport_set_config(h, *[int(x) for x in args.transport_config.split(",")])
        cfg = ctx.transport_config(h)
        log(f"transport config {cfg} (given)")
    else:
        tune_ms = ctx.transport_autotune(h, 0, ns, capi.MIXED)
        cfg = ctx.transport_config(h)
        log(f"autotune {time.perf_counter() - t0:.1f}s -> {cfg} ({tune_ms:.3f} ms / {ns} histories)")
    out = ctx.transport_outputs(nvol, n)
    g = ctx.transport_capture(h, 0, n, out, capi.MIXED)
    try:
        per, _ = time_launches(torch, stream, lambda: ctx.graph_launch(g, sh),
                               max(2, args.steps // 5), 1)
    finally:
        torch.cuda.synchronize()
        ctx.graph_destroy(g)
    avg = sum(per) / len(per)
    ev = out["events"].long().sum().item()
    st = torch.bincount(out["status"].long(), minlength=7).cpu().numpy()
    tally = out["tally_fx"].cpu().numpy().astype(np.float64) / 2.0 ** 32
    ctx.transport_destroy(h)
    res = {"mhistories_s": round(n / avg / 1e6, 4), "ms_per_launch": round(avg * 1e3, 3),
           "histories": n, "events_per_s": ev / avg, "events_per_history": ev / n,
           "absorbed": int(st[1]), "leaked": int(st[2]),
           "lost_stuck_capped_error": int(st[3:].sum()),
           "total_track_length": float(tally.sum()), "config": cfg,
           "launch": "captured hipGraph: start stage (sample, sort, candidate volumes, "
                     "point_in_volume items), then the persistent event kernel"}
    log(f"transport {res['mhistories_s']:.3f} Mhistories/s, {ev / avg / 1e6:.1f} Mevents/s, "
        f"status {st.tolist()}")
    return res


# ---------------------------------------------------------------------------
# ray-distance adjoint
# ---------------------------------------------------------------------------

def bench_adjoint(args, ctx, torch, stream, sh):
    from xdg_amd import capi
    from xdg_amd import geometry as G
    from xdg_amd import philox as P
    from xdg_amd import adjoint as A
    dev = torch.device("cuda", 0)
    m = G.perturbed_sphere(6, 10.0, 0.05)
    n = args.adjoint_rays
    org = torch.from_numpy(P.ball_points(5, P.STREAM_POS, n, 9.5)).to(dev)
    dirs = torch.from_numpy(P.iso_dirs(5, P.STREAM_DIR, n)).to(dev)
    w = torch.from_numpy(P.normals(5, P.STREAM_AUX, n)).to(dev)
    surf = A.DifferentiableSurface(ctx, m.tris)
    verts = torch.from_numpy(m.verts).to(dev)
    tree = surf.tree_for(verts)
    t, s, f = ctx.ray_fire(tree, org, dirs, precision=capi.MIXED)
    gv = torch.zeros_like(verts)
    go = torch.empty_like(org)
    gd = torch.empty_like(dirs)

    def adj():
        with torch.cuda.stream(stream):  # the zeroing is ordered before the atomics
            gv.zero_()
        ctx.ray_distance_adjoint(org, dirs, t, f, w, verts, surf.tris, gv, go, gd, stream=sh)

    per, _ = time_launches(torch, stream, adj, max(2, args.steps // 2), 1)
    avg = sum(per) / len(per)
    # bytes per ray: 48 ray + 8 t + 4 face + 8 dL/dt + 12 ids + 72 verts
    # + 9 x 16 atomic RMW + 48 dL/d(o, d)
    b = 48 + 8 + 4 + 8 + 12 + 72 + 144 + 48
    res = {"mrays_s": round(n / avg / 1e6, 3), "ms_per_launch": round(avg * 1e3, 4),
           "rays": n, "algorithmic_gbs": round(b * n / avg / 1e9, 2),
           "grad_norm": float(gv.norm().item())}
    # one step of a finite-difference or optimiser loop on the 81 920-triangle
    # surface: the tree follows moved vertices by a device refit (the
    # harness's path) instead of a rebuild from host arrays
    moved = verts * (1.0 + 1e-6)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for r in range(reps):
        surf.tree_for(moved if r % 2 == 0 else verts)
    torch.cuda.synchronize()
    refit_ms = (time.perf_counter() - t0) / reps * 1e3
    vh = m.verts.copy()
    nt = m.tris.shape[0]
    t0 = time.perf_counter()
    for _ in range(reps):
        tr = ctx.build_surface_tree(vh, m.tris.astype(np.int32), np.arange(nt, dtype=np.int32),
                                    np.ones(nt, np.int32), np.zeros(nt, np.uint8), G.vertex_bump(vh))
        ctx.release_surface_tree(tr)
    rebuild_ms = (time.perf_counter() - t0) / reps * 1e3
    res.update(refit_ms=round(refit_ms, 3), rebuild_ms=round(rebuild_ms, 3))
    log(f"adjoint {res['mrays_s']:.1f} Mrays/s ({res['ms_per_launch']:.3f} ms); tree refit {refit_ms:.2f} ms, "
        f"rebuild {rebuild_ms:.2f} ms")
    return res


# ---------------------------------------------------------------------------
# CPU baseline: the oracle's BVH path on this host's cores
# ---------------------------------------------------------------------------

def cpu_baseline(args, hl):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    model, vols, pts, dirs = hl["model"], hl["vols"], hl["pts"], hl["dirs"]
    n = hl["n"]
    ncpu = len(os.sched_getaffinity(0))
    m = min(args.cpu_sample, n)
    vol_h = hl["d_vol"][:m].cpu().numpy()
    otrees = {v: O.Tree.from_model(model, v) for v in vols}
    groups = [(v, np.nonzero(vol_h == v)[0]) for v in vols]
    sub = [(v, idx, np.ascontiguousarray(pts[idx]), np.ascontiguousarray(dirs[idx]))
           for v, idx in groups if idx.size]
    res_t = np.empty(m)

    def trace():
        for v, idx, p_, d_ in sub:
            t_, s_, f_ = otrees[v].ray_fire(p_, d_, nthreads=ncpu)
            res_t[idx] = t_

    trace()  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        trace()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    agree = bool(np.array_equal(res_t.view(np.uint64),
                                hl["t_mixed"][:m].cpu().numpy().view(np.uint64)))
    cpu = {"value": round(m / dt / 1e6, 3), "unit": "Mrays/s", "cores": ncpu,
           "kind": "port",
           "sample": f"{m} of the {n} nested-sphere queries, ray_fire in their "
                     f"containing volume with the C oracle (FP64 host BVH, "
                     f"OpenMP dynamic schedule); warm-up then median of 5 "
                     f"passes, {dt:.2f} s each, {sum(times) + dt:.1f} s in all; "
                     f"distances bit-identical to the GPU: {agree}"}
    log(f"cpu baseline {cpu['value']:.2f} Mrays/s on {ncpu} threads ({dt:.2f}s per pass)")
    return cpu


def _host_info():
    ncpu = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ncpu, model


def _median_passes(fn, passes=5):
    fn()  # warm-up
    times = []
    for _ in range(passes):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), sum(times)


def cpu_baseline_sphere(sp):
    """The oracle's ray_fire over the same 1 M single-sphere rays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    m, org, dirs, t_gpu = sp["_p"]
    ncpu, cpu_model = _host_info()
    ot = O.Tree.from_model(m, 1)
    res = {}
    dt, tot = _median_passes(lambda: res.update(r=ot.ray_fire(org, dirs, nthreads=ncpu)))
    agree = bool(np.array_equal(res["r"][0].view(np.uint64), t_gpu.cpu().numpy().view(np.uint64)))
    n = org.shape[0]
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": ncpu, "cpu": cpu_model,
            "kind": "port",
            "sample": f"all {n} single-sphere rays, C oracle (FP64 host BVH, OpenMP dynamic "
                      f"schedule); warm-up then median of 5 passes ({dt:.3f} s each); "
                      f"distances bit-identical to the GPU: {agree}"}


def cpu_baseline_tet(args, tet):
    """The oracle's XDG::segments walk with per-element tallies (the loop
    tools/tally_segments.h:49-88 times) over the first segments of the same
    batch, on every host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    p0, p1, cnt, hsh, mesh = tet["_p"]
    ncpu, cpu_model = _host_info()
    m = min(args.cpu_tet_sample, p0.shape[0])
    a = p0[:m].cpu().numpy()
    b = p1[:m].cpu().numpy()
    om = O.TetMesh(mesh)
    res = {}
    dt, tot = _median_passes(lambda: res.update(r=om.segments(a, b, tally=True, nthreads=ncpu)))
    rc, rh, _ = res["r"]
    agree = bool(np.array_equal(rc, cnt[:m].cpu().numpy())
                 and np.array_equal(rh, hsh[:m].cpu().numpy().view(np.uint64)))
    return {"value": round(m / dt / 1e6, 4), "unit": "Msegs/s", "cores": ncpu, "cpu": cpu_model,
            "kind": "port",
            "sample": f"first {m} of the {p0.shape[0]} segments on the 64^3 x 6 Kuhn mesh, C oracle "
                      f"XDG::segments + tallies (host element BVH, OpenMP dynamic schedule); "
                      f"warm-up then median of 5 passes ({dt:.2f} s each, {tot + dt:.1f} s in all); "
                      f"element sequences (counts and hashes) identical to the GPU: {agree}"}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: this build runs the hot path on a "
              "single GPU only (multi-GPU sharding is a design note, not a "
              f"measured path); got --gpus {args.gpus}, WORLD_SIZE {world}.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from xdg_amd import capi

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    ctx = capi.Context(0)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    sh = capi._vp(stream.cuda_stream)

    if args.copy_probe:
        # a copy of known size: 256 MiB read + 256 MiB written per launch
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.legs:
        legs = {"ray_fire": bench_ray_fire, "sphere": bench_sphere, "tet": bench_tet,
                "transport": bench_transport, "adjoint": bench_adjoint}
        res = {}
        for name in args.legs.split(","):
            r = legs[name](args, ctx, torch, stream, sh)
            res[name] = {k: v for k, v in r.items() if not k.startswith("_") and k not in ("model", "pts", "dirs")}
        print(json.dumps({"legs": res}, default=str), flush=True)
        ctx.close()
        return

    hl = bench_ray_fire(args, ctx, torch, stream, sh)
    sph = None if args.no_sphere else bench_sphere(args, ctx, torch, stream, sh)
    tet = None if args.no_tet else bench_tet(args, ctx, torch, stream, sh)
    tr = None if args.no_transport else bench_transport(args, ctx, torch, stream, sh)
    adj = None if args.no_adjoint else bench_adjoint(args, ctx, torch, stream, sh)
    cpu = None if args.no_cpu else cpu_baseline(args, hl)
    if cpu is not None:
        ncpu, cpu_model = _host_info()
        cpu["cpu"] = cpu_model
        if sph is not None:
            cpu["single_sphere"] = cpu_baseline_sphere(sph)
            log(f"cpu baseline single sphere {cpu['single_sphere']['value']:.2f} Mrays/s")
        if tet is not None:
            cpu["tet_walk"] = cpu_baseline_tet(args, tet)
            log(f"cpu baseline tet walk {cpu['tet_walk']['value']:.4f} Msegs/s")
    for leg in (sph, tet):
        if leg is not None:
            leg.pop("_p", None)

    n = hl["n"]
    line = {
        "metric": METRIC,
        "value": round(hl["mixed_mrays"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hl["avg_m"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "nested_spheres_ray_fire",
                   "geometry": "icospheres r=5/10/15 subdiv 6 (81920 tris each) + cube [-20,20]^3",
                   "queries": n, "query_box": [-16, 16], "seed": 2,
                   "orientation": "EXITING", "box_precision": "f32 outward-rounded",
                   "triangle_test": "f64 Pluecker", "parallelism": "single GPU"},
        "plain_fp64_mrays_s": round(hl["plain_mrays"], 3),
        "mixed_over_plain": round(hl["mixed_mrays"] / hl["plain_mrays"], 4),
        "identical_hits": hl["identical"],
        "max_rel_distance_error": hl["rel"],
        "find_volume_mqueries_s": round(hl["fv_mqs"], 3),
        "find_volume_mixed_equals_plain": hl["vol_identical"],
        "single_sphere": sph,
        "tet_walk": tet,
        "transport": tr,
        "adjoint": adj,
        "roofline": hl["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
