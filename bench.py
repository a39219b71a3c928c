"""Benchmark of the geometry-query hot path on one MI355X.

Workload (the metric's configuration): three nested icospheres r = 5/10/15
(subdivision 6, 81 920 triangles each) inside the cube [-20, 20]^3; 10 M
query points uniform in [-16, 16]^3 with isotropic directions (Philox seed 2).
Every point is located with find_volume, then one step = ray_fire of all
points against the surface tree of the volume that contains them
(EXITING, distance and surface/face ids), on device-resident inputs.

Reported: the mixed FP32-box / FP64-Pluecker kernel's Mrays/s (value), the
plain FP64 kernel's Mrays/s, whether the two produce identical hits,
find_volume Mqueries/s, the tet-walk Msegs/s (64^3 x 6 jittered Kuhn tets,
4 M segments), a roofline object for the dominant kernel and the CPU oracle
baseline timed on this host.

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
    p.add_argument("--cpu-sample", type=int, default=2_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tet", action="store_true")
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
    from xdg_amd import geometry as G
    from xdg_amd import philox as P

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

    # ---- geometry and trees ------------------------------------------------
    t0 = time.perf_counter()
    model = G.nested_spheres(6)
    vols = list(model.volumes)
    trees = []
    for v in vols:
        v9, f, s, fl = model.tree_arrays(v)
        nt = v9.shape[0]
        trees.append(ctx.build_surface_tree(
            v9.reshape(-1, 3), np.arange(3 * nt, dtype=np.int32).reshape(-1, 3),
            f, s, fl, model.bump(v)))
    log(f"built {len(trees)} surface trees in {time.perf_counter() - t0:.2f}s: "
        + ", ".join(f"v{v}:{ctx.tree_info(t)['n_prims']}/d{ctx.tree_info(t)['depth']}"
                    for v, t in zip(vols, trees)))

    n = args.points
    t0 = time.perf_counter()
    pts = P.box_points(2, P.STREAM_POS, n, [-16.0] * 3, [16.0] * 3)
    dirs = P.iso_dirs(2, P.STREAM_DIR, n)
    d_pts = torch.from_numpy(pts).to(dev)
    d_dirs = torch.from_numpy(dirs).to(dev)
    log(f"generated {n} queries in {time.perf_counter() - t0:.2f}s")

    # ---- find_volume (mixed and plain must agree) ---------------------------
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

    # ---- ray_fire: mixed (value) and plain ---------------------------------
    outs = {}
    for prec in (capi.MIXED, capi.PLAIN):
        outs[prec] = (torch.empty(n, dtype=torch.float64, device=dev),
                      torch.empty(n, dtype=torch.int32, device=dev),
                      torch.empty(n, dtype=torch.int32, device=dev))

    def fire(prec):
        ctx.ray_fire(0, d_pts, d_dirs, tree_per_ray=d_tpr, precision=prec,
                     out=outs[prec], stream=sh)

    per_m, wall_m = time_launches(torch, stream, lambda: fire(capi.MIXED),
                                  args.steps, args.warmup)
    per_p, wall_p = time_launches(torch, stream, lambda: fire(capi.PLAIN),
                                  args.steps, args.warmup)
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

    # ---- roofline accounting from the instrumented traversal ---------------
    ns = min(n, 1 << 16)
    st = ctx.ray_fire_stats(0, d_pts[:ns].contiguous(), d_dirs[:ns].contiguous(),
                            tree_per_ray=d_tpr[:ns].contiguous(),
                            precision=capi.MIXED)
    V = st["nodes"] / st["rays"]
    F = st["tri_tests"] / st["rays"]
    E = st["edge_tests"] / st["rays"]
    Cf = st["finishes"] / st["rays"]
    N = st["normals"] / st["rays"]
    # bytes per ray: origin+direction 48, tree id 4, t/surf/face 16;
    # mixed node 64 B (two FP32 child boxes + ids), triangle record 80 B
    bytes_ray = 48 + 4 + 16 + 64 * V + 80 * F
    flops_ray = 9 * F + 23 * E + 26 * Cf + 29 * N  # FP64 (boxes are FP32)
    achieved = bytes_ray * n / avg_m / 1e9
    log(f"per ray: nodes {V:.2f}, tri tests {F:.2f}, edge tests {E:.2f}, "
        f"finishes {Cf:.2f}, normals {N:.2f}, fallbacks {st['fallbacks']}; "
        f"{bytes_ray:.0f} B/ray -> {achieved:.0f} GB/s; "
        f"FP64 {flops_ray * n / avg_m / 1e12:.2f} TFLOP/s")
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "ray_fire_mixed_nested_spheres_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # ---- tet walk -------------------------------------------------------------
    tet = None
    if not args.no_tet:
        t0 = time.perf_counter()
        mesh = G.kuhn_cube(args.tet_n, 0.2, seed=4)
        et = ctx.build_element_tree(mesh.verts, mesh.tets, mesh.adj, mesh.face_order)
        bm = mesh.boundary_model()
        v9, f, s, fl = bm.tree_arrays(bm.ipc)
        stree = ctx.build_surface_tree(
            v9.reshape(-1, 3), np.arange(3 * v9.shape[0], dtype=np.int32).reshape(-1, 3),
            f, s, fl, bm.bump(bm.ipc))
        ctx.attach_boundary(et, stree, mesh.bface_elem)
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
        tet = {"msegs_s": ns_ / avg_w / 1e6, "ms_per_launch": avg_w * 1e3,
               "segments": ns_, "tets": int(mesh.tets.shape[0]),
               "elements_per_segment": steps_per_seg,
               "element_steps_per_s": ns_ * steps_per_seg / avg_w,
               "capped": bool(status & 1)}
        log(f"tet walk {tet['msegs_s']:.2f} Msegs/s, {steps_per_seg:.1f} elements/segment")

    # ---- CPU baseline: the oracle's BVH path on this host's cores ----------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        ncpu = len(os.sched_getaffinity(0))
        m = min(args.cpu_sample, n)
        vol_h = d_vol[:m].cpu().numpy()
        otrees = {v: O.Tree.from_model(model, v) for v in vols if v != model.ipc}
        otrees[model.ipc] = O.Tree.from_model(model, model.ipc)
        groups = [(v, np.nonzero(vol_h == v)[0]) for v in vols]
        t0 = time.perf_counter()
        res_t = np.empty(m)
        for v, idx in groups:
            if idx.size:
                t_, s_, f_ = otrees[v].ray_fire(pts[idx], dirs[idx], nthreads=ncpu)
                res_t[idx] = t_
        dt = time.perf_counter() - t0
        agree = bool(np.array_equal(res_t.view(np.uint64),
                                    tm[:m].cpu().numpy().view(np.uint64)))
        cpu = {"value": m / dt / 1e6, "unit": "Mrays/s", "cores": ncpu,
               "kind": "port",
               "sample": f"first {m} of the {n} nested-sphere queries, ray_fire "
                         f"in their containing volume (FP64 host BVH, OpenMP), "
                         f"{dt:.1f} s; distances bit-identical to the GPU: {agree}"}
        log(f"cpu baseline {cpu['value']:.2f} Mrays/s on {ncpu} threads ({dt:.1f}s)")

    line = {
        "metric": METRIC,
        "value": round(mixed_mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(avg_m * 1e3, 4),
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
        "plain_fp64_mrays_s": round(plain_mrays, 3),
        "mixed_over_plain": round(mixed_mrays / plain_mrays, 4),
        "identical_hits": identical,
        "max_rel_distance_error": rel,
        "find_volume_mqueries_s": round(fv_mqs, 3),
        "find_volume_mixed_equals_plain": vol_identical,
        "tet_walk": tet,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_ray": round(bytes_ray, 2),
                     "nodes_per_ray": round(V, 3), "tri_tests_per_ray": round(F, 3),
                     "fp64_tflops": round(flops_ray * n / avg_m / 1e12, 4),
                     "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
