"""xdg_hip_ray_fire_noalloc against xdg_hip_ray_fire through the C-ABI.

The allocation-free entry point keeps its scratch (work counter, deferral
list) in one array per GPU that persists between calls, so two calls in a
row, the second smaller, must each equal xdg_hip_ray_fire bit for bit (no
state left over from the first).  Far origins exercise the deferral list;
both precisions, per-ray trees and an oversized batch (rejected before any
launch) are covered.
"""
import numpy as np
import pytest

from xdg_amd import geometry as G
from xdg_amd import philox as P

from gpu_util import bits, build_tree, dev, host

pytestmark = pytest.mark.gpu


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


def _rays(seed, n, lim):
    org = P.ball_points(seed, P.STREAM_POS, n, 14.0)
    dirs = P.iso_dirs(seed, P.STREAM_DIR, n)
    far = np.arange(n) % 5 == 0  # beyond o_limit: the FP64 deferral pass
    org[far] = -dirs[far] * (3.0 * lim) + 0.01 * org[far]
    return org, dirs


def _empty(torch, n):
    return (torch.full((n,), -7.0, dtype=torch.float64, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda"))


def test_noalloc_matches_ray_fire_twice(ctx, torch_mod):
    torch = torch_mod
    m = G.nested_spheres(4)
    vols = list(m.volumes)
    trees = [build_tree(ctx, m, v) for v in vols]
    lim = min(ctx.tree_info(t)["o_limit"] for t in trees)
    for prec in (1, 0):
        for n, seed in ((300000, 31), (70001, 32)):  # the second call smaller
            org, dirs = _rays(seed, n, lim)
            d_org, d_dir = dev(torch, org), dev(torch, dirs)
            tpr = dev(torch, np.array(trees, np.int32)[np.arange(n) % len(trees)])
            for tree, per_ray in ((trees[1], None), (0, tpr)):
                ref = ctx.ray_fire(tree, d_org, d_dir, precision=prec, tree_per_ray=per_ray)
                got = ctx.ray_fire_noalloc(tree, d_org, d_dir, _empty(torch, n), precision=prec,
                                           tree_per_ray=per_ray)
                torch.cuda.synchronize()
                assert np.array_equal(host(got[1]), host(ref[1])), (prec, n)
                assert np.array_equal(host(got[2]), host(ref[2])), (prec, n)
                assert np.array_equal(bits(host(got[0])), bits(host(ref[0]))), (prec, n)
                assert (host(ref[1]) >= 0).mean() > 0.5


def test_noalloc_orientations_and_exclusions(ctx, torch_mod):
    torch = torch_mod
    m = G.single_sphere(4, 10.0)
    tree = build_tree(ctx, m, 1)
    n = 5000
    org = P.ball_points(5, 0, n, 9.0)
    dirs = P.iso_dirs(5, 1, n)
    excl = np.tile(np.arange(4, dtype=np.int32), (n, 1))
    for orientation in (-1, 0, 1):
        for prec in (0, 1):
            kw = dict(orientation=orientation, precision=prec, excl=dev(torch, excl))
            ref = ctx.ray_fire(tree, dev(torch, org), dev(torch, dirs), **kw)
            got = ctx.ray_fire_noalloc(tree, dev(torch, org), dev(torch, dirs), _empty(torch, n), **kw)
            torch.cuda.synchronize()
            for a, b in zip(got, ref):
                assert np.array_equal(host(a).view(np.uint8), host(b).view(np.uint8))


def test_noalloc_rejects_oversized_batch(ctx, torch_mod):
    from xdg_amd import capi
    torch = torch_mod
    m = G.single_sphere(2, 10.0)
    tree = build_tree(ctx, m, 1)
    n = capi.NOALLOC_MAX_RAYS + 1
    org = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dirs = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dirs[:, 0] = 1.0
    out = _empty(torch, n)
    with pytest.raises(capi.XdgHipError) as e:
        ctx.ray_fire_noalloc(tree, org, dirs, out)
    assert e.value.code == capi.ERR_INVALID and "XDG_HIP_NOALLOC_MAX_RAYS" in str(e.value)
    assert (out[1] == -7).all().item()  # nothing was launched
