"""Philox4x32-10 counter-based RNG and the sampling protocols shared by the
host input generators, the device transport kernel and the CPU oracle.

Philox4x32-10 follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers:
as easy as 1, 2, 3" (SC'11).  Block = philox(counter, key) with
counter = (item_lo, item_hi, draw, 0) and key = (seed, stream); one block
gives two 53-bit uniforms u = (w >> 11) * 2^-53 from the 64-bit words
(x0<<32|x1) and (x2<<32|x3).

Sampling protocols (identical on host, device and oracle):
  iso_dir   Marsaglia (1972) rejection as in tools/ray_benchmark.h:24-40:
            attempt j uses block draw = base + j; x1 = ua*2-1, x2 = ub*2-1,
            s = x1*x1 + x2*x2, reject s <= 0 or s >= 1,
            dir = (x1*t, x2*t, 1 - 2*s) with t = 2*sqrt(1 - s).
  ball      rejection in the cube [-R,R]^3: attempt j uses blocks 2j, 2j+1
            (x = ua*2-1, y = ub*2-1 from the first, z = ua'*2-1 from the
            second), accept x*x+y*y+z*z < 1, point = R*(x,y,z).
  box       blocks 0 and 1: p = lo + (hi - lo) * (ua, ub, ua').
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)
_TWO_M53 = 2.0 ** -53

# stream tags (key[1]) for the fixed-shape workloads
STREAM_POS = 0
STREAM_DIR = 1
STREAM_AUX = 2


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 over uint32 arrays (broadcasting)."""
    c0 = np.asarray(c0, dtype=np.uint32)
    c1 = np.asarray(c1, dtype=np.uint32)
    c2 = np.asarray(c2, dtype=np.uint32)
    c3 = np.asarray(c3, dtype=np.uint32)
    k0 = np.asarray(k0, dtype=np.uint32)
    k1 = np.asarray(k1, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = k0 + W0
                k1 = k1 + W1
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n1 = (p1 & _MASK32).astype(np.uint32)
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            n3 = (p0 & _MASK32).astype(np.uint32)
            c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def uniform2(seed: int, stream: int, item, draw):
    """Two uniforms in [0,1) per (item, draw)."""
    item = np.asarray(item, dtype=np.uint64)
    lo = (item & _MASK32).astype(np.uint32)
    hi = (item >> np.uint64(32)).astype(np.uint32)
    o0, o1, o2, o3 = philox4x32_10(lo, hi, np.asarray(draw, dtype=np.uint32),
                                   np.uint32(0), np.uint32(seed),
                                   np.uint32(stream))
    a = (o0.astype(np.uint64) << np.uint64(32)) | o1.astype(np.uint64)
    b = (o2.astype(np.uint64) << np.uint64(32)) | o3.astype(np.uint64)
    ua = (a >> np.uint64(11)).astype(np.float64) * _TWO_M53
    ub = (b >> np.uint64(11)).astype(np.float64) * _TWO_M53
    return ua, ub


def iso_dirs(seed: int, stream: int, n: int, base: int = 0,
             first_item: int = 0, max_attempts: int = 64):
    """Isotropic unit directions (Marsaglia rejection), float64 (n, 3)."""
    items = np.arange(first_item, first_item + n, dtype=np.uint64)
    out = np.empty((n, 3), dtype=np.float64)
    todo = np.arange(n)
    for j in range(max_attempts):
        if todo.size == 0:
            break
        ua, ub = uniform2(seed, stream, items[todo], base + j)
        x1 = ua * 2.0 - 1.0
        x2 = ub * 2.0 - 1.0
        s = x1 * x1 + x2 * x2
        ok = (s > 0.0) & (s < 1.0)
        idx = todo[ok]
        t = 2.0 * np.sqrt(1.0 - s[ok])
        out[idx, 0] = x1[ok] * t
        out[idx, 1] = x2[ok] * t
        out[idx, 2] = 1.0 - 2.0 * s[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("iso_dirs: rejection did not converge")
    return out


def ball_points(seed: int, stream: int, n: int, radius: float,
                center=(0.0, 0.0, 0.0), first_item: int = 0,
                max_attempts: int = 64):
    """Uniform points strictly inside a ball (cube rejection), (n, 3)."""
    items = np.arange(first_item, first_item + n, dtype=np.uint64)
    out = np.empty((n, 3), dtype=np.float64)
    todo = np.arange(n)
    c = np.asarray(center, dtype=np.float64)
    for j in range(max_attempts):
        if todo.size == 0:
            break
        ua, ub = uniform2(seed, stream, items[todo], 2 * j)
        uc, _ = uniform2(seed, stream, items[todo], 2 * j + 1)
        x = ua * 2.0 - 1.0
        y = ub * 2.0 - 1.0
        z = uc * 2.0 - 1.0
        ok = (x * x + y * y + z * z) < 1.0
        idx = todo[ok]
        out[idx, 0] = c[0] + radius * x[ok]
        out[idx, 1] = c[1] + radius * y[ok]
        out[idx, 2] = c[2] + radius * z[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("ball_points: rejection did not converge")
    return out


def box_points(seed: int, stream: int, n: int, lo, hi, first_item: int = 0):
    """Uniform points in an axis-aligned box, (n, 3)."""
    items = np.arange(first_item, first_item + n, dtype=np.uint64)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    ua, ub = uniform2(seed, stream, items, 0)
    uc, _ = uniform2(seed, stream, items, 1)
    w = hi - lo
    return np.stack([lo[0] + w[0] * ua, lo[1] + w[1] * ub, lo[2] + w[2] * uc],
                    axis=1)


def normals(seed: int, stream: int, n: int, first_item: int = 0):
    """Standard normal weights by Box-Muller (host-only input data)."""
    items = np.arange(first_item, first_item + n, dtype=np.uint64)
    ua, ub = uniform2(seed, stream, items, 0)
    return np.sqrt(-2.0 * np.log1p(-ua)) * np.cos(2.0 * np.pi * ub)
