"""Seeded synthetic workloads for the UMI operator configurations.

All generators are deterministic functions of ``seed`` (numpy PCG64).  They
produce the inputs only; results come from the library (and, in tests, from
the oracle).
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def box5_u8(img: np.ndarray) -> np.ndarray:
    """5x5 box filter with edge replication, integer rounding (sum + 12) // 25."""
    p = np.pad(img.astype(np.int32), 2, mode="edge")
    H, W = img.shape
    acc = np.zeros((H, W), np.int32)
    for dy in range(5):
        for dx in range(5):
            acc += p[dy:dy + H, dx:dx + W]
    return ((acc + 12) // 25).astype(np.uint8)


def smooth_noise_u8(H: int, W: int, g: np.random.Generator) -> np.ndarray:
    return box5_u8(g.integers(0, 256, size=(H, W), dtype=np.int32).astype(np.uint8))


def me_pair(H: int = 1080, W: int = 1920, R: int = 16, tile: int = 16, sigma: float = 2.0, seed: int = 1):
    """Frame pair for full-search ME.

    frame0 = uniform noise smoothed by a 5x5 box; frame1(y, x) =
    clip(round(frame0[cl(y+my)][cl(x+mx)] + N(0, sigma^2))) with (my, mx)
    uniform in [-R, R]^2, constant over each ``tile`` x ``tile`` block.
    Returns (cur=frame1, ref=frame0, true_mv int32 [H//tile+1][W//tile+1][2]).
    """
    g = rng(seed)
    f0 = smooth_noise_u8(H, W, g)
    ty, tx = -(-H // tile), -(-W // tile)
    mv = g.integers(-R, R + 1, size=(ty, tx, 2), dtype=np.int32)
    yy, xx = np.mgrid[0:H, 0:W]
    my = mv[yy // tile, xx // tile, 0]
    mx = mv[yy // tile, xx // tile, 1]
    src = f0[np.clip(yy + my, 0, H - 1), np.clip(xx + mx, 0, W - 1)].astype(np.float64)
    f1 = np.clip(np.rint(src + g.normal(0.0, sigma, size=(H, W))), 0, 255).astype(np.uint8)
    return f1, f0, mv


def nv12_pair(H: int = 1080, W: int = 1920, R: int = 16, tile: int = 16, sigma: float = 2.0, seed: int = 1):
    """NV12 frame pair: the luma of me_pair and a 4:2:0 chroma plane pair
    moved with it (half the luma vector, floor), each plane smoothed noise
    plus N(0, sigma^2).  Returns (cur_y, cur_uv, ref_y, ref_uv, true_mv)."""
    cy, ry, mv = me_pair(H, W, R, tile, sigma, seed)
    g = rng(seed + 7)
    HC, WC = H // 2, W // 2
    yy, xx = np.mgrid[0:HC, 0:WC]
    my = mv[(2 * yy) // tile, (2 * xx) // tile, 0] >> 1
    mx = mv[(2 * yy) // tile, (2 * xx) // tile, 1] >> 1
    ruv = np.empty((HC, 2 * WC), np.uint8)
    cuv = np.empty_like(ruv)
    for ch in range(2):
        p0 = smooth_noise_u8(HC, WC, g)
        src = p0[np.clip(yy + my, 0, HC - 1), np.clip(xx + mx, 0, WC - 1)].astype(np.float64)
        ruv[:, ch::2] = p0
        cuv[:, ch::2] = np.clip(np.rint(src + g.normal(0.0, sigma, size=(HC, WC))), 0, 255).astype(np.uint8)
    return cy, cuv, ry, ruv, mv


def smooth_motion_pair(H: int, W: int, R: int = 16, sigma: float = 2.0, seed: int = 1):
    """Frame pair with a smooth affine motion field: frame1(y, x) =
    clip(round(frame0[cl(y+my)][cl(x+mx)] + N(0, sigma^2))) with
    (my, mx) = round(t + A (y/H - 1/2, x/W - 1/2)), |t| <= R/2, |A| <= R/2.
    Returns (cur=frame1, ref=frame0)."""
    g = rng(seed)
    f0 = smooth_noise_u8(H, W, g)
    t = g.uniform(-R / 2, R / 2, size=2)
    A = g.uniform(-R / 2, R / 2, size=(2, 2))
    yy, xx = np.mgrid[0:H, 0:W]
    u, v = yy / H - 0.5, xx / W - 0.5
    my = np.clip(np.rint(t[0] + A[0, 0] * u + A[0, 1] * v), -R, R).astype(np.int64)
    mx = np.clip(np.rint(t[1] + A[1, 0] * u + A[1, 1] * v), -R, R).astype(np.int64)
    src = f0[np.clip(yy + my, 0, H - 1), np.clip(xx + mx, 0, W - 1)].astype(np.float64)
    f1 = np.clip(np.rint(src + g.normal(0.0, sigma, size=(H, W))), 0, 255).astype(np.uint8)
    return f1, f0


def stereo_pair(H: int = 375, W: int = 1242, D: int = 128, tile: int = 32, sigma: float = 2.0, seed: int = 1):
    """Left/right pair: R[y][x] = L[y][cl(x + d)] + N(0, sigma^2), d piecewise
    constant over ``tile`` x ``tile`` tiles, uniform in [0, D-1]."""
    g = rng(seed)
    L = smooth_noise_u8(H, W, g)
    ty, tx = -(-H // tile), -(-W // tile)
    d = g.integers(0, D, size=(ty, tx), dtype=np.int32)
    yy, xx = np.mgrid[0:H, 0:W]
    dd = d[yy // tile, xx // tile]
    src = L[yy, np.clip(xx + dd, 0, W - 1)].astype(np.float64)
    R = np.clip(np.rint(src + g.normal(0.0, sigma, size=(H, W))), 0, 255).astype(np.uint8)
    return L, R, d


def uneven_sizes(n: int = 256, lo: int = 64, hi: int = 1024, seed: int = 1):
    """n (H, W) pairs, each side log-uniform in [lo, hi] independently."""
    g = rng(seed)
    e = g.uniform(np.log(lo), np.log(hi), size=(n, 2))
    return [(int(h), int(w)) for h, w in np.clip(np.rint(np.exp(e)), lo, hi).astype(int)]


def uneven_batch(n: int = 256, lo: int = 64, hi: int = 1024, seed: int = 1):
    """float32 images, values uniform [0, 1), with uneven log-uniform sizes; plus
    their uint8 copies floor(256 x)."""
    g = rng(seed + 1000)
    imgs = [g.random(s, dtype=np.float32) for s in uneven_sizes(n, lo, hi, seed)]
    u8 = [np.minimum(np.floor(im * 256.0), 255).astype(np.uint8) for im in imgs]
    return imgs, u8
