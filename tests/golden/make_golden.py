"""Writes the small known-answer fixtures in tests/golden/*.npz.

The expected outputs come from brute-force numpy written independently of
the C oracle and of the HIP kernels: frames are edge-padded with np.pad
(instead of clamping indices) and every offset is evaluated by slicing.  The
closed-form cases are asserted against their formulas before saving.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def brute_me(cur, ref, B, R):
    """SAD volume [nby][nbx][S][S] and first-min (dy, dx, sad) per block."""
    H, W = cur.shape
    nby, nbx, S = H // B, W // B, 2 * R + 1
    pad = np.pad(ref.astype(np.int64), ((R, R + B), (R, R + B)), mode="edge")
    vol = np.zeros((nby, nbx, S, S), np.int32)
    for by in range(nby):
        for bx in range(nbx):
            blk = cur[by * B:(by + 1) * B, bx * B:(bx + 1) * B].astype(np.int64)
            for sy in range(S):
                for sx in range(S):
                    win = pad[by * B + sy:by * B + sy + B, bx * B + sx:bx * B + sx + B]
                    vol[by, bx, sy, sx] = np.abs(blk - win).sum()
    flat = vol.reshape(nby, nbx, S * S)
    idx = flat.argmin(axis=2)  # numpy argmin returns the first minimum
    mv = np.stack([idx // S - R, idx % S - R, flat.min(axis=2)], axis=2).astype(np.int32)
    return mv, vol


def brute_cv(L, R, D, s0, step, KY, KX):
    H, W = L.shape
    smax = max(abs(s0), abs(s0 + step * (D - 1))) + KX + 1
    Lp = np.pad(L.astype(np.int64), ((KY, KY), (KX, KX)), mode="edge")
    Rp = np.pad(R.astype(np.int64), ((KY, KY), (smax, smax)), mode="edge")
    out = np.zeros((H, W, D), np.int32)
    for i in range(D):
        s = s0 + step * i
        acc = np.zeros((H, W), np.int64)
        for dy in range(2 * KY + 1):
            for dx in range(-KX, KX + 1):
                lv = Lp[dy:dy + H, KX + dx:KX + dx + W]
                rv = Rp[dy:dy + H, smax + dx + s:smax + dx + s + W]
                acc += np.abs(lv - rv)
        out[:, :, i] = acc
    return out


def main():
    g = np.random.Generator(np.random.PCG64(20170101))
    # --- ME: linear ramp, closed form SAD = B^2 |dx + 2 dy| away from edges
    H, W, B, R = 48, 64, 8, 4
    yy, xx = np.mgrid[0:H, 0:W]
    ramp = (xx + 2 * yy).astype(np.uint8)
    mv, vol = brute_me(ramp, ramp, B, R)
    for by in range(1, H // B - 1):
        for bx in range(1, W // B - 1):
            for sy in range(2 * R + 1):
                for sx in range(2 * R + 1):
                    assert vol[by, bx, sy, sx] == B * B * abs((sx - R) + 2 * (sy - R))
    np.savez_compressed(OUT / "me_ramp.npz", cur=ramp, ref=ramp, B=B, R=R, mv=mv, vol=vol)
    # --- ME: textured frame shifted by a known offset -> SAD 0 at the offset
    H, W, B, R = 40, 48, 8, 5
    ref = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    my, mx = 2, -3
    cur = ref[np.clip(yy[:H, :W] + my, 0, H - 1), np.clip(xx[:H, :W] + mx, 0, W - 1)]
    mv, vol = brute_me(cur, ref, B, R)
    assert (mv[1:-1, 1:-1, 0] == my).all() and (mv[1:-1, 1:-1, 1] == mx).all() and (mv[1:-1, 1:-1, 2] == 0).all()
    np.savez_compressed(OUT / "me_shift.npz", cur=cur, ref=ref, B=B, R=R, mv=mv, vol=vol)
    # --- ME: random frames with partial blocks at the right/bottom edges
    for name, (H, W, B, R) in {"me_random8": (27, 35, 8, 3), "me_random16": (48, 40, 16, 2)}.items():
        cur = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
        ref = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
        mv, vol = brute_me(cur, ref, B, R)
        np.savez_compressed(OUT / f"{name}.npz", cur=cur, ref=ref, B=B, R=R, mv=mv, vol=vol)
    # --- cost volume: the reference example's shape family on a ramp (1x7, s in -7..7):
    #     closed form 7 |s| away from the edges
    H, W = 6, 80
    ramp = np.tile(np.arange(W, dtype=np.uint8), (H, 1))
    cv = brute_cv(ramp, ramp, 15, -7, 1, 0, 3)
    for x in range(10, W - 10):
        for i in range(15):
            assert cv[2, x, i] == 7 * abs(i - 7)
    np.savez_compressed(OUT / "cv_ramp.npz", L=ramp, R=ramp, D=15, s0=-7, step=1, KY=0, KX=3, cost=cv)
    # --- cost volume: random stereo-shaped case (5x5 window, d = 0..D-1)
    H, W, D = 13, 29, 9
    L = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    Rr = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    cv = brute_cv(L, Rr, D, 0, -1, 2, 2)
    np.savez_compressed(OUT / "cv_random.npz", L=L, R=Rr, D=D, s0=0, step=-1, KY=2, KX=2, cost=cv)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
