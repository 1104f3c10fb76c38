"""Pins the NV12 motion-search oracle (CPU only).

* chroma planes equal to zero in both frames: the chroma cost vanishes and
  the search is the luma full search (itself pinned by the reference's
  check loop), vector, cost and tie order included;
* flat luma and an integer chroma shift (even luma vector): the shifted copy
  is found with cost 0, in both directions (the floor of a negative half);
* the H.264 chroma predictor against hand-computed values of its formula
  (rounding terms, which samples enter each half-pel phase, clamping), then
  a current frame predicted at an odd vector is found there with cost 0;
* an independent numpy restatement (general bilinear weights) on tiny frames.
"""
import numpy as np
import pytest

import oracle
from umi_amd import workloads


def _rand(shape, g):
    return g.integers(0, 256, size=shape, dtype=np.int64).astype(np.uint8)


@pytest.mark.parametrize("R", [1, 4, 16])
def test_zero_chroma_is_luma_search(R):
    g = workloads.rng(10 + R)
    H, W = 48, 64
    cy, ry = _rand((H, W), g), _rand((H, W), g)
    z = np.zeros((H // 2, W), np.uint8)
    out = oracle.me_nv12(cy, z, ry, z, R)
    ref, _ = oracle.me_full_search(cy, ry, 8, R)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("dy,dx", [(2, 4), (-4, -2), (0, -6), (6, 0)])
def test_integer_chroma_shift_found(dy, dx):
    g = workloads.rng(20)
    H, W = 64, 64
    flat = np.full((H, W), 90, np.uint8)
    ruv = _rand((H // 2, W), g)
    U, V = ruv[:, 0::2], ruv[:, 1::2]
    yy, xx = np.mgrid[0:H // 2, 0:W // 2]
    sy, sx = np.clip(yy + dy // 2, 0, H // 2 - 1), np.clip(xx + dx // 2, 0, W // 2 - 1)
    cuv = np.empty_like(ruv)
    cuv[:, 0::2], cuv[:, 1::2] = U[sy, sx], V[sy, sx]
    out = oracle.me_nv12(flat, cuv, flat, ruv, 8)
    inner = out[2:-2, 2:-2]
    assert (inner[..., 0] == dy).all() and (inner[..., 1] == dx).all() and (inner[..., 2] == 0).all()


def _plane(vals):
    """A 2x2 chroma plane (U = A B / C D, V all 0) as an NV12 uv array."""
    uv = np.zeros((2, 4), np.uint8)
    uv[0, 0], uv[0, 2], uv[1, 0], uv[1, 2] = vals
    return uv


@pytest.mark.parametrize("vals,fy,fx,want", [
    ((200, 0, 0, 0), 0, 4, 100),    # (32*200 + 32) >> 6 = 100 (100.5 rounds down here)
    ((201, 0, 0, 0), 0, 4, 101),    # 100.5 + 0.5 -> 101: the +32 rounding term
    ((201, 0, 0, 0), 4, 0, 101),    # vertical half: C enters, B does not
    ((0, 201, 0, 0), 4, 0, 0),
    ((0, 0, 201, 0), 0, 4, 0),      # horizontal half: B enters, C does not
    ((200, 0, 0, 0), 4, 4, 50),     # (16*200 + 32) >> 6
    ((10, 11, 12, 13), 4, 4, 12),   # (46 + 2) >> 2: rounding of the four-sample mean
    ((10, 11, 12, 14), 4, 4, 12),   # (47 * 16 + 32) >> 6 = 12
    ((255, 255, 255, 255), 4, 4, 255),
    ((7, 9, 11, 13), 0, 0, 7),      # integer position: A itself
    ((0, 64, 0, 0), 0, 2, 16),      # quarter weights of the general formula: (16 * 64 + 32) >> 6
])
def test_chroma_predictor_known_answers(vals, fy, fx, want):
    assert oracle.nv12_chroma_pred(_plane(vals), 0, 0, 0, fy, fx) == want
    assert oracle.nv12_chroma_pred(_plane(vals), 1, 0, 0, fy, fx) == 0  # V plane is 0


def test_chroma_predictor_clamps_to_the_plane():
    uv = _plane((10, 20, 30, 40))
    # beyond the right and bottom edges B, C, D repeat the last column / row
    assert oracle.nv12_chroma_pred(uv, 0, 1, 1, 4, 4) == 40
    assert oracle.nv12_chroma_pred(uv, 0, 0, 1, 0, 4) == 20
    # above / left: clamped to row / column 0
    assert oracle.nv12_chroma_pred(uv, 0, -3, -3, 4, 4) == 10
    assert oracle.nv12_chroma_pred(uv, 0, -1, 0, 4, 0) == (10 + 10 + 1) >> 1


@pytest.mark.parametrize("dy,dx", [(0, 3), (3, 0), (-1, -1), (5, -3)])
def test_half_pel_vector_found(dy, dx):
    # flat luma; current chroma = the reference predicted at an odd vector
    # (through the predictor pinned above): that vector is found at cost 0
    g = workloads.rng(40)
    H, W = 64, 64
    HC, WC = H // 2, W // 2
    flat = np.full((H, W), 7, np.uint8)
    ruv = _rand((HC, W), g)
    cuv = np.zeros_like(ruv)
    oy, fy, ox, fx = (4 * dy) >> 3, (4 * dy) & 7, (4 * dx) >> 3, (4 * dx) & 7
    for y in range(HC):
        for x in range(WC):
            for ch in range(2):
                cuv[y, 2 * x + ch] = oracle.nv12_chroma_pred(ruv, ch, y + oy, x + ox, fy, fx)
    out = oracle.me_nv12(flat, cuv, flat, ruv, 6)
    assert (out[..., 0] == dy).all() and (out[..., 1] == dx).all() and (out[..., 2] == 0).all()


def _numpy_nv12(cy, cuv, ry, ruv, R):
    H, W = cy.shape
    HC, WC = H // 2, W // 2
    out = np.zeros((H // 8, W // 8, 3), np.int64)
    ref_c = [ruv[:, 0::2].astype(np.int64), ruv[:, 1::2].astype(np.int64)]
    cur_c = [cuv[:, 0::2].astype(np.int64), cuv[:, 1::2].astype(np.int64)]
    for by in range(H // 8):
        for bx in range(W // 8):
            best = None
            for dy in range(-R, R + 1):
                for dx in range(-R, R + 1):
                    ys = np.clip(np.arange(8) + 8 * by + dy, 0, H - 1)
                    xs = np.clip(np.arange(8) + 8 * bx + dx, 0, W - 1)
                    cost = np.abs(cy[8 * by:8 * by + 8, 8 * bx:8 * bx + 8].astype(np.int64) - ry[np.ix_(ys, xs)]).sum()
                    # eighth-pel chroma vector (4 dy, 4 dx), general bilinear weights
                    vy, vx = 4 * dy, 4 * dx
                    Y0 = 4 * by + np.arange(4) + vy // 8
                    X0 = 4 * bx + np.arange(4) + vx // 8
                    wy, wx = vy % 8, vx % 8
                    y0, y1 = np.clip(Y0, 0, HC - 1), np.clip(Y0 + 1, 0, HC - 1)
                    x0, x1 = np.clip(X0, 0, WC - 1), np.clip(X0 + 1, 0, WC - 1)
                    for P, C in zip(ref_c, cur_c):
                        pred = ((8 - wx) * (8 - wy) * P[np.ix_(y0, x0)] + wx * (8 - wy) * P[np.ix_(y0, x1)]
                                + (8 - wx) * wy * P[np.ix_(y1, x0)] + wx * wy * P[np.ix_(y1, x1)] + 32) >> 6
                        cost += np.abs(C[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] - pred).sum()
                    if best is None or cost < best[2]:
                        best = (dy, dx, cost)
            out[by, bx] = best
    return out.astype(np.int32)


@pytest.mark.parametrize("R", [1, 3])
def test_numpy_restatement(R):
    g = workloads.rng(30 + R)
    H, W = 24, 32
    cy, ry = _rand((H, W), g), _rand((H, W), g)
    cuv, ruv = _rand((H // 2, W), g), _rand((H // 2, W), g)
    np.testing.assert_array_equal(oracle.me_nv12(cy, cuv, ry, ruv, R), _numpy_nv12(cy, cuv, ry, ruv, R))
