"""Pins the CPU oracle for motion estimation and the cost volume.

Checks, each chosen so that a plausible slip (swapped dy/dx, a missing clamp,
a wrong tie-break, an off-by-one offset) fails one of them:
  * golden fixtures (tests/golden, brute force over an edge-padded frame);
  * closed forms: int32 iota frames give SAD = B^2 |W dy + dx| in the interior
    (the known answer of main/full_me.cu:29,66-87); a ramp x + 2y gives
    B^2 |dx + 2 dy| and a first minimum at (dy, dx) = (-R/2, R) (tie-break);
    a 1x7 cost volume of a ramp gives 7 |s| (main/cost_volume.cu:30,67-80);
  * invariants: a constant frame -> every SAD 0 -> MV (-R, -R).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["me_ramp", "me_shift", "me_random8", "me_random16"])
def test_oracle_me_matches_golden(name):
    f = np.load(GOLD / f"{name}.npz")
    mv, vol = oracle.me_full_search(f["cur"], f["ref"], int(f["B"]), int(f["R"]), volume=True)
    np.testing.assert_array_equal(vol, f["vol"])
    np.testing.assert_array_equal(mv, f["mv"])


def test_oracle_sad_volume_iota_closed_form():
    # main/full_me.cu: iota frame searched against itself, interior SAD = B^2 |W dy + dx|
    H, W, B, S, ofs = 64, 200, 8, 15, -7
    f = np.arange(H * W, dtype=np.int32).reshape(H, W)
    vol = oracle.sad_volume_i32(f, f, B, S, ofs)
    for by in range(1, H // B - 1):
        for bx in range(1, W // B - 1):
            dy = np.arange(S)[:, None] + ofs
            dx = np.arange(S)[None, :] + ofs
            np.testing.assert_array_equal(vol[by, bx], B * B * np.abs(W * dy + dx))


def test_oracle_me_ramp_tie_break():
    H, W, B, R = 64, 64, 8, 6
    yy, xx = np.mgrid[0:H, 0:W]
    ramp = (xx + 2 * yy).astype(np.uint8)
    mv, _ = oracle.me_full_search(ramp, ramp, B, R)
    # zero SAD on the line dx = -2 dy; the first in raster order has the smallest dy
    assert tuple(mv[3, 3]) == (-R // 2, R, 0)


def test_oracle_me_constant_frame():
    f = np.full((32, 32), 77, np.uint8)
    mv, _ = oracle.me_full_search(f, f, 8, 3)
    assert (mv[..., 0] == -3).all() and (mv[..., 1] == -3).all() and (mv[..., 2] == 0).all()


def test_oracle_me_row_range_only_touches_rows():
    f = np.load(GOLD / "me_random8.npz")
    mv, _ = oracle.me_full_search(f["cur"], f["ref"], 8, 3, by_range=(1, 2))
    np.testing.assert_array_equal(mv[1], f["mv"][1])
    assert (mv[0] == 0).all() and (mv[2] == 0).all()


def test_fixture_discriminates_swapped_offsets():
    # A transposed search (dy <-> dx) must not pass the ramp fixture.
    f = np.load(GOLD / "me_ramp.npz")
    swapped = np.swapaxes(f["vol"], 2, 3)
    assert not np.array_equal(swapped, f["vol"])


@pytest.mark.parametrize("name", ["cv_ramp", "cv_random"])
def test_oracle_cost_volume_matches_golden(name):
    f = np.load(GOLD / f"{name}.npz")
    out = oracle.cost_volume(f["L"], f["R"], int(f["D"]), int(f["s0"]), int(f["step"]), int(f["KY"]), int(f["KX"]))
    np.testing.assert_array_equal(out, f["cost"])


def test_oracle_cost_volume_ramp_closed_form():
    W = 120
    ramp = np.tile(np.arange(W, dtype=np.uint8), (4, 1))
    out = oracle.cost_volume(ramp, ramp, 15, -7, 1, 0, 3)
    for i in range(15):
        assert (out[:, 12:W - 12, i] == 7 * abs(i - 7)).all()
