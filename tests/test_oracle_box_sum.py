"""Pins the CPU oracle of the box sums read from an integral image.

  * closed form: an all-ones image gives the clipped window's area,
    (min(y+ry, H-1) - max(y-ry, 0) + 1) x (min(x+rx, W-1) - max(x-rx, 0) + 1);
  * windows larger than the image give the whole image's sum everywhere;
  * radius 0 gives the image itself;
  * random images against sums of numpy slices.
"""
import numpy as np
import pytest

import oracle
from umi_amd import workloads


def test_box_sum_ones_is_clipped_area():
    H, W, ry, rx = 9, 14, 2, 5
    out = oracle.box_sum(np.ones((H, W), np.uint8), ry, rx)
    y = np.arange(H)[:, None]
    x = np.arange(W)[None, :]
    area = (np.minimum(y + ry, H - 1) - np.maximum(y - ry, 0) + 1) * (np.minimum(x + rx, W - 1) - np.maximum(x - rx, 0) + 1)
    np.testing.assert_array_equal(out, area.astype(np.uint32))


def test_box_sum_large_window_and_radius_zero():
    img = workloads.rng(1).integers(0, 256, size=(7, 11), dtype=np.int64).astype(np.uint8)
    assert (oracle.box_sum(img, 50, 50) == img.astype(np.uint32).sum()).all()
    np.testing.assert_array_equal(oracle.box_sum(img, 0, 0), img.astype(np.uint32))


@pytest.mark.parametrize("ry,rx", [(1, 1), (0, 3), (4, 2)])
def test_box_sum_random_vs_slices(ry, rx):
    img = workloads.rng(10 + ry * 7 + rx).integers(0, 256, size=(13, 19), dtype=np.int64).astype(np.uint8)
    out = oracle.box_sum(img, ry, rx)
    for y in range(13):
        for x in range(19):
            ref = img[max(y - ry, 0):y + ry + 1, max(x - rx, 0):x + rx + 1].astype(np.uint32).sum()
            assert out[y, x] == ref


# ---------------------------------------------------------------------------
# Rectangle (Haar-like) features
# ---------------------------------------------------------------------------
def _feats(*rects_per_feature):
    f = np.zeros((len(rects_per_feature), 3, 5), np.int32)
    for i, rects in enumerate(rects_per_feature):
        for k, r in enumerate(rects):
            f[i, k] = r
    return f


def test_haar_oracle_window_sum_and_step_edge():
    g = workloads.rng(2)
    img = g.integers(0, 256, size=(20, 30), dtype=np.int64).astype(np.uint8)
    f = _feats([(0, 0, 8, 6, 1)])                       # the whole 6x8 window
    out = oracle.haar(img, f, 6, 8)
    ref = np.array([[img[y:y + 6, x:x + 8].astype(np.int64).sum() for x in range(23)] for y in range(15)])
    np.testing.assert_array_equal(out[0], ref.astype(np.int32))
    # a vertical step (left 0, right 255) under a two-rectangle edge feature
    step = np.zeros((10, 20), np.uint8)
    step[:, 10:] = 255
    edge = _feats([(0, 0, 2, 4, -1), (2, 0, 2, 4, 1)])  # right half minus left half of a 4x4 window
    out = oracle.haar(step, edge, 4, 4)
    x = np.arange(17)
    right = 4 * 255 * (np.clip(x + 4, 10, None) - np.clip(x + 2, 10, None))
    left = 4 * 255 * (np.clip(x + 2, 10, None) - np.clip(x, 10, None))
    np.testing.assert_array_equal(out[0], np.tile(right - left, (7, 1)).astype(np.int32))


def test_haar_oracle_weights_wrap_like_int32():
    img = np.full((8, 8), 255, np.uint8)
    f = _feats([(0, 0, 8, 8, 2**30)])                   # 2^30 x 16320 wraps mod 2^32
    out = oracle.haar(img, f, 8, 8)
    want = np.array([(2**30 * 16320) % 2**32], np.uint32).view(np.int32)[0]
    assert out[0, 0, 0] == want
