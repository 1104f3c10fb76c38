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
