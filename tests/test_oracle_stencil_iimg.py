"""Pins the CPU oracle for the 3x3 filters, the bilateral filter and the integral image.

  * the reference's own known answers: iota images through filter3x3
    (main/filter3x3.cu:21-78: init -5, mul 2, pad 999) and wfilter3x3
    (main/wfilter3x3.cu:19-89: weights 1 2 1 / 2 4 2 / 1 2 1, init -5, pad 56),
    evaluated here as closed forms over in-image/out-of-image taps; the
    all-ones integral image (main/iimg.cu:31,84-112) -> (y+1)(x+1);
  * exactness: float filters on small integers equal the integer oracle;
  * invariants: bilateral of a constant image with clamp = the constant;
    with zero padding a closed form; radius 0 = identity;
  * independent numpy: cumulative sums for the integral image, float64 sums
    for the weighted filter.
"""
import numpy as np
import pytest

import oracle


def taps_closed_form(H, W, weights, init, pad, mul=1):
    """sum over 3x3 taps of w * (iota value or pad), vectorised numpy."""
    yy, xx = np.mgrid[0:H, 0:W].astype(np.int64)
    acc = np.full((H, W), init, np.int64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            y2, x2 = yy + dy, xx + dx
            inside = (y2 >= 0) & (y2 < H) & (x2 >= 0) & (x2 < W)
            acc += weights[dy + 1][dx + 1] * np.where(inside, y2 * W + x2, pad)
    return (acc * mul).astype(np.int64)


def test_filter3x3_reference_example_known_answer():
    H, W = 300, 500
    img = np.arange(H * W, dtype=np.int32).reshape(H, W)
    out = oracle.filter3x3(img, -5, 2, 999)
    ones = [[1, 1, 1]] * 3
    np.testing.assert_array_equal(out, taps_closed_form(H, W, ones, -5, 999, mul=2).astype(np.int32))
    assert out[100, 100] == 2 * (-5 + 9 * (500 * 100 + 100))


def test_wfilter3x3_reference_example_known_answer():
    H, W = 250, 300
    wts = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]])
    img = np.arange(H * W, dtype=np.int32).reshape(H, W)
    out = oracle.wfilter3x3(img, wts, -5, 56)
    np.testing.assert_array_equal(out, taps_closed_form(H, W, wts, -5, 56).astype(np.int32))
    assert out[10, 10] == -5 + 16 * (300 * 10 + 10)


def test_float_filters_equal_integer_oracle_on_small_integers():
    g = np.random.Generator(np.random.PCG64(3))
    img = g.integers(-50, 50, size=(37, 45)).astype(np.int32)
    wts = g.integers(-4, 5, size=9).astype(np.int32)
    np.testing.assert_array_equal(oracle.wfilter3x3(img.astype(np.float32), wts.astype(np.float32), 3.0, 7.0),
                                  oracle.wfilter3x3(img, wts, 3, 7).astype(np.float32))
    np.testing.assert_array_equal(oracle.filter3x3(img.astype(np.float32), -2.0, 3.0, 5.0),
                                  oracle.filter3x3(img, -2, 3, 5).astype(np.float32))
    for mode in (0, 1):
        np.testing.assert_array_equal(oracle.wfilter3x3(img.astype(np.float32), wts.astype(np.float32), 0.0, 2.0, mode),
                                      oracle.wfilter3x3(img, wts, 0, 2, mode).astype(np.float32))


def test_clamp_mode_matches_edge_padding():
    g = np.random.Generator(np.random.PCG64(4))
    img = g.integers(0, 1000, size=(20, 30)).astype(np.int32)
    wts = np.arange(9, dtype=np.int32).reshape(3, 3)
    p = np.pad(img.astype(np.int64), 1, mode="edge")
    ref = sum(wts[dy, dx] * p[dy:dy + 20, dx:dx + 30] for dy in range(3) for dx in range(3))
    np.testing.assert_array_equal(oracle.wfilter3x3(img, wts, 0, 0, mode=1), ref.astype(np.int32))


def test_wfilter_f32_close_to_f64():
    g = np.random.Generator(np.random.PCG64(5))
    img = g.random((64, 80), dtype=np.float32)
    w = g.normal(size=9)
    w = (w / w.sum()).astype(np.float32)
    out = oracle.wfilter3x3(img, w, 0.0, 0.0)
    ref, mag = oracle.wfilter3x3_f64(img, w, 0.0, 0.0)
    assert (np.abs(out - ref) <= 1e-6 * mag).all()


def test_bilateral_invariants():
    r, sigma_s = 2, 2.0
    d = np.arange(-r, r + 1)
    ws = np.exp(-d * d / (2 * sigma_s ** 2)).astype(np.float32)
    norm = np.float32(-50.0)
    c = np.float32(0.3)
    img = np.full((9, 11), c, np.float32)
    out = oracle.bilateral(img, r, ws, norm, 0.0, mode=1)
    np.testing.assert_allclose(out, c, rtol=2e-7)
    # zero padding: corner pixel sees 9 of 25 taps inside, the rest have n = 0
    out0 = oracle.bilateral(img, r, ws, norm, 0.0, mode=0)
    W2 = np.outer(ws, ws).astype(np.float64)
    inside = np.zeros((5, 5), bool)
    inside[2:, 2:] = True
    wout = np.exp(float(c) ** 2 * float(norm))
    expect = (W2[inside].sum() * float(c)) / (W2[inside].sum() + wout * W2[~inside].sum())
    assert abs(out0[0, 0] - expect) <= 1e-6 * expect
    g = np.random.Generator(np.random.PCG64(6))
    rnd = g.random((10, 12), dtype=np.float32)
    np.testing.assert_array_equal(oracle.bilateral(rnd, 0, np.ones(1, np.float32), norm), rnd)


def test_bilateral_f32_close_to_f64():
    g = np.random.Generator(np.random.PCG64(7))
    img = g.random((40, 50), dtype=np.float32)
    d = np.arange(-2, 3)
    ws = np.exp(-d * d / 8.0).astype(np.float32)
    out = oracle.bilateral(img, 2, ws, -50.0)
    ref = oracle.bilateral(img, 2, ws, -50.0, f64=True)
    assert np.max(np.abs(out - ref) / np.abs(ref)) < 1e-5


def test_iimg_all_ones_known_answer():
    H, W = 300, 500
    out = oracle.iimg(np.ones((H, W), np.int32))
    yy, xx = np.mgrid[0:H, 0:W]
    np.testing.assert_array_equal(out, ((yy + 1) * (xx + 1)).astype(np.int32))


def test_iimg_matches_numpy_cumsum():
    g = np.random.Generator(np.random.PCG64(8))
    u8 = g.integers(0, 256, size=(70, 90), dtype=np.int64).astype(np.uint8)
    np.testing.assert_array_equal(oracle.iimg(u8), np.cumsum(np.cumsum(u8.astype(np.uint64), 0), 1).astype(np.uint32))
    i32 = g.integers(-2**31, 2**31 - 1, size=(30, 40), dtype=np.int64).astype(np.int32)
    ref = (np.cumsum(np.cumsum(i32.astype(np.int64), 0), 1) & 0xffffffff).astype(np.uint32).view(np.int32)
    np.testing.assert_array_equal(oracle.iimg(i32), ref)
    f = g.random((50, 60), dtype=np.float32)
    np.testing.assert_allclose(oracle.iimg(f), np.cumsum(np.cumsum(f.astype(np.float64), 0), 1), rtol=1e-12)
