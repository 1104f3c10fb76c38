"""GPU parity of the 3x3 filters, the bilateral filter and the integral image."""
import numpy as np
import pytest

import oracle
from umi_amd import DeviceBuffer, ops, workloads

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 70), (37, 45), (130, 257), (300, 500)]


@pytest.mark.parametrize("H,W", SHAPES)
@pytest.mark.parametrize("mode", [0, 1])
def test_filters_int_bit_exact(gpu, H, W, mode):
    g = workloads.rng(H * 7 + W + mode)
    img = g.integers(-1000, 1000, size=(H, W)).astype(np.int32)
    wts = g.integers(-9, 10, size=9).astype(np.int32)
    print(f"[filter] i32 {H}x{W} mode={mode}", flush=True)
    np.testing.assert_array_equal(ops.filter3x3_np(img, -5, 2, 999, mode), oracle.filter3x3(img, -5, 2, 999, mode))
    np.testing.assert_array_equal(ops.wfilter3x3_np(img, wts, -5, 56, mode), oracle.wfilter3x3(img, wts, -5, 56, mode))


@pytest.mark.parametrize("H,W", SHAPES)
@pytest.mark.parametrize("mode", [0, 1])
def test_filters_float_bit_exact(gpu, H, W, mode):
    g = workloads.rng(H * 11 + W + mode)
    img = g.random((H, W), dtype=np.float32)
    w = g.normal(size=9)
    w = (w / w.sum()).astype(np.float32)
    print(f"[filter] f32 {H}x{W} mode={mode}", flush=True)
    np.testing.assert_array_equal(ops.wfilter3x3_np(img, w, 0.25, 0.5, mode), oracle.wfilter3x3(img, w, 0.25, 0.5, mode))
    np.testing.assert_array_equal(ops.filter3x3_np(img, -0.5, 3.0, 2.0, mode), oracle.filter3x3(img, -0.5, 3.0, 2.0, mode))


def test_reference_examples_int(gpu):
    # main/filter3x3.cu and main/wfilter3x3.cu: iota images, their constants
    img = np.arange(300 * 500, dtype=np.int32).reshape(300, 500)
    np.testing.assert_array_equal(ops.filter3x3_np(img, -5, 2, 999), oracle.filter3x3(img, -5, 2, 999))
    img = np.arange(250 * 300, dtype=np.int32).reshape(250, 300)
    wts = np.array([1, 2, 1, 2, 4, 2, 1, 2, 1], np.int32)
    np.testing.assert_array_equal(ops.wfilter3x3_np(img, wts, -5, 56), oracle.wfilter3x3(img, wts, -5, 56))


def test_wfilter_1024_config(gpu):
    # 1024x1024 f32 U[0,1), 3x3 weights N(0,1) normalised to sum 1, zero padding:
    # bit-exact against the fmaf oracle and within 1e-6 of the magnitude sum vs FP64
    g = workloads.rng(1)
    img = g.random((1024, 1024), dtype=np.float32)
    w = g.normal(size=9)
    w = (w / w.sum()).astype(np.float32)
    out = ops.wfilter3x3_np(img, w, 0.0, 0.0)
    np.testing.assert_array_equal(out, oracle.wfilter3x3(img, w, 0.0, 0.0))
    ref, mag = oracle.wfilter3x3_f64(img, w, 0.0, 0.0)
    err = np.abs(out - ref)
    print(f"[filter] 1024^2 max |err|/mag = {np.max(err / mag):.3e}", flush=True)
    assert (err <= 1e-6 * mag).all()


@pytest.mark.parametrize("H,W,r,ss,sr", [(37, 45, 2, 2.0, 0.1), (250, 300, 2, 2.0, 0.1), (64, 80, 12, 4.0, 10.0),
                                         (5, 9, 0, 1.0, 0.1), (129, 67, 5, 3.0, 0.2)])
@pytest.mark.parametrize("mode", [0, 1])
def test_bilateral_vs_oracle(gpu, H, W, r, ss, sr, mode):
    # tolerance: the device weight 2^(d^2 norm log2 e) on v_exp_f32 vs host expf differ by a few ulp -> 2e-6 relative
    g = workloads.rng(H + W + r)
    img = g.random((H, W), dtype=np.float32)
    print(f"[bilateral] {H}x{W} r={r} mode={mode}", flush=True)
    out = ops.bilateral_np(img, r, ss, sr, 0.0, mode)
    ws = ops.spatial_weights(r, ss)
    norm = np.float32(-1.0 / (2.0 * sr * sr))
    ref = oracle.bilateral(img, r, ws, float(norm), 0.0, mode)
    rel = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-30)
    print(f"[bilateral] max rel vs f32 oracle {rel.max():.3e}", flush=True)
    assert rel.max() <= 2e-6
    ref64 = oracle.bilateral(img, r, ws, float(norm), 0.0, mode, f64=True)
    assert np.max(np.abs(out - ref64) / np.abs(ref64)) <= 1e-5


def test_bilateral_reference_example(gpu):
    # main/bilateral.cu: 300x250 iota, sigma_1 = 4 (k = 25), sigma_2 = 10, PAD 0.0; 0.1 % there
    img = np.arange(250 * 300, dtype=np.float32).reshape(250, 300)
    out = ops.bilateral_np(img, 12, 4.0, 10.0, 0.0, 0)
    ref = oracle.bilateral(img, 12, ops.spatial_weights(12, 4.0), float(np.float32(-1 / 200.0)), 0.0, 0)
    np.testing.assert_allclose(out, ref, rtol=1e-5)


@pytest.mark.parametrize("H,W", [(1, 1), (1, 700), (700, 1), (31, 65), (300, 500), (1080, 1920), (64, 5000), (257, 1024)])
def test_iimg_u8_i32_exact(gpu, H, W):
    g = workloads.rng(H * 3 + W)
    u8 = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    print(f"[iimg] u8 {H}x{W}", flush=True)
    np.testing.assert_array_equal(ops.iimg_np(u8), oracle.iimg(u8))
    i32 = g.integers(-2**31, 2**31 - 1, size=(H, W), dtype=np.int64).astype(np.int32)
    np.testing.assert_array_equal(ops.iimg_np(i32, use_workspace=False), oracle.iimg(i32))


def test_iimg_max_value_and_ones(gpu):
    H, W = 1024, 1024
    full = np.full((H, W), 255, np.uint8)  # 255 * 1024^2 < 2^32: no wrap
    np.testing.assert_array_equal(ops.iimg_np(full)[-1, -1], np.uint32(255 * H * W))
    ones = np.ones((300, 500), np.int32)  # main/iimg.cu's input
    yy, xx = np.mgrid[0:300, 0:500]
    np.testing.assert_array_equal(ops.iimg_np(ones), ((yy + 1) * (xx + 1)).astype(np.int32))


@pytest.mark.parametrize("H,W,pad", [(37, 300, 3), (5, 20000, 0), (300, 513, 13), (16, 256, 16)])
def test_iimg_pitched_and_wide(gpu, H, W, pad):
    # rows `pad` elements past W (pad 3 and 13 leave rows unaligned, so the tiles take the
    # element path), a width of 79 tiles, and a single whole tile
    g = workloads.rng(H * 7 + W + pad)
    u8 = g.integers(0, 256, size=(H, W + pad), dtype=np.int64).astype(np.uint8)
    f32 = g.random((H, W + pad), dtype=np.float32)
    print(f"[iimg] pitched {H}x{W} pitch {W + pad}", flush=True)
    o = DeviceBuffer(H * (W + 5) * 4)
    ops.iimg(DeviceBuffer.from_array(u8), H, W, W + pad, o, W + 5, np.uint8)
    np.testing.assert_array_equal(o.download((H, W + 5), np.uint32)[:, :W], oracle.iimg(u8[:, :W]))
    ops.iimg(DeviceBuffer.from_array(f32), H, W, W + pad, o, W + 5, np.float32)
    ref = oracle.iimg(np.ascontiguousarray(f32[:, :W]))
    got = o.download((H, W + 5), np.float32)[:, :W]
    assert np.max(np.abs(got.astype(np.float64) - ref) / ref) <= 4e-6


@pytest.mark.parametrize("H,W", [(31, 65), (1024, 1024), (1080, 1920), (100, 4000)])
def test_iimg_f32_fp64_column_pass(gpu, H, W):
    # fp32 row scan + fp64 column pass: relative error vs the fp64 oracle <= 4e-6
    g = workloads.rng(H + 5 * W)
    img = g.random((H, W), dtype=np.float32)
    out = ops.iimg_np(img)
    ref = oracle.iimg(img)
    rel = np.abs(out.astype(np.float64) - ref) / ref
    print(f"[iimg] f32 {H}x{W} max rel {rel.max():.3e} mean {rel.mean():.3e}", flush=True)
    assert rel.max() <= 4e-6


# ---------------------------------------------------------------------------
# Box sums read from the integral image (umi_box_sum_iimg_u32, batch form)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("H,W", [(1, 1), (3, 5), (17, 70), (64, 64), (200, 333)])
@pytest.mark.parametrize("ry,rx", [(0, 0), (1, 2), (4, 4), (100, 3)])
def test_box_sum_from_gpu_iimg(gpu, H, W, ry, rx):
    img = workloads.rng(H * 7 + W + ry).integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    print(f"[box] {H}x{W} r=({ry},{rx})", flush=True)
    S = ops.iimg_np(img)
    np.testing.assert_array_equal(ops.box_sum_np(S, ry, rx), oracle.box_sum(img, ry, rx))


def test_box_sum_wraps_mod_2_32(gpu):
    # an arbitrary uint32 table: every output is the corner formula mod 2^32
    g = workloads.rng(5)
    H, W, ry, rx = 37, 90, 3, 6
    S = g.integers(0, 2**32, size=(H, W), dtype=np.uint64).astype(np.uint32)
    print("[box] wrap", flush=True)
    got = ops.box_sum_np(S, ry, rx)
    Sp = np.zeros((H + 1, W + 1), np.uint32)
    Sp[1:, 1:] = S
    y = np.arange(H)[:, None]
    x = np.arange(W)[None, :]
    y1, y2 = np.maximum(y - ry, 0), np.minimum(y + ry, H - 1) + 1
    x1, x2 = np.maximum(x - rx, 0), np.minimum(x + rx, W - 1) + 1
    with np.errstate(over="ignore"):
        ref = Sp[y2, x2] - Sp[y1, x2] - Sp[y2, x1] + Sp[y1, x1]
    np.testing.assert_array_equal(got, ref)


def test_box_sum_1080p(gpu):
    img = workloads.rng(9).integers(0, 256, size=(1080, 1920), dtype=np.int64).astype(np.uint8)
    print("[box] 1080p r=4", flush=True)
    np.testing.assert_array_equal(ops.box_sum_np(ops.iimg_np(img), 4, 4), oracle.box_sum(img, 4, 4))


def test_batch_box_sum(gpu):
    from umi_amd import batch

    _, u8 = workloads.uneven_batch(n=20, lo=30, hi=400, seed=8)
    b = batch.ImageBatch(u8, out_dtype=np.uint32)
    ws = b.iimg_workspace()
    b.iimg(ws)
    box = DeviceBuffer(b.total * 4)
    print("[box] batch", flush=True)
    b.box_sum(3, 5, box)
    flat = box.download((b.total,), np.uint32)
    for im, o in zip(u8, b.offsets):
        h, w = im.shape
        np.testing.assert_array_equal(flat[o:o + h * w].reshape(h, w), oracle.box_sum(im, 3, 5))


def test_box_sum_pitched(gpu):
    # input and output pitches differ from the width (and from each other)
    H, W, ip, op = 45, 77, 84, 80
    img = workloads.rng(12).integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    S = np.zeros((H, ip), np.uint32)
    S[:, :W] = oracle.iimg(img)
    out = DeviceBuffer(H * op * 4)
    print("[box] pitched", flush=True)
    ops.box_sum(DeviceBuffer.from_array(S), H, W, ip, 3, 2, out, op)
    np.testing.assert_array_equal(out.download((H, op), np.uint32)[:, :W], oracle.box_sum(img, 3, 2))


def _random_features(g, n, win_h, win_w):
    f = np.zeros((n, 3, 5), np.int32)
    for i in range(n):
        for k in range(int(g.integers(1, 4))):
            w, h = int(g.integers(1, win_w + 1)), int(g.integers(1, win_h + 1))
            x, y = int(g.integers(0, win_w - w + 1)), int(g.integers(0, win_h - h + 1))
            f[i, k] = (x, y, w, h, int(g.integers(-4, 5)) or 1)
    return f


# (200, 300, 96, 96) and (120, 400, 40, 160) stage more than 48 KB and take the global-memory kernel;
# (100, 130, 24, 24) with 70 features spans three launches of at most 32 features
@pytest.mark.parametrize("H,W,wh,ww,nf", [(100, 130, 24, 24, 70), (24, 24, 24, 24, 12), (37, 300, 5, 19, 12),
                                          (270, 480, 24, 24, 12), (200, 300, 96, 96, 5), (120, 400, 40, 160, 5),
                                          (90, 100, 80, 80, 4)])
def test_haar_features_from_gpu_iimg(gpu, H, W, wh, ww, nf):
    g = workloads.rng(H + W + wh)
    img = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    feats = _random_features(g, nf, wh, ww)
    feats[0, 0] = (0, 0, ww, wh, 2**29)  # a large weight: the sums wrap as int32
    print(f"[haar] {H}x{W} window {wh}x{ww} features {len(feats)}", flush=True)
    got = ops.haar_features_np(ops.iimg_np(img), feats, wh, ww)
    np.testing.assert_array_equal(got, oracle.haar(img, feats, wh, ww))
