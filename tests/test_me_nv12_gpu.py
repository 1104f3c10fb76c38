"""GPU parity of the NV12 motion search and the motion-vector export
(bit-exact against the oracle and the published record layouts)."""
import numpy as np
import pytest

import oracle
from umi_amd import DeviceBuffer, ops, workloads

pytestmark = pytest.mark.gpu


def _frames(H, W, seed):
    g = workloads.rng(seed)
    cy = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    ry = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    cuv = g.integers(0, 256, size=(H // 2, W), dtype=np.int64).astype(np.uint8)
    ruv = g.integers(0, 256, size=(H // 2, W), dtype=np.int64).astype(np.uint8)
    return cy, cuv, ry, ruv


@pytest.mark.parametrize("R", [1, 2, 5, 8, 15, 16])
def test_nv12_search_radii(gpu, R):
    cy, cuv, ry, ruv = _frames(48, 272, 50 + R)
    print(f"[nv12] 48x272 R={R}", flush=True)
    np.testing.assert_array_equal(ops.me_full_search_nv12_np(cy, cuv, ry, ruv, R), oracle.me_nv12(cy, cuv, ry, ruv, R))


@pytest.mark.parametrize("H,W", [(8, 8), (10, 18), (70, 102), (64, 136), (136, 256)])
def test_nv12_shapes(gpu, H, W):
    # partial workgroups (nbx not a multiple of 16), sizes not multiples of 8
    cy, cuv, ry, ruv = _frames(H, W, 60 + H + W)
    print(f"[nv12] {H}x{W}", flush=True)
    np.testing.assert_array_equal(ops.me_full_search_nv12_np(cy, cuv, ry, ruv, 16),
                                  oracle.me_nv12(cy, cuv, ry, ruv, 16))


def test_nv12_zero_chroma_equals_luma_search(gpu):
    cy, _, ry, _ = _frames(64, 128, 70)
    z = np.zeros((32, 128), np.uint8)
    out = ops.me_full_search_nv12_np(cy, z, ry, z, 16)
    luma, _ = ops.me_full_search_np(cy, ry, 8, 16)
    np.testing.assert_array_equal(out, luma)


@pytest.mark.parametrize("H,W,py,puv,R", [(40, 96, 128, 112, 7), (48, 100, 128, 112, 16), (40, 264, 272, 264, 16)])
def test_nv12_pitched_planes(gpu, H, W, py, puv, R):
    # 16-byte aligned luma pitch: 16-byte staging, with a chunk across W when W % 16 != 0
    cy, cuv, ry, ruv = _frames(H, py, 80 + W)
    cuv, ruv = cuv[:, :puv], ruv[:, :puv]
    d = [DeviceBuffer.from_array(np.ascontiguousarray(a)) for a in (cy, cuv, ry, ruv)]
    out = DeviceBuffer((H // 8) * (W // 8) * 12)
    print(f"[nv12] pitched {H}x{W} pitch {py}/{puv} R={R}", flush=True)
    ops.me_full_search_nv12(*d, H, W, py, puv, R, out)
    want = oracle.me_nv12(np.ascontiguousarray(cy[:, :W]), np.ascontiguousarray(cuv[:, :W]),
                          np.ascontiguousarray(ry[:, :W]), np.ascontiguousarray(ruv[:, :W]), R)
    np.testing.assert_array_equal(out.download((H // 8, W // 8, 3), np.int32), want)


def test_nv12_1080p_vs_oracle(gpu):
    # 1920x1080 NV12 pair, luma moved per 16x16 tile in [-16, 16]^2 and the
    # chroma with it: every block bit-exact
    cy, cuv, ry, ruv, _ = workloads.nv12_pair(1080, 1920, seed=3)
    print("[nv12] 1080p R=16", flush=True)
    out = ops.me_full_search_nv12_np(cy, cuv, ry, ruv, 16)
    try:
        oracle.set_threads(16)
        want = oracle.me_nv12(cy, cuv, ry, ruv, 16)
    finally:
        oracle.set_threads(1)
    np.testing.assert_array_equal(out, want)


@pytest.mark.parametrize("units", [1, 4])
def test_export_q4_int16(gpu, units):
    g = workloads.rng(90 + units)
    mv = np.zeros((7, 11, 3), np.int32)
    mv[..., :2] = g.integers(-64, 65, size=(7, 11, 2))
    mv[..., 2] = g.integers(0, 5000, size=(7, 11))
    out = ops.export_mvs_np(mv, 8, units, ops.MV_Q4_INT16)
    scale = 4 if units == 1 else 1
    np.testing.assert_array_equal(out[..., 0], mv[..., 1] * scale)  # x first
    np.testing.assert_array_equal(out[..., 1], mv[..., 0] * scale)


@pytest.mark.parametrize("B,units", [(8, 1), (16, 1), (8, 4)])
def test_export_avmotionvector(gpu, B, units):
    g = workloads.rng(95 + B + units)
    nby, nbx = 5, 9
    mv = np.zeros((nby, nbx, 3), np.int32)
    mv[..., :2] = g.integers(-63, 64, size=(nby, nbx, 2))
    rec = ops.export_mvs_np(mv, B, units, ops.MV_AVMOTIONVECTOR)
    q = 4 if units == 1 else 1
    by, bx = np.mgrid[0:nby, 0:nbx]
    assert (rec["source"] == -1).all() and (rec["w"] == B).all() and (rec["h"] == B).all()
    assert (rec["flags"] == 0).all() and (rec["motion_scale"] == 4).all()
    np.testing.assert_array_equal(rec["dst_x"], bx * B + B // 2)
    np.testing.assert_array_equal(rec["dst_y"], by * B + B // 2)
    np.testing.assert_array_equal(rec["motion_x"], mv[..., 1] * q)
    np.testing.assert_array_equal(rec["motion_y"], mv[..., 0] * q)
    # src = dst + motion / motion_scale (floor for quarter pels)
    np.testing.assert_array_equal(rec["src_x"], rec["dst_x"] + ((mv[..., 1] * q) >> 2))
    np.testing.assert_array_equal(rec["src_y"], rec["dst_y"] + ((mv[..., 0] * q) >> 2))


def test_subpel_then_export(gpu):
    # the encoder path: NV12 luma search -> quarter-pel refinement -> records
    cy, cuv, ry, ruv = _frames(64, 96, 99)
    mv = ops.me_full_search_nv12_np(cy, cuv, ry, ruv, 8)
    q = ops.me_subpel_refine_np(cy, ry, mv)
    rec = ops.export_mvs_np(q, 8, 4, ops.MV_AVMOTIONVECTOR)
    np.testing.assert_array_equal(rec["motion_x"], q[..., 1])
    np.testing.assert_array_equal(rec["motion_y"], q[..., 0])
