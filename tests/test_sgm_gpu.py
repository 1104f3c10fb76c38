"""GPU parity of semi-global matching against the oracle (bit-exact).

The oracle's four path recurrences are pinned in test_oracle_sgm.py; here the
two-launch HIP path (vertical sweeps, then horizontal sweeps with the argmin)
must reproduce its (index, aggregated cost) per pixel exactly: every window
size, both shift directions, disparity counts around the 64-lane boundaries,
one-row / one-column / chunk-crossing images, and the 1242x375 pair with 128
disparities.
"""
import numpy as np
import pytest

import oracle
from umi_amd import DeviceBuffer, ops, workloads

pytestmark = pytest.mark.gpu


def _pair(H, W, seed):
    g = workloads.rng(seed)
    L = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    R = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    return L, R


@pytest.mark.parametrize("KY", [0, 1, 2, 3])
@pytest.mark.parametrize("KX", [0, 1, 2, 3])
def test_sgm_windows(gpu, KY, KX):
    L, R = _pair(37, 83, 300 + 4 * KY + KX)
    for s0, step, D in [(0, -1, 48), (-20, 1, 70)]:
        print(f"[sgm] KY={KY} KX={KX} s0={s0} step={step} D={D}", flush=True)
        out = ops.sgm_np(L, R, D, s0, step, KY, KX, 7, 90)
        np.testing.assert_array_equal(out, oracle.sgm(L, R, D, s0, step, KY, KX, 7, 90))


@pytest.mark.parametrize("D", [1, 2, 63, 64, 65, 127, 128])
def test_sgm_disparity_counts(gpu, D):
    L, R = _pair(23, 150, 400 + D)
    print(f"[sgm] D={D}", flush=True)
    out = ops.sgm_np(L, R, D, 0, -1, 1, 1, 12, 200)
    np.testing.assert_array_equal(out, oracle.sgm(L, R, D, 0, -1, 1, 1, 12, 200))


@pytest.mark.parametrize("H,W", [(1, 1), (1, 57), (61, 1), (2, 3), (64, 9), (65, 10), (129, 13), (200, 66)])
def test_sgm_shapes(gpu, H, W):
    L, R = _pair(H, W, 500 + H * 7 + W)
    print(f"[sgm] {H}x{W}", flush=True)
    out = ops.sgm_np(L, R, 40, 0, -1, 2, 2, 10, 120)
    np.testing.assert_array_equal(out, oracle.sgm(L, R, 40, 0, -1, 2, 2, 10, 120))


@pytest.mark.parametrize("P1,P2", [(0, 0), (0, 500), (15, 15), (8, 240)])
def test_sgm_penalties(gpu, P1, P2):
    L, R, _ = workloads.stereo_pair(48, 160, 32, tile=16, sigma=4.0, seed=9)
    print(f"[sgm] P1={P1} P2={P2}", flush=True)
    out = ops.sgm_np(L, R, 32, 0, -1, 2, 2, P1, P2)
    np.testing.assert_array_equal(out, oracle.sgm(L, R, 32, 0, -1, 2, 2, P1, P2))


def test_sgm_zero_penalties_is_winner_take_all(gpu):
    # P1 = P2 = 0: every path cost is the matching cost, S = 4 C (GPU against GPU)
    L, R, _ = workloads.stereo_pair(90, 300, 64, seed=4)
    wta = ops.stereo_wta_np(L, R, 64, 0, -1, 2, 2)
    out = ops.sgm_np(L, R, 64, 0, -1, 2, 2, 0, 0)
    np.testing.assert_array_equal(out[..., 0], wta[..., 0])
    np.testing.assert_array_equal(out[..., 1], 4 * wta[..., 1])


def test_sgm_pitched_rows_and_reused_workspace(gpu):
    # images inside wider rows; the workspace is reused without clearing
    H, W, pitch, D = 70, 100, 128, 50
    L, R = _pair(H, pitch, 77)
    dl, dr = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
    ws = DeviceBuffer(ops.sgm_workspace_bytes(H, W, D, 0, -1, 2, 2))
    ws.fill(0xAB)
    out = DeviceBuffer(H * W * 8)
    ref = oracle.sgm(np.ascontiguousarray(L[:, :W]), np.ascontiguousarray(R[:, :W]), D, 0, -1, 2, 2, 10, 120)
    for _ in range(2):
        ops.sgm(dl, dr, H, W, pitch, D, 0, -1, 2, 2, 10, 120, ws, out)
        np.testing.assert_array_equal(out.download((H, W, 2), np.int32), ref)


def test_sgm_kitti_size_vs_oracle(gpu):
    # 1242x375 pair, 128 disparities, 5x5 window, P1 = 10, P2 = 120: every pixel bit-exact
    L, R, _ = workloads.stereo_pair(375, 1242, 128, seed=1)
    print("[sgm] 375x1242 D=128", flush=True)
    out = ops.sgm_np(L, R, 128, 0, -1, 2, 2, 10, 120)
    np.testing.assert_array_equal(out, oracle.sgm(L, R, 128, 0, -1, 2, 2, 10, 120))


@pytest.mark.parametrize("s0,step,D", [(25, 1, 40), (-200, 1, 16), (300, -1, 20), (-5, -1, 64), (3, 1, 33)])
def test_sgm_shifts_past_the_image(gpu, s0, step, D):
    # shifts beyond the image width read clamped columns only
    L, R = _pair(21, 30, 600 + D)
    print(f"[sgm] s0={s0} step={step} D={D}", flush=True)
    out = ops.sgm_np(L, R, D, s0, step, 1, 2, 6, 70)
    np.testing.assert_array_equal(out, oracle.sgm(L, R, D, s0, step, 1, 2, 6, 70))
