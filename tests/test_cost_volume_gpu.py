"""GPU parity of the SAD cost volume against the oracle (bit-exact)."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from umi_amd import ops, workloads

GOLD = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cv_ramp", "cv_random"])
def test_cost_volume_golden(gpu, name):
    f = np.load(GOLD / f"{name}.npz")
    print(f"[cv] golden {name}", flush=True)
    out = ops.cost_volume_np(f["L"], f["R"], int(f["D"]), int(f["s0"]), int(f["step"]), int(f["KY"]), int(f["KX"]))
    np.testing.assert_array_equal(out, f["cost"])


@pytest.mark.parametrize("KY", [0, 1, 2, 3])
@pytest.mark.parametrize("KX", [0, 1, 2, 3])
def test_cost_volume_windows(gpu, KY, KX):
    g = workloads.rng(100 + 4 * KY + KX)
    H, W, D = 41, 75, 70
    L = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    R = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    for s0, step in [(0, -1), (-20, 1)]:
        print(f"[cv] KY={KY} KX={KX} s0={s0} step={step}", flush=True)
        out = ops.cost_volume_np(L, R, D, s0, step, KY, KX)
        np.testing.assert_array_equal(out, oracle.cost_volume(L, R, D, s0, step, KY, KX))


def test_cost_volume_reference_example_shape(gpu):
    # main/cost_volume.cu: one frame, 1x7 window, shifts -7..7 (here uint8 and 135x240)
    g = workloads.rng(7)
    f = g.integers(0, 256, size=(135, 240), dtype=np.int64).astype(np.uint8)
    out = ops.cost_volume_np(f, f, 15, -7, 1, 0, 3)
    np.testing.assert_array_equal(out, oracle.cost_volume(f, f, 15, -7, 1, 0, 3))


def test_cost_volume_kitti_size_vs_oracle(gpu):
    # 1242x375 pair, 128 disparities, 5x5 window: the whole volume bit-exact
    L, R, _ = workloads.stereo_pair(375, 1242, 128, seed=1)
    print("[cv] 375x1242 D=128", flush=True)
    out = ops.cost_volume_np(L, R, 128, 0, -1, 2, 2)
    np.testing.assert_array_equal(out, oracle.cost_volume(L, R, 128, 0, -1, 2, 2))


# ---------------------------------------------------------------------------
# Stereo winner-take-all fused with the cost volume (umi_stereo_wta_u8)
# ---------------------------------------------------------------------------
def _first_min(vol):
    idx = np.argmin(vol, axis=-1).astype(np.int32)   # numpy returns the first minimum
    return np.stack([idx, np.take_along_axis(vol, idx[..., None], -1)[..., 0]], axis=-1)


@pytest.mark.parametrize("name", ["cv_ramp", "cv_random"])
def test_stereo_wta_golden(gpu, name):
    f = np.load(GOLD / f"{name}.npz")
    print(f"[wta] golden {name}", flush=True)
    out = ops.stereo_wta_np(f["L"], f["R"], int(f["D"]), int(f["s0"]), int(f["step"]), int(f["KY"]), int(f["KX"]))
    np.testing.assert_array_equal(out, _first_min(f["cost"]))


@pytest.mark.parametrize("KY", [0, 1, 2, 3])
@pytest.mark.parametrize("KX", [0, 1, 2, 3])
def test_stereo_wta_windows(gpu, KY, KX):
    g = workloads.rng(200 + 4 * KY + KX)
    H, W = 41, 75
    L = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    R = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    # D = 20 one chunk, 70 two chunks (ragged), 200 two passes, 1 a single shift
    for s0, step, D in [(0, -1, 20), (-20, 1, 70), (0, -1, 200), (3, 1, 1)]:
        print(f"[wta] KY={KY} KX={KX} s0={s0} step={step} D={D}", flush=True)
        out = ops.stereo_wta_np(L, R, D, s0, step, KY, KX)
        np.testing.assert_array_equal(out, oracle.stereo_wta(L, R, D, s0, step, KY, KX))


@pytest.mark.parametrize("H,W", [(3, 5), (7, 8), (33, 31), (64, 96)])
def test_stereo_wta_small_and_ragged(gpu, H, W):
    g = workloads.rng(H * 1000 + W)
    L = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    R = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    print(f"[wta] {H}x{W}", flush=True)
    out = ops.stereo_wta_np(L, R, 48, 0, -1, 2, 2)
    np.testing.assert_array_equal(out, oracle.stereo_wta(L, R, 48, 0, -1, 2, 2))


def test_stereo_wta_pitched(gpu):
    from umi_amd import DeviceBuffer

    g = workloads.rng(17)
    H, W, P, D = 37, 70, 85, 64
    L = g.integers(0, 256, size=(H, P), dtype=np.int64).astype(np.uint8)
    R = g.integers(0, 256, size=(H, P), dtype=np.int64).astype(np.uint8)
    out = DeviceBuffer(H * W * 8)
    print("[wta] pitched", flush=True)
    ops.stereo_wta(DeviceBuffer.from_array(L), DeviceBuffer.from_array(R), H, W, P, D, 0, -1, 2, 2, out)
    got = out.download((H, W, 2), np.int32)
    np.testing.assert_array_equal(got, oracle.stereo_wta(L[:, :W], R[:, :W], D, 0, -1, 2, 2))


def test_stereo_wta_kitti_size(gpu):
    # 1242x375 pair, 128 disparities, 5x5 window: equal to the first minimum of
    # the (bit-exact) volume, and to the oracle on every row
    L, R, _ = workloads.stereo_pair(375, 1242, 128, seed=1)
    print("[wta] 375x1242 D=128", flush=True)
    out = ops.stereo_wta_np(L, R, 128, 0, -1, 2, 2)
    np.testing.assert_array_equal(out, _first_min(ops.cost_volume_np(L, R, 128, 0, -1, 2, 2)))
    np.testing.assert_array_equal(out, oracle.stereo_wta(L, R, 128, 0, -1, 2, 2))
