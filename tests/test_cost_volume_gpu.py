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
