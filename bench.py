#!/usr/bin/env python3
"""Benchmark of the UMI operator hot path on one MI355X.

Workload (one "step"): full-search block motion estimation of one
1920x1080 uint8 frame pair, 8x8 blocks, +-16 search (33x33 candidates),
int32 SAD, MV + SAD per block (32,400 blocks) — the plain full-search ME
configuration.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line (metric, value, roofline, cpu_baseline, ...).
`--gpus N` with N > 1 exits: this build runs on a single GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H, W, BLK, RAD = 1080, 1920, 8, 16
VALU_LANE_OPS = 256 * 64 * 2.4e9  # MI355X: 256 CUs x 64 lanes/clk x 2.4 GHz
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def cpu_baseline(cur, ref, seconds):
    """The oracle (a C restatement of main/full_me.cu:66-87 plus argmin), one
    thread, timed over whole frames of the same pair until `seconds` pass."""
    import oracle

    t0 = time.perf_counter()
    frames = 0
    mv = None
    while True:
        mv, _ = oracle.me_full_search(cur, ref, BLK, RAD)
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    blocks = frames * (H // BLK) * (W // BLK)
    return {"value": blocks / el, "unit": "blocks/s", "cores": 1, "kind": "port",
            "sample": f"{frames} full 1920x1080 frame pair(s), +-16, 8x8 blocks, {el:.1f} s, single thread"}, mv


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)

    import umi_amd
    from umi_amd import DeviceBuffer, Stream, Timer, ops, workloads

    if umi_amd.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    cur, ref, _ = workloads.me_pair(H, W, RAD, seed=a.seed)
    nby, nbx = H // BLK, W // BLK
    nblocks = nby * nbx
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    mv = DeviceBuffer(nblocks * 3 * 4)
    st = Stream()
    for _ in range(a.warmup):
        ops.me_full_search(dc, dr, H, W, W, BLK, RAD, mv, stream=st)
    st.synchronize()
    umi_amd.synchronize()
    t = Timer()
    w0 = time.perf_counter()
    t.start(st)
    for _ in range(a.steps):
        ops.me_full_search(dc, dr, H, W, W, BLK, RAD, mv, stream=st)
    t.stop(st)
    st.synchronize()
    umi_amd.synchronize()
    wall = time.perf_counter() - w0
    ms = t.ms() / a.steps
    gpu_mv = mv.download((nby, nbx, 3), np.int32)

    # roofline: algorithmic work per launch = blocks x (2R+1)^2 candidates x B^2 |a-b|
    absdiff = nblocks * (2 * RAD + 1) ** 2 * BLK * BLK
    achieved = absdiff / (ms * 1e-3) / 1e12            # T abs-diff/s
    peak = VALU_LANE_OPS * 4 / 1e12                    # v_sad_u8: 4 abs-diff per lane-op
    in_bytes = 2 * H * W
    out_bytes = nblocks * 3 * 4

    line = {
        "metric": "ME blocks/s",
        "value": nblocks / (ms * 1e-3),
        "unit": "blocks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded smoothed-noise frame pair, per-16x16-block motion in [-16,16], noise sigma 2)",
        "config": {"workload": "full-search ME 1920x1080 u8, 8x8 blocks, +-16 (33x33), int32 SAD, MV+SAD out",
                   "blocks": nblocks, "candidates_per_block": (2 * RAD + 1) ** 2, "seed": a.seed},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "T abs-diff/s",
                     "frac": achieved / peak, "traffic": None,
                     "hbm_gbs": (in_bytes + out_bytes) / (ms * 1e-3) / 1e9, "hbm_peak_gbs": HBM_PEAK_GBS},
        "wall_s_timed_region": wall,
        "device": umi_amd.device_name(),
    }
    if not a.no_cpu:
        cb, cpu_mv = cpu_baseline(cur, ref, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["parity_vs_oracle"] = bool(np.array_equal(cpu_mv, gpu_mv))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
