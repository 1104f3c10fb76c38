#!/usr/bin/env python3
"""Benchmark of the UMI operator hot path on one MI355X.

Workload (one "step"): full-search block motion estimation of one
1920x1080 uint8 frame pair, 8x8 blocks, +-16 search (33x33 candidates),
int32 SAD, MV + SAD per block (32,400 blocks) — the plain full-search ME
configuration.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line (metric, value, roofline, cpu_baseline, ...).
`--gpus N` with N > 1 exits: the library and this benchmark use a single GPU.
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
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary configurations")
    return ap.parse_args()


def timed(fn, st, n, w=3):
    from umi_amd import Timer

    for _ in range(w):
        fn()
    st.synchronize()
    t = Timer()
    t.start(st)
    for _ in range(n):
        fn()
    t.stop(st)
    st.synchronize()
    return t.ms() / n


def hbm_copy_probe(st, mb=256):
    """Streaming copy kernel of known size: the measured HBM roof, and the byte
    reference for the FETCH_SIZE / WRITE_SIZE counters of profiling runs."""
    from umi_amd import DeviceBuffer
    from umi_amd._lib import call

    n = mb << 20
    a, b = DeviceBuffer(n), DeviceBuffer(n)
    ms = timed(lambda: call("umi_copy_d2d", b.ptr, a.ptr, n, st.handle), st, 20)
    return {"bytes": 2 * n, "ms": ms, "gbs": 2 * n / (ms * 1e-3) / 1e9}


def learned_proposer(st):
    """Learned-proposal ME on held-out 1920x1080 pairs (+-32 search, 8x8
    blocks; smooth motion from 4 Gaussian-weighted affine bumps, |mv| <= 32,
    noise sigma 2): the committed network, its refinement and fallbacks,
    against the plain +-32 full search; the same refinement from zero
    proposals shows what the network adds."""
    from umi_amd import proposer

    if not proposer.CHECKPOINT.exists():
        return None
    net = proposer.Network.load(proposer.CHECKPOINT)
    prop = proposer.Proposer(net)
    pairs, seed0 = 3, 2_000_000
    with_net = proposer.evaluate(prop, pairs, seed0, proposals=True, reps=20)
    without = proposer.evaluate(prop, pairs, seed0, proposals=False, reps=20)
    # four and six pairs in flight, one stream each (graph-captured), both methods alike
    concurrent = proposer.throughput(net, 4, seed0, reps=20, gate=150)
    concurrent6 = proposer.throughput(net, 6, seed0, reps=20, gate=150)
    return {"pairs": pairs, "first_seed": seed0, "search": "+-32, 8x8 blocks, 32400 blocks per pair",
            "settings": {"iters": prop.iters, "local_rounds": prop.local_rounds, "rounds": prop.rounds,
                         "slack": prop.slack},
            "blocks_per_s_pipeline": prop.nblk / (with_net["ms_pipeline"] * 1e-3),
            "blocks_per_s_full_search": prop.nblk / (with_net["ms_full_search"] * 1e-3),
            "with_network": with_net, "without_network": without, "four_streams": concurrent,
            "six_streams": concurrent6,
            # a second held-out set, one pair at a time (the ratio moves with the pairs' motion)
            "second_set_6_pairs_from_3000000": proposer.evaluate(prop, 6, 3_000_000, proposals=True, reps=20)}


def extra_configs(st):
    """Secondary workloads, each on its own data (not part of `value`): GEMM
    4096^3 and conv 16x64x56x56 -> 64, 3x3, in five precision forms with FP64
    error; stereo cost volume 1242x375, 128 disparities, 5x5; weighted 3x3
    filter on 1024^2 float32; 256 uneven images in graph-captured batched
    launches; learned-proposal motion estimation on 1920x1080 pairs."""
    import oracle
    from umi_amd import DeviceBuffer, batch, ops, workloads

    out = {}
    g = workloads.rng(3)
    # --- quarter-pel refinement of the +-16 full search's vectors, 1080p
    cur, ref_f, _ = workloads.me_pair(1080, 1920, 16, seed=2)
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref_f)
    mvi, mvq = DeviceBuffer(32400 * 12), DeviceBuffer(32400 * 12)
    t_fs = timed(lambda: ops.me_full_search(dc, dr, 1080, 1920, 1920, 8, 16, mvi, stream=st), st, 50)
    t_sp = timed(lambda: ops.me_subpel_refine(dc, dr, 1080, 1920, 1920, mvi, mvq, stream=st), st, 50)
    q = mvq.download((135, 240, 3), np.int32)
    i = mvi.download((135, 240, 3), np.int32)
    out["subpel_refine_1080p_after_r16"] = {
        "us": t_sp * 1e3, "full_search_us": t_fs * 1e3, "blocks_per_s": 32400 / (t_sp * 1e-3),
        "mean_sad_integer": float(i[..., 2].mean()), "mean_sad_quarter_pel": float(q[..., 2].mean()),
        "blocks_moved": float(np.mean((q[..., 0] != 4 * i[..., 0]) | (q[..., 1] != 4 * i[..., 1])))}
    # --- GEMM 4096^3 (error on 64 sampled rows against FP64)
    M = N = K = 4096
    A64, B64 = g.standard_normal((M, K)), g.standard_normal((K, N))
    dA, dB = DeviceBuffer.from_array(A64.astype(np.float32)), DeviceBuffer.from_array(B64.astype(np.float32))
    rows = np.arange(0, M, M // 64)
    ref = A64[rows] @ B64
    gemm = {}
    for mode in (ops.F32, ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3):
        oa = ops.gemm_pack(dA, M, K, K, False, mode, stream=st)
        ob = ops.gemm_pack(dB, N, K, N, True, mode, stream=st)
        dC = DeviceBuffer(M * N * 4)
        ms = timed(lambda: ops.gemm(mode, M, N, K, oa, ob, dC, stream=st), st, 10)
        C = dC.download((M, N), np.float32)[rows]
        e = oracle.contraction_error(C, ref, np.linalg.norm(A64[rows], axis=1), np.linalg.norm(B64, axis=0))
        gemm[ops.MODE_NAMES[mode]] = {"ms": ms, "tflops": 2 * M * N * K / (ms * 1e-3) / 1e12, **e}
    out["gemm_4096"] = gemm
    # --- conv 16x64x56x56 -> 64, 3x3 (full FP64 oracle)
    x64, w64 = g.standard_normal((16, 64, 56, 56)), g.standard_normal((64, 64, 3, 3))
    ref = oracle.conv_f64(x64, w64, 1, 1)
    cols, _ = oracle.im2col_f64(x64, 3, 3, 1, 1)
    scale = np.linalg.norm(w64.reshape(64, -1), axis=1)[None, :, None, None] * \
        np.linalg.norm(cols, axis=1).reshape(16, 56, 56)[:, None]
    dx, dw = DeviceBuffer.from_array(x64.astype(np.float32)), DeviceBuffer.from_array(w64.astype(np.float32))
    conv = {}
    flops = 2 * 16 * 64 * 56 * 56 * 64 * 9
    for mode in (ops.F32, ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3):
        ox = ops.conv_pack_input(dx, 16, 64, 56, 56, mode, stream=st)
        ow = ops.conv_pack_filter(dw, 64, 64, 3, 3, mode, stream=st)
        o = DeviceBuffer(16 * 64 * 56 * 56 * 4)
        ms = timed(lambda: ops.conv_fprop(mode, 16, 64, 56, 56, 64, 3, 3, 1, 1, ox, ow, o, stream=st), st, 20)
        rel = np.abs(o.download((16, 64, 56, 56), np.float32) - ref) / scale
        conv[ops.MODE_NAMES[mode]] = {"us": ms * 1e3, "tflops": flops / (ms * 1e-3) / 1e12,
                                      "max_rel": float(rel.max()), "mean_rel": float(rel.mean())}
    out["conv_16x64x56x56_k64_3x3"] = conv
    # --- stereo cost volume
    L, R, _ = workloads.stereo_pair(375, 1242, 128, seed=1)
    dl, dr, cv = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R), DeviceBuffer(375 * 1242 * 128 * 4)
    ms = timed(lambda: ops.cost_volume(dl, dr, 375, 1242, 1242, 128, 0, -1, 2, 2, cv, stream=st), st, 20)
    out["cost_volume_1242x375_d128_5x5"] = {"ms": ms, "gbs": (375 * 1242 * 128 * 4 + 2 * 375 * 1242) / (ms * 1e-3) / 1e9}
    # --- stereo winner-take-all fused with the cost volume (the volume never reaches HBM);
    # VALU-bound: 25 |a-b| per (pixel, shift) counted as the 5x5 window defines them
    wta = DeviceBuffer(375 * 1242 * 8)
    us = timed(lambda: ops.stereo_wta(dl, dr, 375, 1242, 1242, 128, 0, -1, 2, 2, wta, stream=st), st, 50) * 1e3
    ok = bool(np.array_equal(wta.download((375, 1242, 2), np.int32)[..., 0],
                             np.argmin(cv.download((375, 1242, 128), np.int32), axis=-1)))
    absdiff = 375 * 1242 * 128 * 25
    out["stereo_wta_1242x375_d128_5x5"] = {
        "us": us, "vs_volume": ms * 1e3 / us, "gpx_s": 375 * 1242 / (us * 1e-6) / 1e9,
        "t_absdiff_s": absdiff / (us * 1e-6) / 1e12, "valu_roof_t_absdiff_s": VALU_LANE_OPS * 4 / 1e12,
        "frac": absdiff / (us * 1e-6) / (VALU_LANE_OPS * 4), "equals_volume_argmin": ok}
    # --- weighted 3x3 filter
    img = g.random((1024, 1024), dtype=np.float32)
    wt = g.normal(size=9)
    di, do = DeviceBuffer.from_array(img), DeviceBuffer(img.nbytes)
    dwt = DeviceBuffer.from_array((wt / wt.sum()).astype(np.float32))
    ms = timed(lambda: ops.wfilter3x3(di, 1024, 1024, 1024, dwt, 0.0, 0.0, 0, do, 1024, stream=st), st, 100)
    out["wfilter3x3_1024_f32"] = {"us": ms * 1e3, "gbs": 8 * 1024 * 1024 / (ms * 1e-3) / 1e9}
    # --- 256 uneven images, graph-captured batched launches
    imgs, u8 = workloads.uneven_batch(256, 64, 1024, seed=1)
    bf, bu = batch.ImageBatch(imgs), batch.ImageBatch(u8, out_dtype=np.uint32)
    px = bf.total
    wsd = DeviceBuffer.from_array(ops.spatial_weights(2, 2.0))
    wsu, wsf = bu.iimg_workspace(), bf.iimg_workspace()
    t_wf = batch.time_graph(lambda s: bf.wfilter3x3(dwt, 0.0, 0.0, 0, 0, s), st)
    t_bl = batch.time_graph(lambda s: bf.bilateral(2, wsd, -50.0, 0.0, 0, 1, s), st)
    t_iu = batch.time_graph(lambda s: bu.iimg(wsu, s), st)
    t_if = batch.time_graph(lambda s: bf.iimg(wsf, s), st)
    out["batch_256_uneven"] = {"pixels": px, "wfilter_ms": t_wf, "wfilter_gbs": 8 * px / (t_wf * 1e-3) / 1e9,
                               "bilateral_ms": t_bl, "bilateral_gpx_s": px / (t_bl * 1e-3) / 1e9,
                               "iimg_u8_ms": t_iu, "iimg_u8_gbs": 5 * px / (t_iu * 1e-3) / 1e9,
                               "iimg_f32_ms": t_if, "iimg_f32_gbs": 8 * px / (t_if * 1e-3) / 1e9}
    # --- box sums read from the integral images (9x9 windows): one 1080p image and the batch
    bu.iimg(wsu, st)
    bbox = DeviceBuffer(px * 4)
    t_bx = batch.time_graph(lambda s: bu.box_sum(4, 4, bbox, s), st)
    one = g.integers(0, 256, size=(1080, 1920), dtype=np.int64).astype(np.uint8)
    d1, s1, o1 = DeviceBuffer.from_array(one), DeviceBuffer(1080 * 1920 * 4), DeviceBuffer(1080 * 1920 * 4)
    ws1 = ops.iimg_workspace(1080, 1920, np.uint8)
    ops.iimg(d1, 1080, 1920, 1920, s1, 1920, np.uint8, ws1, stream=st)
    us1 = timed(lambda: ops.box_sum(s1, 1080, 1920, 1920, 4, 4, o1, 1920, stream=st), st, 100) * 1e3
    out["box_sum_from_iimg_r4"] = {"one_1080p_us": us1, "one_1080p_gbs": 8 * 1080 * 1920 / (us1 * 1e-6) / 1e9,
                                   "batch_ms": t_bx, "batch_gbs": 8 * px / (t_bx * 1e-3) / 1e9}
    # --- 16 rectangle (Haar-like) features of 2-3 rectangles, 24x24 window, on the same 1080p integral image
    gf = workloads.rng(24)
    feats = np.zeros((16, 3, 5), np.int32)
    for i in range(16):
        for k in range(2 + i % 2):
            w_, h_ = int(gf.integers(2, 25)), int(gf.integers(2, 25))
            feats[i, k] = (int(gf.integers(0, 25 - w_)), int(gf.integers(0, 25 - h_)), w_, h_, (-1, 2, -1)[k])
    npos = (1080 - 23) * (1920 - 23)
    hf = DeviceBuffer(16 * npos * 4)
    us_h = timed(lambda: ops.haar_features(s1, 1080, 1920, 1920, feats, 24, 24, hf, stream=st), st, 50) * 1e3
    out["haar_16_features_1080p_24x24"] = {"us": us_h, "g_feature_values_s": 16 * npos / (us_h * 1e-6) / 1e9,
                                           "output_gbs": 16 * npos * 4 / (us_h * 1e-6) / 1e9}
    lp = learned_proposer(st)
    if lp is not None:
        out["learned_proposal_me_1080p_r32"] = lp
    return out


def measured_traffic():
    """HBM bytes per launch of the ME kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel reference)."""
    p = os.path.join(ROOT, "profiles", "me_fullsearch_1080p_r16_pmc.json")
    try:
        with open(p) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


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
        print(f"bench.py: measures one GPU only (asked for --gpus {a.gpus}); the frame-band sharding of the "
              "search across GPUs is a design, not run here.", file=sys.stderr)
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
    copy = hbm_copy_probe(st)
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
                     "frac": achieved / peak, "traffic": measured_traffic(),
                     "algorithmic_bytes": in_bytes + out_bytes,
                     "hbm_gbs": (in_bytes + out_bytes) / (ms * 1e-3) / 1e9, "hbm_peak_gbs": HBM_PEAK_GBS,
                     "hbm_copy_probe_gbs": copy["gbs"]},
        "wall_s_timed_region": wall,
        "device": umi_amd.device_name(),
    }
    if not a.no_extra:
        line["extra"] = extra_configs(st)
    if not a.no_cpu:
        cb, cpu_mv = cpu_baseline(cur, ref, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["parity_vs_oracle"] = bool(np.array_equal(cpu_mv, gpu_mv))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
