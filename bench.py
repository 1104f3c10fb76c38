#!/usr/bin/env python3
"""Benchmark of the UMI operator hot path on one MI355X.

Workload (one "step"): full-search block motion estimation of one
1920x1080 uint8 frame pair, 8x8 blocks, +-16 search (33x33 candidates),
int32 SAD, MV + SAD per block (32,400 blocks) -- the plain full-search ME
configuration.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line (metric, value, roofline, cpu_baseline, extra, ...).
`--gpus N` with N > 1 exits: the library and this benchmark use a single GPU.
`--only a,b` runs only the named secondary sections (see SECTIONS).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H, W, BLK, RAD = 1080, 1920, 8, 16
VALU_LANE_OPS = 256 * 64 * 2.4e9  # MI355X: 256 CUs x 64 lanes/clk x 2.4 GHz
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0, "fp8": 5000.0, "fp8x3": 5000.0}
CPU_THREADS_CAP = 16  # the host share of one GPU


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline sample length (all cores)")
    ap.add_argument("--no-cpu", action="store_true", help="skip every CPU leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary configurations")
    ap.add_argument("--only", default="", help="comma-separated secondary sections to run (default: all)")
    ap.add_argument("--warm-s", type=float, default=0.3,
                    help="seconds of back-to-back calls before a steady-state timing (GEMM, conv)")
    return ap.parse_args()


# ---------------------------------------------------------------- helpers
def timed(fn, st, n, w=3):
    """Mean ms per call of fn() over n calls, hipEvents on stream st."""
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


WARM_S = 0.3  # --warm-s


def timed_steady(fn, st, n, warm_s=None):
    """timed() after about warm_s seconds (default --warm-s) of back-to-back
    calls, so that the engine clock has settled at its sustained value for
    this load (MFMA-bound kernels run below the peak clock once the power
    limit is reached)."""
    warm_s = WARM_S if warm_s is None else warm_s
    ms = timed(fn, st, 3, 1)
    return timed(fn, st, n, max(3, int(warm_s * 1e3 / max(ms, 1e-3))))


def graph_median_us(run, st, per_graph=10, reps=21):
    """Median over `reps` timed replays of a graph holding `per_graph` calls of
    run(stream); returns microseconds per call."""
    from umi_amd import Graph, Timer

    run(st)
    st.synchronize()
    g = Graph(st)
    g.begin()
    for _ in range(per_graph):
        run(st)
    g.end()
    g.launch()
    st.synchronize()
    vals = []
    for _ in range(reps):
        t = Timer()
        t.start(st)
        g.launch()
        t.stop(st)
        st.synchronize()
        vals.append(t.ms() * 1e3 / per_graph)
    return statistics.median(vals)


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return {"nproc": n, "model": model, "threads_used": max(1, min(n, CPU_THREADS_CAP))}


def cpu_timed(fn, seconds, max_reps=1000):
    """Repeats fn() until `seconds` pass (at least once); returns (reps, elapsed s)."""
    t0 = time.perf_counter()
    reps = 0
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds or reps >= max_reps:
            return reps, el


def roofline(bound, achieved, peak, unit, traffic=None, **more):
    d = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
         "traffic": traffic}
    d.update(more)
    return d


# ---------------------------------------------------------------- headline
def me_sources_sha():
    """Fingerprint of the ME kernel's sources (the PMC summary records the one it was taken on)."""
    h = hashlib.sha1()
    for f in ("umi_amd/csrc/me.hip", "umi_amd/csrc/stage.h", "umi_amd/csrc/common.h"):
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic():
    """HBM bytes per launch of the ME kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel reference), used
    only when the summary was taken on the current ME sources."""
    p = os.path.join(ROOT, "profiles", "me_fullsearch_1080p_r16_pmc.json")
    try:
        with open(p) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, "no PMC summary"
    if d.get("sources_sha") != me_sources_sha():
        return None, "PMC summary predates the current ME sources"
    return d.get("traffic_bytes_per_launch"), f"{os.path.relpath(p, ROOT)} (sources {d['sources_sha']})"


def cpu_baseline_me(cur, ref, seconds):
    """The oracle (a C restatement of main/full_me.cu:66-87 plus argmin) over
    whole frames of the same pair: all host cores (OpenMP over block rows) for
    `seconds`, and one thread on a band of 16 block rows for half as long."""
    import oracle

    info = cpu_info()
    nt = info["threads_used"]
    mv = None
    try:
        oracle.set_threads(nt)
        out = {}

        def whole():
            out["mv"] = oracle.me_full_search(cur, ref, BLK, RAD)[0]

        frames, el = cpu_timed(whole, seconds)
        mv = out["mv"]
        oracle.set_threads(1)
        rows = 16
        bands, el1 = cpu_timed(lambda: oracle.me_full_search(cur, ref, BLK, RAD, by_range=(0, rows)), seconds / 2)
    finally:
        oracle.set_threads(1)
    nbx = W // BLK
    multi = frames * (H // BLK) * nbx / el
    single = bands * rows * nbx / el1
    return {"value": multi, "unit": "blocks/s", "cores": nt, "kind": "port",
            "sample": f"{frames} whole 1920x1080 frame pair(s) on {nt} OpenMP threads in {el:.1f} s; "
                      f"single thread: {bands} band(s) of {rows} block rows in {el1:.1f} s",
            "single_thread_value": single, "nproc": info["nproc"], "cpu_model": info["model"]}, mv


# ---------------------------------------------------------------- sections
def sec_subpel(st, g, cpu):
    """Quarter-pel refinement of the +-16 full search's vectors, 1080p."""
    from umi_amd import DeviceBuffer, ops, workloads

    cur, ref_f, _ = workloads.me_pair(1080, 1920, 16, seed=2)
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref_f)
    mvi, mvq = DeviceBuffer(32400 * 12), DeviceBuffer(32400 * 12)
    t_fs = timed(lambda: ops.me_full_search(dc, dr, 1080, 1920, 1920, 8, 16, mvi, stream=st), st, 50)
    t_sp = timed(lambda: ops.me_subpel_refine(dc, dr, 1080, 1920, 1920, mvi, mvq, stream=st), st, 50)
    q = mvq.download((135, 240, 3), np.int32)
    i = mvi.download((135, 240, 3), np.int32)
    out = {"us": t_sp * 1e3, "full_search_us": t_fs * 1e3, "blocks_per_s": 32400 / (t_sp * 1e-3),
           "mean_sad_integer": float(i[..., 2].mean()), "mean_sad_quarter_pel": float(q[..., 2].mean()),
           "blocks_moved": float(np.mean((q[..., 0] != 4 * i[..., 0]) | (q[..., 1] != 4 * i[..., 1])))}
    # NV12 1080p: luma + chroma full search (+-16) and the FFmpeg-record export
    cy, cuv, ry, ruv, _ = workloads.nv12_pair(1080, 1920, seed=2)
    planes = [DeviceBuffer.from_array(a) for a in (cy, cuv, ry, ruv)]
    mvn, rec = DeviceBuffer(32400 * 12), DeviceBuffer(32400 * 40)
    t_nv = timed(lambda: ops.me_full_search_nv12(*planes, 1080, 1920, 1920, 1920, 16, mvn, stream=st), st, 50)
    t_ex = timed(lambda: ops.export_mvs(mvq, 135, 240, 8, 4, ops.MV_AVMOTIONVECTOR, rec, stream=st), st, 50)
    absdiff = 32400 * 1089 * 96
    out["nv12_full_search_1080p_r16"] = {
        "us": t_nv * 1e3, "blocks_per_s": 32400 / (t_nv * 1e-3), "vs_luma_only": t_nv / t_fs,
        "roofline": roofline("valu", absdiff / (t_nv * 1e-3) / 1e12, VALU_LANE_OPS * 4 / 1e12, "T abs-diff/s")}
    out["mv_export_avmotionvector_us"] = t_ex * 1e3
    return out


def sec_gemm(st, g, cpu):
    """GEMM 4096^3 in five precision forms: error on 64 sampled rows against
    FP64, MFMA and HBM roofline per form, and the float loop of
    main/gemm.cu:69-78 on the host for 16 rows."""
    import oracle
    from umi_amd import DeviceBuffer, ops

    M = N = K = 4096
    A64, B64 = g.standard_normal((M, K)), g.standard_normal((K, N))
    A32, B32 = A64.astype(np.float32), B64.astype(np.float32)
    dA, dB = DeviceBuffer.from_array(A32), DeviceBuffer.from_array(B32)
    rows = np.arange(0, M, M // 64)
    ref = A64[rows] @ B64
    out = {}
    flops = 2 * M * N * K
    for mode in (ops.F32, ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3):
        name = ops.MODE_NAMES[mode]
        oa = ops.gemm_pack(dA, M, K, K, False, mode, stream=st)
        ob = ops.gemm_pack(dB, N, K, N, True, mode, stream=st)
        dC = DeviceBuffer(M * N * 4)
        ms = timed_steady(lambda: ops.gemm(mode, M, N, K, oa, ob, dC, stream=st), st, 50)
        C = dC.download((M, N), np.float32)[rows]
        e = oracle.contraction_error(C, ref, np.linalg.norm(A64[rows], axis=1), np.linalg.norm(B64, axis=0))
        eb = ops.elem_bytes(mode) * (2 if mode in (ops.BF16X3, ops.FP8X3) else 1)
        hbm = (M * K + N * K) * eb + M * N * 4
        tf = flops / (ms * 1e-3) / 1e12
        out[name] = {"ms": ms, "tflops": tf, **e,
                     "roofline": roofline("mfma", tf, MFMA_PEAK_TFLOPS[name], "TFLOP/s",
                                          products=3 if mode in (ops.BF16X3, ops.FP8X3) else 1,
                                          hbm_bytes=hbm, hbm_gbs=hbm / (ms * 1e-3) / 1e9,
                                          hbm_frac=hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)}
    if cpu:
        try:
            oracle.set_threads(cpu_info()["threads_used"])
            r = 16
            reps, el = cpu_timed(lambda: oracle.gemm_f32_rows(A32, B32, 0, r), 2.0, 5)
        finally:
            oracle.set_threads(1)
        cf = reps * 2 * r * N * K / el / 1e9
        out["cpu_baseline"] = {"value": cf, "unit": "GFLOP/s", "cores": cpu_info()["threads_used"], "kind": "port",
                               "sample": f"{reps} x {r} rows of 4096^3 by the f32 loop of main/gemm.cu:69-78, {el:.1f} s",
                               "extrapolated_s_per_gemm": flops / (cf * 1e9)}
    return out


def sec_conv(st, g, cpu):
    """Conv 16x64x56x56 -> 64, 3x3, stride 1, pad 1 in five forms against the
    FP64 oracle, with MFMA and HBM rooflines; the float loop of
    main/fprop.cu:84-108 on the host for one image."""
    import oracle
    from umi_amd import DeviceBuffer, ops

    x64, w64 = g.standard_normal((16, 64, 56, 56)), g.standard_normal((64, 64, 3, 3))
    ref = oracle.conv_f64(x64, w64, 1, 1)
    cols, _ = oracle.im2col_f64(x64, 3, 3, 1, 1)
    scale = np.linalg.norm(w64.reshape(64, -1), axis=1)[None, :, None, None] * \
        np.linalg.norm(cols, axis=1).reshape(16, 56, 56)[:, None]
    x32, w32 = x64.astype(np.float32), w64.astype(np.float32)
    dx, dw = DeviceBuffer.from_array(x32), DeviceBuffer.from_array(w32)
    out = {}
    flops = 2 * 16 * 64 * 56 * 56 * 64 * 9
    for mode in (ops.F32, ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3):
        name = ops.MODE_NAMES[mode]
        ox = ops.conv_pack_input(dx, 16, 64, 56, 56, mode, stream=st)
        ow = ops.conv_pack_filter(dw, 64, 64, 3, 3, mode, stream=st)
        o = DeviceBuffer(16 * 64 * 56 * 56 * 4)
        ms = timed_steady(lambda: ops.conv_fprop(mode, 16, 64, 56, 56, 64, 3, 3, 1, 1, ox, ow, o, stream=st), st, 100)
        rel = np.abs(o.download((16, 64, 56, 56), np.float32) - ref) / scale
        eb = ops.elem_bytes(mode) * (2 if mode in (ops.BF16X3, ops.FP8X3) else 1)
        hbm = (16 * 64 * 56 * 56 + 64 * 64 * 9) * eb + 16 * 64 * 56 * 56 * 4
        tf = flops / (ms * 1e-3) / 1e12
        gbs = hbm / (ms * 1e-3) / 1e9
        out[name] = {"us": ms * 1e3, "tflops": tf, "max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
                     "roofline": roofline("hbm", gbs, HBM_PEAK_GBS, "GB/s", algorithmic_bytes=hbm,
                                          mfma_tflops=tf, mfma_peak=MFMA_PEAK_TFLOPS[name],
                                          mfma_frac=tf / MFMA_PEAK_TFLOPS[name])}
    if cpu:
        try:
            oracle.set_threads(cpu_info()["threads_used"])
            reps, el = cpu_timed(lambda: oracle.conv_f32(x32, w32, 1, 1, 0, 1), 2.0, 5)
        finally:
            oracle.set_threads(1)
        cf = reps * flops / 16 / el / 1e9
        out["cpu_baseline"] = {"value": cf, "unit": "GFLOP/s", "cores": cpu_info()["threads_used"], "kind": "port",
                               "sample": f"{reps} x one of the 16 images by the f32 loop of main/fprop.cu:84-108, "
                                         f"{el:.1f} s", "extrapolated_s_per_batch": flops / (cf * 1e9)}
    return out


def sec_stereo(st, g, cpu):
    """Stereo cost volume 1242x375, 128 disparities, 5x5, and its fused
    winner-take-all; the oracle on a band of rows on the host."""
    import oracle
    from umi_amd import DeviceBuffer, ops, workloads

    L, R, _ = workloads.stereo_pair(375, 1242, 128, seed=1)
    dl, dr, cv = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R), DeviceBuffer(375 * 1242 * 128 * 4)
    ms = timed(lambda: ops.cost_volume(dl, dr, 375, 1242, 1242, 128, 0, -1, 2, 2, cv, stream=st), st, 20)
    vol_bytes = 375 * 1242 * 128 * 4 + 2 * 375 * 1242
    out = {"cost_volume_1242x375_d128_5x5": {
        "ms": ms, "gbs": vol_bytes / (ms * 1e-3) / 1e9,
        "roofline": roofline("hbm", vol_bytes / (ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s", algorithmic_bytes=vol_bytes)}}
    wta = DeviceBuffer(375 * 1242 * 8)
    us = timed(lambda: ops.stereo_wta(dl, dr, 375, 1242, 1242, 128, 0, -1, 2, 2, wta, stream=st), st, 50) * 1e3
    ok = bool(np.array_equal(wta.download((375, 1242, 2), np.int32)[..., 0],
                             np.argmin(cv.download((375, 1242, 128), np.int32), axis=-1)))
    absdiff = 375 * 1242 * 128 * 25
    ach = absdiff / (us * 1e-6) / 1e12
    out["stereo_wta_1242x375_d128_5x5"] = {
        "us": us, "vs_volume": ms * 1e3 / us, "gpx_s": 375 * 1242 / (us * 1e-6) / 1e9,
        "equals_volume_argmin": ok,
        "roofline": roofline("valu", ach, VALU_LANE_OPS * 4 / 1e12, "T abs-diff/s")}
    # semi-global matching, 4 paths, P1 = 10, P2 = 120: the uint16 partial sums
    # S [H][W][D] are written 3 times and read 3 times across the two launches
    ws = DeviceBuffer(ops.sgm_workspace_bytes(375, 1242, 128, 0, -1, 2, 2))
    sg = DeviceBuffer(375 * 1242 * 8)
    us_sgm = timed(lambda: ops.sgm(dl, dr, 375, 1242, 1242, 128, 0, -1, 2, 2, 10, 120, ws, sg, stream=st),
                   st, 20) * 1e3
    sgm_bytes = 6 * 375 * 1242 * 128 * 2 + 2 * 375 * 1242 + 375 * 1242 * 8
    out["sgm_1242x375_d128_5x5_4paths"] = {
        "us": us_sgm, "gpx_s": 375 * 1242 / (us_sgm * 1e-6) / 1e9,
        "roofline": roofline("hbm", sgm_bytes / (us_sgm * 1e-6) / 1e9, HBM_PEAK_GBS, "GB/s",
                             algorithmic_bytes=sgm_bytes)}
    if cpu:
        try:
            oracle.set_threads(cpu_info()["threads_used"])
            reps, el = cpu_timed(lambda: oracle.sgm(L, R, 128, 0, -1, 2, 2, 10, 120), 1.0, 3)
        finally:
            oracle.set_threads(1)
        out["sgm_1242x375_d128_5x5_4paths"]["cpu_baseline"] = {
            "value": reps * 375 * 1242 / el, "unit": "px/s", "cores": cpu_info()["threads_used"], "kind": "port",
            "sample": f"{reps} whole pairs (volume on all threads, the four paths on one), {el:.1f} s"}
    if cpu:
        rows = 8
        try:
            oracle.set_threads(cpu_info()["threads_used"])
            reps, el = cpu_timed(lambda: oracle.cost_volume(L, R, 128, 0, -1, 2, 2, y_range=(0, rows)), 2.0, 50)
        finally:
            oracle.set_threads(1)
        v = reps * rows * 1242 / el
        out["cost_volume_1242x375_d128_5x5"]["cpu_baseline"] = {
            "value": v, "unit": "px/s (128 shifts each)", "cores": cpu_info()["threads_used"], "kind": "port",
            "sample": f"{reps} x {rows} rows by the loop of main/cost_volume.cu:67-80 (5x5), {el:.1f} s",
            "gpu_px_s": 375 * 1242 / (ms * 1e-3)}
    return out


def sec_filters(st, g, cpu):
    """Weighted 3x3 filter on 1024^2 float32 (eager and graph replay), and
    the 1080p integral images (one launch) per tile height."""
    import oracle
    from umi_amd import DeviceBuffer, ops

    img = g.random((1024, 1024), dtype=np.float32)
    wt = g.normal(size=9)
    wts = (wt / wt.sum()).astype(np.float32)
    di, do = DeviceBuffer.from_array(img), DeviceBuffer(img.nbytes)
    dwt = DeviceBuffer.from_array(wts)
    run = lambda s: ops.wfilter3x3(di, 1024, 1024, 1024, dwt, 0.0, 0.0, 0, do, 1024, stream=s)
    us = timed(lambda: run(st), st, 100) * 1e3
    usg = graph_median_us(run, st)
    b = 8 * 1024 * 1024
    out = {"wfilter3x3_1024_f32": {"us": us, "graph_median_us": usg, "gbs": b / (us * 1e-6) / 1e9,
                                   "roofline": roofline("hbm", b / (usg * 1e-6) / 1e9, HBM_PEAK_GBS, "GB/s",
                                                        algorithmic_bytes=b, timing="graph replay median")}}
    if cpu:
        try:
            oracle.set_threads(cpu_info()["threads_used"])
            reps, el = cpu_timed(lambda: oracle.wfilter3x3(img, wts, 0.0, 0.0), 2.0, 500)
        finally:
            oracle.set_threads(1)
        out["wfilter3x3_1024_f32"]["cpu_baseline"] = {
            "value": reps * 1024 * 1024 / el / 1e6, "unit": "Mpx/s", "cores": cpu_info()["threads_used"],
            "kind": "port", "sample": f"{reps} images by the loop of main/wfilter3x3.cu:71-89 (fmaf), {el:.1f} s",
            "gpu_mpx_s": 1024 * 1024 / (usg * 1e-6) / 1e6}
    one8 = g.integers(0, 256, size=(1080, 1920), dtype=np.int64).astype(np.uint8)
    onef = g.random((1080, 1920), dtype=np.float32)
    ii = {}
    for dt, im, bpp in ((np.uint8, one8, 5), (np.float32, onef, 8)):
        d = DeviceBuffer.from_array(im)
        o = DeviceBuffer(1080 * 1920 * 4)
        ws = ops.iimg_workspace(1080, 1920, dt)
        per = {}
        for rpl in (1, 2, 4):
            ops.iimg_set_config(rpl)
            per[rpl] = graph_median_us(lambda s: ops.iimg(d, 1080, 1920, 1920, o, 1920, dt, ws, stream=s), st)
        ops.iimg_set_config(0)
        best = min(per, key=per.get)
        ii[np.dtype(dt).name] = {"us_by_rows_per_lane": per, "best_rows_per_lane": best, "us": per[best],
                                 "gbs": bpp * 1080 * 1920 / (per[best] * 1e-6) / 1e9}
    out["iimg_1080p_one_launch"] = ii
    return out


def sec_batch(st, g, cpu):
    """256 uneven images (sides log-uniform in [64, 1024]) in graph-captured
    batched launches, each operator at the tile configuration the autotuner
    picks; box sums and rectangle features from the integral images."""
    import oracle
    from umi_amd import DeviceBuffer, batch, ops, workloads

    imgs, u8 = workloads.uneven_batch(256, 64, 1024, seed=1)
    bf, bu = batch.ImageBatch(imgs), batch.ImageBatch(u8, out_dtype=np.uint32)
    px = bf.total
    wt = g.normal(size=9)
    dwt = DeviceBuffer.from_array((wt / wt.sum()).astype(np.float32))
    wsd = DeviceBuffer.from_array(ops.spatial_weights(2, 2.0))
    wsu, wsf = bu.iimg_workspace(), bf.iimg_workspace()
    tune = {
        "wfilter": batch.autotune(batch.OP_FILTER, bf, lambda c, s: bf.wfilter3x3(dwt, 0.0, 0.0, 0, c, s),
                                  use_cache=False),
        "bilateral": batch.autotune(batch.OP_BILATERAL, bf, lambda c, s: bf.bilateral(2, wsd, -50.0, 0.0, 0, c, s),
                                    use_cache=False),
        "iimg_u8": batch.autotune(batch.OP_IIMG, bu, lambda c, s: bu.iimg(wsu, s, config=c), use_cache=False),
        "iimg_f32": batch.autotune(batch.OP_IIMG, bf, lambda c, s: bf.iimg(wsf, s, config=c), use_cache=False),
    }
    t = {k: v[1][v[0]] for k, v in tune.items()}  # ms at the picked configuration
    bpp = {"wfilter": 8, "bilateral": 8, "iimg_u8": 5, "iimg_f32": 8}
    out = {"pixels": px, "tuned": {k: {"config": v[0], "ms_by_config": v[1]} for k, v in tune.items()}}
    for k in t:
        gbs = bpp[k] * px / (t[k] * 1e-3) / 1e9
        out[k] = {"ms": t[k], "gbs": gbs, "gpx_s": px / (t[k] * 1e-3) / 1e9,
                  "roofline": roofline("hbm", gbs, HBM_PEAK_GBS, "GB/s", algorithmic_bytes=bpp[k] * px)}
    # box sums read from the integral images (9x9 windows): one 1080p image and the batch
    bu.iimg(wsu, st, config=tune["iimg_u8"][0])
    bbox = DeviceBuffer(px * 4)
    t_bx = batch.time_graph(lambda s: bu.box_sum(4, 4, bbox, s), st)
    one = g.integers(0, 256, size=(1080, 1920), dtype=np.int64).astype(np.uint8)
    d1, s1, o1 = DeviceBuffer.from_array(one), DeviceBuffer(1080 * 1920 * 4), DeviceBuffer(1080 * 1920 * 4)
    ws1 = ops.iimg_workspace(1080, 1920, np.uint8)
    ops.iimg(d1, 1080, 1920, 1920, s1, 1920, np.uint8, ws1, stream=st)
    us1 = timed(lambda: ops.box_sum(s1, 1080, 1920, 1920, 4, 4, o1, 1920, stream=st), st, 100) * 1e3
    out["box_sum_from_iimg_r4"] = {"one_1080p_us": us1, "one_1080p_gbs": 8 * 1080 * 1920 / (us1 * 1e-6) / 1e9,
                                   "batch_ms": t_bx, "batch_gbs": 8 * px / (t_bx * 1e-3) / 1e9}
    # 16 rectangle (Haar-like) features of 2-3 rectangles, 24x24 window, on the 1080p integral image
    gf = np.random.default_rng(24)
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
    if cpu:
        nt = cpu_info()["threads_used"]
        sub = list(range(0, 256, 8))  # 32 of the 256 images
        spx = sum(imgs[i].size for i in sub)
        w9 = (wt / wt.sum()).astype(np.float32)
        ws_h = ops.spatial_weights(2, 2.0)
        legs = {}
        try:
            oracle.set_threads(nt)
            for k, fn in (("wfilter", lambda: [oracle.wfilter3x3(imgs[i], w9, 0.0, 0.0) for i in sub]),
                          ("bilateral", lambda: [oracle.bilateral(imgs[i], 2, ws_h, -50.0, 0.0, 0) for i in sub])):
                reps, el = cpu_timed(fn, 1.0, 50)
                legs[k] = {"value": reps * spx / el / 1e6, "unit": "Mpx/s", "cores": nt, "kind": "port",
                           "sample": f"{reps} x 32 of the 256 images ({spx} px), {el:.1f} s"}
        finally:
            oracle.set_threads(1)
        reps, el = cpu_timed(lambda: [oracle.iimg(u8[i]) for i in sub], 1.0, 50)
        legs["iimg_u8"] = {"value": reps * spx / el / 1e6, "unit": "Mpx/s", "cores": 1, "kind": "port",
                           "sample": f"{reps} x 32 of the 256 images by the loop of main/iimg.cu:84-112, {el:.1f} s"}
        out["cpu_baseline"] = legs
    return out


def sec_engine(st, g, cpu):
    """The generic UMI::Execute<> engine on the reference examples' strategies
    and descriptors at larger sizes (tests/cpp/bin/engine_bench, its own process)."""
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "engine_bench")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    for d in rows:
        d["hbm_frac"] = d["gbs"] / HBM_PEAK_GBS
    return {"rc": r.returncode, "cases": rows}


def sec_proposer(st, g, cpu):
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
    with_net = proposer.evaluate(prop, pairs, seed0, proposals=True, reps=20, graphs=True)
    without = proposer.evaluate(prop, pairs, seed0, proposals=False, reps=20)
    concurrent = proposer.throughput(net, 4, seed0, reps=20, gate=150)
    concurrent6 = proposer.throughput(net, 6, seed0, reps=20, gate=150)
    return {"pairs": pairs, "first_seed": seed0, "search": "+-32, 8x8 blocks, 32400 blocks per pair",
            "settings": {"iters": prop.iters, "local_rounds": prop.local_rounds, "rounds": prop.rounds,
                         "slack": prop.slack},
            "blocks_per_s_pipeline": prop.nblk / (with_net["ms_pipeline"] * 1e-3),
            "blocks_per_s_full_search": prop.nblk / (with_net["ms_full_search"] * 1e-3),
            "with_network": with_net, "without_network": without, "four_streams": concurrent,
            "six_streams": concurrent6,
            "second_set_6_pairs_from_3000000": proposer.evaluate(prop, 6, 3_000_000, proposals=True, reps=20,
                                                                       graphs=True)}


SECTIONS = {"subpel": sec_subpel, "gemm": sec_gemm, "conv": sec_conv, "stereo": sec_stereo,
            "filters": sec_filters, "batch": sec_batch, "engine": sec_engine, "proposer": sec_proposer}


def extra_configs(st, only, cpu):
    """Secondary workloads, each on its own data (none enters `value`)."""
    from umi_amd import workloads

    g = workloads.rng(3)
    out = {}
    for name, fn in SECTIONS.items():
        if only and name not in only:
            continue
        t0 = time.perf_counter()
        r = fn(st, g, cpu)
        if r is not None:
            out[name] = r
            if isinstance(r, dict):
                r["section_wall_s"] = time.perf_counter() - t0
        print(f"[bench] section {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return out


def main():
    a = parse()
    global WARM_S
    WARM_S = a.warm_s
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {a.gpus}); the frame-band sharding of the "
              "search across GPUs is a design, not run here.", file=sys.stderr)
        sys.exit(2)

    import umi_amd
    from umi_amd import DeviceBuffer, HostBuffer, Stream, Timer, ops, workloads

    if umi_amd.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    cur, ref, _ = workloads.me_pair(H, W, RAD, seed=a.seed)
    nby, nbx = H // BLK, W // BLK
    nblocks = nby * nbx
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    mv = DeviceBuffer(nblocks * 3 * 4)
    st = Stream()
    from umi_amd._lib import call

    n = 256 << 20
    ca, cb = DeviceBuffer(n), DeviceBuffer(n)
    copy_ms = timed(lambda: call("umi_copy_d2d", cb.ptr, ca.ptr, n, st.handle), st, 20)
    copy_gbs = 2 * n / (copy_ms * 1e-3) / 1e9
    del ca, cb
    step = lambda s: ops.me_full_search(dc, dr, H, W, W, BLK, RAD, mv, stream=s)
    for _ in range(a.warmup):
        step(st)
    st.synchronize()
    umi_amd.synchronize()
    t = Timer()
    w0 = time.perf_counter()
    t.start(st)
    for _ in range(a.steps):
        step(st)
    t.stop(st)
    st.synchronize()
    umi_amd.synchronize()
    wall = time.perf_counter() - w0
    ms = t.ms() / a.steps
    gpu_mv = mv.download((nby, nbx, 3), np.int32)
    graph_us = graph_median_us(step, st)
    # host buffers in, motion vectors out: the PCIe-inclusive rate (never `value`)
    hc, hr, hm = HostBuffer(H * W), HostBuffer(H * W), HostBuffer(nblocks * 12)
    hc.array((H, W), np.uint8)[:] = cur
    hr.array((H, W), np.uint8)[:] = ref

    def with_copies():
        call("umi_memcpy_async", dc.ptr, hc.ptr, H * W, 1, st.handle)
        call("umi_memcpy_async", dr.ptr, hr.ptr, H * W, 1, st.handle)
        step(st)
        call("umi_memcpy_async", hm.ptr, mv.ptr, nblocks * 12, 2, st.handle)

    pcie_ms = timed(with_copies, st, 20)
    assert np.array_equal(hm.array((nby, nbx, 3), np.int32), gpu_mv)

    # roofline: algorithmic work per launch = blocks x (2R+1)^2 candidates x B^2 |a-b|
    absdiff = nblocks * (2 * RAD + 1) ** 2 * BLK * BLK
    achieved = absdiff / (ms * 1e-3) / 1e12            # T abs-diff/s
    peak = VALU_LANE_OPS * 4 / 1e12                    # v_sad_u8: 4 abs-diff per lane-op
    in_bytes = 2 * H * W
    out_bytes = nblocks * 3 * 4
    traffic, traffic_src = measured_traffic()

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
                     "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": in_bytes + out_bytes,
                     "hbm_gbs": (in_bytes + out_bytes) / (ms * 1e-3) / 1e9, "hbm_peak_gbs": HBM_PEAK_GBS,
                     "hbm_copy_probe_gbs": copy_gbs},
        "graph_replay_median_us": graph_us,
        "pcie_inclusive": {"ms": pcie_ms, "blocks_per_s": nblocks / (pcie_ms * 1e-3),
                           "what": "pinned host frames in, ME, motion vectors out, one stream"},
        "wall_s_timed_region": wall,
        "device": umi_amd.device_name(),
    }
    if not a.no_extra:
        only = {s for s in a.only.split(",") if s}
        line["extra"] = extra_configs(st, only, not a.no_cpu)
    if not a.no_cpu:
        cb_, cpu_mv = cpu_baseline_me(cur, ref, a.cpu_seconds)
        cb_["gpu_over_cpu"] = line["value"] / cb_["value"]
        line["cpu_baseline"] = cb_
        line["parity_vs_oracle"] = bool(np.array_equal(cpu_mv, gpu_mv))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
