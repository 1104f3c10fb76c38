"""Batched operators over images of uneven sizes (one persistent launch per operator).

Every batched result equals the one-image kernel's result on each image
(same tile body, so bit-exact for all tile configurations) and the oracle;
graph-captured replays equal the eager result on every replay.
"""
import numpy as np
import pytest

import oracle
from umi_amd import DeviceBuffer, Graph, Stream, batch, ops, workloads

pytestmark = pytest.mark.gpu


def small_batch(seed=3, n=24):
    return workloads.uneven_batch(n=n, lo=40, hi=300, seed=seed)


def weights():
    g = workloads.rng(77)
    w = g.normal(size=9)
    return (w / w.sum()).astype(np.float32)


@pytest.mark.parametrize("config", range(4))
def test_batch_wfilter_equals_single(gpu, config):
    imgs, _ = small_batch()
    b = batch.ImageBatch(imgs)
    w = weights()
    b.wfilter3x3(DeviceBuffer.from_array(w), 0.0, 0.0, 0, config)
    print(f"[batch] wfilter config {config}", flush=True)
    for im, got in zip(imgs, b.outputs()):
        np.testing.assert_array_equal(got, ops.wfilter3x3_np(im, w, 0.0, 0.0))
    np.testing.assert_array_equal(b.outputs()[0], oracle.wfilter3x3(imgs[0], w, 0.0, 0.0))


@pytest.mark.parametrize("config", range(4))
def test_batch_bilateral_equals_single(gpu, config):
    imgs, _ = small_batch(seed=4)
    b = batch.ImageBatch(imgs)
    ws = ops.spatial_weights(2, 2.0)
    norm = float(np.float32(-1.0 / (2 * 0.1 ** 2)))
    b.bilateral(2, DeviceBuffer.from_array(ws), norm, 0.0, 0, config)
    print(f"[batch] bilateral config {config}", flush=True)
    for im, got in zip(imgs, b.outputs()):
        np.testing.assert_array_equal(got, ops.bilateral_np(im, 2, 2.0, 0.1))


@pytest.mark.parametrize("config", range(3))
def test_batch_iimg_u8_and_f32(gpu, config):
    imgs, u8 = small_batch(seed=5)
    bu = batch.ImageBatch(u8, out_dtype=np.uint32)
    bu.iimg(bu.iimg_workspace(), config=config)
    for im, got in zip(u8, bu.outputs()):
        np.testing.assert_array_equal(got, oracle.iimg(im))
    bf = batch.ImageBatch(imgs)
    bf.iimg(bf.iimg_workspace(), config=config)
    for im, got in zip(imgs, bf.outputs()):
        ref = oracle.iimg(im)
        assert np.max(np.abs(got - ref) / ref) <= 4e-6


def test_full_uneven_batch_all_ops_in_graphs(gpu, tmp_path, monkeypatch):
    # 256 float32 images, sides log-uniform in [64, 1024]: bilateral 5x5 (sigma_s 2, sigma_r 0.1),
    # wfilter3x3 and integral images of uint8 and float copies, each op one graph-captured launch
    # at the tile configuration the autotuner picks for this batch
    monkeypatch.setenv("UMI_TUNE_CACHE", str(tmp_path / "tune.json"))
    imgs, u8 = workloads.uneven_batch(256, 64, 1024, seed=1)
    print(f"[batch] {sum(i.size for i in imgs)} px in 256 images", flush=True)
    bf, bu = batch.ImageBatch(imgs), batch.ImageBatch(u8, out_dtype=np.uint32)
    w = DeviceBuffer.from_array(weights())
    ws = DeviceBuffer.from_array(ops.spatial_weights(2, 2.0))
    norm = float(np.float32(-50.0))
    wsp_u, wsp_f = bu.iimg_workspace(), bf.iimg_workspace()
    outs = {k: DeviceBuffer(bf.total * 4) for k in ("wf", "bl", "if")}
    st = Stream()
    pick = {
        "wf": batch.autotune(batch.OP_FILTER, bf, lambda c, s: bf.wfilter3x3(w, 0.0, 0.0, 0, c, s, out=outs["wf"]))[0],
        "bl": batch.autotune(batch.OP_BILATERAL, bf,
                             lambda c, s: bf.bilateral(2, ws, norm, 0.0, 0, c, s, out=outs["bl"]))[0],
        "if": batch.autotune(batch.OP_IIMG, bf, lambda c, s: bf.iimg(wsp_f, s, out=outs["if"], config=c))[0],
        "iu": batch.autotune(batch.OP_IIMG, bu, lambda c, s: bu.iimg(wsp_u, s, config=c))[0],
    }
    print(f"[batch] tuned configurations {pick}", flush=True)

    def run(s):
        bf.wfilter3x3(w, 0.0, 0.0, 0, pick["wf"], s, out=outs["wf"])
        bf.bilateral(2, ws, norm, 0.0, 0, pick["bl"], s, out=outs["bl"])
        bf.iimg(wsp_f, s, out=outs["if"], config=pick["if"])
        bu.iimg(wsp_u, s, config=pick["iu"])

    run(st)
    st.synchronize()
    eager = {k: v.download((bf.total,), np.float32) for k, v in outs.items()}
    eager["iu"] = bu.out.download((bu.total,), np.uint32)
    g = Graph(st)
    g.begin()
    run(st)
    g.end()
    for rep in range(3):
        for v in list(outs.values()) + [bu.out]:
            v.fill(0xAB)  # eager work between replays
        g.launch()
        st.synchronize()
        for k, v in outs.items():
            np.testing.assert_array_equal(v.download((bf.total,), np.float32), eager[k], err_msg=f"{k} replay {rep}")
        np.testing.assert_array_equal(bu.out.download((bu.total,), np.uint32), eager["iu"])
    # every image against the oracle: wfilter bit-exact (same fmaf order), bilateral within 2e-6
    # of the fp32 oracle, u8 integral image bit-exact, f32 integral image within 4e-6 of FP64
    o = bf.offsets
    wt = weights()
    ws_h = ops.spatial_weights(2, 2.0)
    worst = {"bl": 0.0, "if": 0.0}
    for i in range(256):
        h, wd = bf.shapes[i]
        sl = slice(o[i], o[i] + h * wd)
        np.testing.assert_array_equal(eager["wf"][sl].reshape(h, wd), oracle.wfilter3x3(imgs[i], wt, 0.0, 0.0),
                                      err_msg=f"wfilter image {i}")
        np.testing.assert_array_equal(eager["iu"][sl].reshape(h, wd), oracle.iimg(u8[i]), err_msg=f"iimg u8 {i}")
        ref = oracle.bilateral(imgs[i], 2, ws_h, norm, 0.0, 0)
        worst["bl"] = max(worst["bl"], float(np.max(np.abs(eager["bl"][sl].reshape(h, wd) - ref) /
                                                    np.maximum(np.abs(ref), 1e-30))))
        ref = oracle.iimg(imgs[i])
        worst["if"] = max(worst["if"], float(np.max(np.abs(eager["if"][sl].reshape(h, wd) - ref) / ref)))
    print(f"[batch] all 256 images: bilateral max rel {worst['bl']:.3e}, iimg f32 max rel {worst['if']:.3e}",
          flush=True)
    assert worst["bl"] <= 2e-6
    assert worst["if"] <= 4e-6


def test_autotune_picks_a_config(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("UMI_TUNE_CACHE", str(tmp_path / "tune.json"))
    imgs, _ = small_batch(seed=6, n=16)
    b = batch.ImageBatch(imgs)
    w = DeviceBuffer.from_array(weights())
    best, times = batch.autotune(batch.OP_FILTER, b, lambda c, s: b.wfilter3x3(w, 0.0, 0.0, 0, c, s), replays=3)
    assert best in times and len(times) == batch.n_configs(batch.OP_FILTER)
    again, _ = batch.autotune(batch.OP_FILTER, b, lambda c, s: b.wfilter3x3(w, 0.0, 0.0, 0, c, s))
    assert again == best  # served from the cache
