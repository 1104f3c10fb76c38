"""The learned motion-proposal pipeline on the GPU.

* training kernels: activations and every layer's weight/bias gradients
  match the FP64 oracle (oracle/proposer_ref.py, pinned by finite
  differences) on a small input;
* Adam matches its formula; training lowers the loss;
* refinement and fallback match the oracle rule; the listed full search
  equals the full search on the listed blocks;
* with the trained network (umi_amd/data checkpoint): >= 99 % of blocks reach
  the full search's minimum SAD on held-out pairs, with fewer fallbacks than
  the same refinement around zero motion (the method without the network).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import proposer_ref as PR
from umi_amd import DeviceBuffer, ops, proposer, workloads
from umi_amd._lib import call

pytestmark = pytest.mark.gpu


def test_training_kernels_match_fp64_oracle(gpu):
    N, H4, W4 = 2, 18, 22
    g = workloads.rng(5)
    net = proposer.Network(seed=3)
    params = net.params()
    x = g.standard_normal((N, 2, H4, W4)).astype(np.float32)
    labels = g.integers(-32, 33, size=(N, H4 // 2, W4 // 2, 3)).astype(np.int32)
    tr = proposer.Trainer(net, N, H4, W4)
    dx, dl = DeviceBuffer.from_array(x), DeviceBuffer.from_array(labels)
    tr.forward(dx)
    tr.backward(dx, dl)
    acts = PR.forward(x, params, proposer.LAYERS)
    for li, (ci, co, _, _) in enumerate(proposer.LAYERS):
        got = tr.act[li].download((N, co, H4, W4), np.float32)
        np.testing.assert_allclose(got, acts[li], rtol=2e-4, atol=2e-4 * np.abs(acts[li]).max())
    loss, gout = PR.huber(acts[-1], labels)
    assert abs(tr.loss() - loss) <= 1e-4 * max(1.0, loss)
    grads = PR.backward(x, params, proposer.LAYERS, acts, gout)
    for li, (ci, co, _, _) in enumerate(proposer.LAYERS):
        got = tr.net.gw[li].download((co * ci * 9 + co,), np.float32)
        dw, db = grads[li]
        scale = max(np.abs(dw).max(), np.abs(db).max())
        print(f"[proposer] layer {li} max |grad err| / max|grad| = "
              f"{max(np.abs(got[:co*ci*9] - dw.ravel()).max(), np.abs(got[co*ci*9:] - db).max()) / scale:.2e}",
              flush=True)
        np.testing.assert_allclose(got[:co * ci * 9], dw.ravel(), atol=1e-4 * scale)
        np.testing.assert_allclose(got[co * ci * 9:], db, atol=1e-4 * scale)


def test_adam_formula(gpu):
    g = workloads.rng(6)
    w, gr = g.standard_normal(100).astype(np.float32), g.standard_normal(100).astype(np.float32)
    dw, dg = DeviceBuffer.from_array(w), DeviceBuffer.from_array(gr)
    m, v = DeviceBuffer.zeros(400), DeviceBuffer.zeros(400)
    for step in (1, 2):
        call("umi_adam", dw.ptr, dg.ptr, m.ptr, v.ptr, 100, C.c_float(0.01), C.c_float(0.9), C.c_float(0.999),
             C.c_float(1e-8), step, None)
    mm = 0.1 * gr
    vv = 0.001 * gr * gr
    w1 = w - 0.01 * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
    mm = 0.9 * mm + 0.1 * gr
    vv = 0.999 * vv + 0.001 * gr * gr
    w2 = w1 - 0.01 * (mm / (1 - 0.9 ** 2)) / (np.sqrt(vv / (1 - 0.999 ** 2)) + 1e-8)
    np.testing.assert_allclose(dw.download((100,), np.float32), w2, rtol=1e-5, atol=1e-6)


def test_training_lowers_loss(gpu):
    data = proposer.Dataset(16, seed0=500)
    net = proposer.Network(seed=1)
    hist = proposer.train(net, data, epochs=4, batch=4, lr=2e-3)
    assert hist[-1] < 0.8 * hist[0], hist


def test_refine_rule_and_listed_search_match_oracle(gpu):
    H, W, R = 96, 136, 32
    cur, ref, _ = workloads.me_pair(H, W, 16, seed=9)
    nby, nbx = H // 8, W // 8
    g = workloads.rng(10)
    prop = g.integers(-R, R + 1, size=(nby * nbx, 2)).astype(np.int32)
    prop[::3] = 0
    dc, dr, dp = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref), DeviceBuffer.from_array(prop)
    mv = DeviceBuffer(nby * nbx * 12)
    fl, fc = DeviceBuffer(nby * nbx * 4), DeviceBuffer(4)
    call("umi_me_refine", dc.ptr, dr.ptr, H, W, W, R, dp.ptr, C.c_float(1.1), mv.ptr, fl.ptr, fc.ptr, None)
    got = mv.download((nby * nbx, 3), np.int32)
    n = int(fc.download((1,), np.int32)[0])
    ref_mv, ref_fb = PR.refine(cur, ref, prop, R, 1.1)
    np.testing.assert_array_equal(got, ref_mv)
    assert set(fl.download((n,), np.int32).tolist()) == ref_fb
    call("umi_me_full_search_listed_u8", dc.ptr, dr.ptr, H, W, W, 8, R, fl.ptr, fc.ptr, nby * nbx, mv.ptr, None)
    full, _ = ops.me_full_search_np(cur, ref, 8, R)
    after = mv.download((nby, nbx, 3), np.int32)
    for b in ref_fb:
        np.testing.assert_array_equal(after[b // nbx, b % nbx], full[b // nbx, b % nbx])


def test_synthetic_pairs_deterministic(gpu):
    pm = proposer.PairMaker(128, 192)
    a = [DeviceBuffer(128 * 192) for _ in range(4)]
    pm.make(7, a[0], a[1])
    pm.make(7, a[2], a[3])
    f = [b.download((128, 192), np.uint8) for b in a]
    np.testing.assert_array_equal(f[0], f[2])
    np.testing.assert_array_equal(f[1], f[3])
    assert not np.array_equal(f[0], f[1])


def test_trained_proposer_quality_and_gain(gpu):
    if not proposer.CHECKPOINT.exists():
        pytest.skip("no trained checkpoint committed yet")
    net = proposer.Network.load(proposer.CHECKPOINT)
    prop = proposer.Proposer(net)
    with_net = proposer.evaluate(prop, 3, 2_000_000, proposals=True, reps=3)
    without = proposer.evaluate(prop, 3, 2_000_000, proposals=False, reps=3)
    print(f"[proposer] with network {with_net}\n[proposer] without {without}", flush=True)
    assert with_net["at_min_sad"] >= 0.99
    assert with_net["fallback_frac"] < without["fallback_frac"]
