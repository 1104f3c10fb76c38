"""The learned motion-proposal pipeline on the GPU.

* training kernels: activations and every layer's weight/bias gradients
  match the FP64 oracle (oracle/proposer_ref.py, pinned by finite
  differences) on a small input;
* Adam matches its formula; training lowers the loss;
* the bf16 MFMA inference layers match the FP64 forward pass with the same
  bf16 roundings; the fused first (from the frames) and last (with the
  proposals) layers equal the unfused ones bit for bit;
* refinement (seed, descent, in-tile and global propagation rounds, fallback
  check) matches the oracle exactly, including ragged frames, padded rows,
  the largest radius and edge settings; the
  listed full search equals the full search on the listed blocks and leaves
  the others alone;
* with the trained network (umi_amd/data checkpoint): >= 99 % of blocks reach
  the full search's minimum SAD on held-out pairs, more than the same
  refinement from zero proposals (the method without the network).
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
    """64 Adam steps on 8 full-size pairs (batch 4) lower the training loss."""
    data = proposer.Dataset(8, seed0=500)
    net = proposer.Network(seed=1)
    hist = proposer.train(net, data, epochs=32, batch=4, lr=2e-3)
    print(f"[proposer] loss history {hist}", flush=True)
    assert hist[-1] < 0.85 * hist[0], hist


@pytest.mark.parametrize("N,H4,W4", [(1, 24, 40), (2, 18, 37), (1, 7, 5)])
def test_inference_layers_match_bf16_oracle(gpu, N, H4, W4):
    """The four bf16 MFMA inference layers against the FP64 forward pass with
    the same bf16 roundings (weights, inputs, activations between layers).
    The GPU sums in fp32 in MFMA order; the tolerance, 2e-3 x max |flow|,
    covers that and the rare activations whose bf16 rounding it flips."""
    g = workloads.rng(N * 100 + H4 + W4)
    net = proposer.Network(seed=4)
    for li in range(len(proposer.LAYERS)):   # nonzero biases exercise the bias path
        b = g.standard_normal(proposer.LAYERS[li][1]).astype(np.float32) * 0.1
        net.b[li].upload(b)
    params = net.params()
    x = g.standard_normal((N, 2, H4, W4)).astype(np.float32)
    dx = DeviceBuffer.from_array(x)
    acts = [DeviceBuffer(N * H4 * W4 * 32 * 2) for _ in range(3)]
    flow = DeviceBuffer(N * 2 * H4 * W4 * 4)
    packed = [proposer.pack_layer(net, li) for li in range(len(proposer.LAYERS))]
    src = dx
    for li, (ci, co, d, _) in enumerate(proposer.LAYERS):
        kind = proposer.layer_kind(li)
        dst = flow if kind == 2 else acts[li]
        print(f"[infer] N={N} H4={H4} W4={W4} layer {li} kind {kind}", flush=True)
        call("umi_cnn_infer_layer", kind, src.ptr, N, H4, W4, d, packed[li].ptr, net.b[li].ptr, dst.ptr, None)
        src = dst
    got = flow.download((N, 2, H4, W4), np.float32)
    want = PR.forward_bf16(x, params, proposer.LAYERS)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[infer] max |err| / max |flow| = {err:.2e}", flush=True)
    assert err <= 2e-3
    # the first layer alone is exact up to fp32 summation order
    call("umi_cnn_infer_layer", 0, dx.ptr, N, H4, W4, 1, packed[0].ptr, net.b[0].ptr, acts[0].ptr, None)
    a0 = acts[0].download((N, H4, W4, 32), np.uint16).astype(np.uint32) << 16
    a0 = a0.view(np.float32).transpose(0, 3, 1, 2)
    w0 = PR.bf16(PR.conv(PR.bf16(x), PR.bf16(params[0][0]), params[0][1], 1).clip(0))
    assert np.abs(a0 - w0).max() <= 2 ** -7 * np.abs(w0).max()


def test_downsample_matches_numpy(gpu):
    """Quarter-resolution input: (mean of each 4x4 block - 128) / 8 of cur and
    ref; every step is exact in float32, so the comparison is bit-exact."""
    H, W = 92, 138
    g = workloads.rng(12)
    cur = g.integers(0, 256, size=(H, W)).astype(np.uint8)
    ref = g.integers(0, 256, size=(H, W)).astype(np.uint8)
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    H4, W4 = H // 4, W // 4
    x = DeviceBuffer(2 * 2 * H4 * W4 * 4)
    call("umi_cnn_downsample_pair", dc.ptr, dr.ptr, H, W, W, x.ptr, 1, None)  # pair slot 1 of 2
    got = x.download((2, 2, H4, W4), np.float32)[1]
    for ch, f in enumerate((cur, ref)):
        m = f[:H4 * 4, :W4 * 4].reshape(H4, 4, W4, 4).astype(np.float64).sum((1, 3))
        np.testing.assert_array_equal(got[ch], ((m / 16 - 128) / 8).astype(np.float32))


@pytest.mark.parametrize("H,W", [(72, 136), (1080, 1920), (44, 70)])
def test_fused_first_layer_equals_downsample_path(gpu, H, W):
    """umi_cnn_first_layer_frames = umi_cnn_downsample_pair + layer kind 0, bit for bit."""
    g = workloads.rng(H + W)
    cur = g.integers(0, 256, size=(H, W)).astype(np.uint8)
    ref = g.integers(0, 256, size=(H, W)).astype(np.uint8)
    net = proposer.Network(seed=5)
    packed = proposer.pack_layer(net, 0)
    H4, W4 = H // 4, W // 4
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    x = DeviceBuffer(2 * H4 * W4 * 4)
    a, b = DeviceBuffer(H4 * W4 * 64), DeviceBuffer(H4 * W4 * 64)
    print(f"[fused] first layer H={H} W={W}", flush=True)
    call("umi_cnn_downsample_pair", dc.ptr, dr.ptr, H, W, W, x.ptr, 0, None)
    call("umi_cnn_infer_layer", 0, x.ptr, 1, H4, W4, 1, packed.ptr, net.b[0].ptr, a.ptr, None)
    call("umi_cnn_first_layer_frames", dc.ptr, dr.ptr, H, W, W, packed.ptr, net.b[0].ptr, b.ptr, None)
    np.testing.assert_array_equal(a.download((H4 * W4 * 32,), np.uint16), b.download((H4 * W4 * 32,), np.uint16))


@pytest.mark.parametrize("H,W,d,pitch", [(1080, 1920, 2, 1920), (72, 136, 2, 136), (44, 70, 2, 70),
                                         (60, 100, 3, 101), (36, 200, 1, 200)])
def test_fused_first_two_layers_equal_two_launches(gpu, H, W, d, pitch):
    """umi_cnn_first_two_layers_frames = umi_cnn_first_layer_frames + layer
    kind 1 of dilation d, bit for bit (pitch 101: the byte-load path)."""
    g = workloads.rng(H * W + d)
    cur = g.integers(0, 256, size=(H, pitch)).astype(np.uint8)
    ref = g.integers(0, 256, size=(H, pitch)).astype(np.uint8)
    net = proposer.Network(seed=7)
    p0, p1 = proposer.pack_layer(net, 0), proposer.pack_layer(net, 1)
    H4, W4 = H // 4, W // 4
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    a, b, c = (DeviceBuffer(H4 * W4 * 64) for _ in range(3))
    print(f"[fused] first two layers H={H} W={W} d={d} pitch={pitch}", flush=True)
    call("umi_cnn_first_layer_frames", dc.ptr, dr.ptr, H, W, pitch, p0.ptr, net.b[0].ptr, a.ptr, None)
    call("umi_cnn_infer_layer", 1, a.ptr, 1, H4, W4, d, p1.ptr, net.b[1].ptr, b.ptr, None)
    call("umi_cnn_first_two_layers_frames", dc.ptr, dr.ptr, H, W, pitch, p0.ptr, net.b[0].ptr, d, p1.ptr,
         net.b[1].ptr, c.ptr, None)
    want = b.download((H4 * W4 * 32,), np.uint16)
    assert np.count_nonzero(want) > want.size // 4  # ReLU leaves a good share of the outputs
    np.testing.assert_array_equal(c.download((H4 * W4 * 32,), np.uint16), want)


@pytest.mark.parametrize("H4,W4,R", [(270, 480, 32), (19, 37, 5), (8, 16, 32)])
def test_fused_last_layer_equals_flow_and_proposals(gpu, H4, W4, R):
    """umi_cnn_last_layer_proposals writes layer kind 2's flow and
    umi_me_proposals' block proposals, bit for bit."""
    g = workloads.rng(H4 * W4)
    act = (g.standard_normal((H4, W4, 32)).astype(np.float32) * 3.0)
    act16 = (act.view(np.uint32) >> 16).astype(np.uint16)
    net = proposer.Network(seed=6)
    net.b[3].upload(np.array([0.3, -0.7], np.float32))
    packed = proposer.pack_layer(net, 3)
    di = DeviceBuffer.from_array(act16)
    f1, f2 = DeviceBuffer(2 * H4 * W4 * 4), DeviceBuffer(2 * H4 * W4 * 4)
    nb = (H4 // 2) * (W4 // 2)
    p1, p2 = DeviceBuffer(max(nb, 1) * 8), DeviceBuffer(max(nb, 1) * 8)
    print(f"[fused] last layer H4={H4} W4={W4} R={R}", flush=True)
    call("umi_cnn_infer_layer", 2, di.ptr, 1, H4, W4, 8, packed.ptr, net.b[3].ptr, f1.ptr, None)
    call("umi_me_proposals", f1.ptr, H4, W4, R, p1.ptr, None)
    call("umi_cnn_last_layer_proposals", di.ptr, H4, W4, 8, packed.ptr, net.b[3].ptr, R, f2.ptr, p2.ptr, None)
    np.testing.assert_array_equal(f1.download((2, H4, W4), np.float32), f2.download((2, H4, W4), np.float32))
    np.testing.assert_array_equal(p1.download((nb, 2), np.int32), p2.download((nb, 2), np.int32))
    flow = f1.download((2, H4, W4), np.float32).astype(np.float64)
    want = np.rint(flow[:, :H4 // 2 * 2:2, :W4 // 2 * 2:2] + flow[:, :H4 // 2 * 2:2, 1:W4 // 2 * 2:2]
                   + flow[:, 1:H4 // 2 * 2:2, :W4 // 2 * 2:2] + flow[:, 1:H4 // 2 * 2:2, 1:W4 // 2 * 2:2])
    want = np.clip(want, -R, R).transpose(1, 2, 0).reshape(nb, 2)
    # float32 rounding of the quad sum can move a value sitting at .5 by one
    assert np.abs(p2.download((nb, 2), np.int32) - want).max() <= 1


def _refine_gpu(cur, ref, prop, R, iters, local_rounds, rounds, slack, pitch=None):
    H, W = cur.shape
    pitch = pitch or W
    n = (H // 8) * (W // 8)
    pad = lambda f: np.pad(f, ((0, 0), (0, pitch - W)))  # noqa: E731
    dc, dr = DeviceBuffer.from_array(pad(cur)), DeviceBuffer.from_array(pad(ref))
    dp = DeviceBuffer.from_array(np.ascontiguousarray(prop, np.int32))
    mv, fl, fc = DeviceBuffer(n * 12), DeviceBuffer(n * 4), DeviceBuffer(4)
    nbytes = C.c_size_t()
    call("umi_me_refine_workspace", H, W, rounds, C.byref(nbytes))
    ws = DeviceBuffer(nbytes.value)
    print(f"[refine] H={H} W={W} pitch={pitch} R={R} iters={iters} local={local_rounds} rounds={rounds} "
          f"slack={slack}", flush=True)
    call("umi_me_refine", dc.ptr, dr.ptr, H, W, pitch, R, dp.ptr, iters, local_rounds, rounds, C.c_float(slack), mv.ptr,
         fl.ptr, fc.ptr, ws.ptr, ws.nbytes, None)
    got = mv.download((n, 3), np.int32)
    k = int(fc.download((1,), np.int32)[0])
    fb = fl.download((max(k, 1),), np.int32)[:k].tolist()
    raw = ws.download((ws.nbytes,), np.uint8)
    off = 2 * ((n * 16 + 255) // 256 * 256)
    changed = np.frombuffer(raw[off:off + 4 * rounds].tobytes(), np.int32) != 0
    return got, fb, changed.tolist(), (dc, dr, fl, fc, mv)


@pytest.mark.parametrize("case", [
    dict(H=96, W=136, R=32, iters=8, local=8, rounds=8, slack=1.1, kind="tiles"),
    dict(H=100, W=138, R=16, iters=8, local=3, rounds=8, slack=1.1, kind="smooth"),    # ragged, unaligned rows
    dict(H=96, W=136, R=16, iters=8, local=0, rounds=8, slack=1.1, kind="smooth", pitch=140),
    dict(H=256, W=384, R=24, iters=8, local=8, rounds=3, slack=1.1, kind="smooth"),
    dict(H=256, W=384, R=24, iters=8, local=8, rounds=12, slack=1.1, kind="tiles"),
    dict(H=64, W=64, R=8, iters=1, local=0, rounds=0, slack=1.1, kind="smooth"),       # no rounds, one step
    dict(H=64, W=72, R=8, iters=4, local=2, rounds=3, slack=0.0, kind="smooth"),       # slack 0 forces fallbacks
    dict(H=200, W=264, R=64, iters=8, local=8, rounds=4, slack=1.1, kind="smooth"),    # largest radius
    dict(H=200, W=264, R=32, iters=4, local=4, rounds=4, slack=1.1, kind="tiles"),     # the Proposer defaults
])
def test_refine_matches_oracle(gpu, case):
    H, W, R = case["H"], case["W"], case["R"]
    if case["kind"] == "tiles":
        cur, ref, _ = workloads.me_pair(H, W, min(R, 16), seed=9)
    else:
        cur, ref = workloads.smooth_motion_pair(H, W, R, seed=H + W)
    n = (H // 8) * (W // 8)
    g = workloads.rng(10)
    prop = g.integers(-R - 3, R + 4, size=(n, 2)).astype(np.int32)   # out-of-range proposals are clamped
    prop[::3] = 0
    got, fb, changed, _ = _refine_gpu(cur, ref, prop, R, case["iters"], case["local"], case["rounds"], case["slack"],
                                      case.get("pitch"))
    want, want_fb, want_changed = PR.refine(cur, ref, prop, R, case["iters"], case["local"], case["rounds"],
                                            case["slack"])
    np.testing.assert_array_equal(got, want)
    assert len(fb) == len(set(fb)) and set(fb) == want_fb
    assert changed == want_changed


def _refine_fused_gpu(cur, ref, props, R, iters, local_rounds, rounds, slack, pitch=None):
    """umi_me_refine_fused on one zeroed workspace for each proposal set in
    `props` in turn (consecutive calls on one workspace); per call (mv_sad,
    fallback list, round change flags)."""
    H, W = cur.shape
    pitch = pitch or W
    n = (H // 8) * (W // 8)
    pad = lambda f: np.pad(f, ((0, 0), (0, pitch - W)))  # noqa: E731
    dc, dr = DeviceBuffer.from_array(pad(cur)), DeviceBuffer.from_array(pad(ref))
    mv, fl, fc = DeviceBuffer(n * 12), DeviceBuffer(n * 4), DeviceBuffer(4)
    nbytes = C.c_size_t()
    call("umi_me_refine_fused_workspace", H, W, rounds, C.byref(nbytes))
    ws = DeviceBuffer.zeros(nbytes.value)
    out = []
    for i, prop in enumerate(props):
        dp = DeviceBuffer.from_array(np.ascontiguousarray(prop, np.int32))
        print(f"[refine fused] call {i}: H={H} W={W} pitch={pitch} R={R} iters={iters} local={local_rounds} "
              f"rounds={rounds} slack={slack}", flush=True)
        mv.fill(0)
        call("umi_me_refine_fused", dc.ptr, dr.ptr, H, W, pitch, R, dp.ptr, iters, local_rounds, rounds,
             C.c_float(slack), mv.ptr, fl.ptr, fc.ptr, ws.ptr, ws.nbytes, None)
        got = mv.download((n, 3), np.int32)
        k = int(fc.download((1,), np.int32)[0])
        fb = fl.download((max(k, 1),), np.int32)[:k].tolist()
        changed = (C.c_int32 * max(rounds, 1))()
        timed_out = C.c_int()
        call("umi_me_refine_fused_status", ws.ptr, rounds, C.cast(changed, C.c_void_p), C.byref(timed_out))
        assert timed_out.value == 0, "a wait of the fused refinement ran out"
        out.append((got, fb, [bool(changed[p]) for p in range(rounds)]))
    return out


@pytest.mark.parametrize("case", [
    dict(H=96, W=136, R=32, iters=8, local=8, rounds=8, slack=1.1, kind="tiles"),
    dict(H=100, W=138, R=16, iters=8, local=3, rounds=8, slack=1.1, kind="smooth"),    # ragged, unaligned rows
    dict(H=96, W=136, R=16, iters=8, local=0, rounds=8, slack=1.1, kind="smooth", pitch=140),
    dict(H=256, W=384, R=24, iters=8, local=8, rounds=3, slack=1.1, kind="smooth"),
    dict(H=256, W=384, R=24, iters=8, local=8, rounds=12, slack=1.1, kind="tiles"),
    dict(H=64, W=64, R=8, iters=1, local=0, rounds=0, slack=1.1, kind="smooth"),       # no rounds, one step
    dict(H=64, W=72, R=8, iters=4, local=2, rounds=3, slack=0.0, kind="smooth"),       # slack 0 forces fallbacks
    dict(H=200, W=264, R=64, iters=8, local=8, rounds=4, slack=1.1, kind="smooth"),    # largest radius
    dict(H=200, W=264, R=32, iters=4, local=4, rounds=4, slack=1.1, kind="tiles"),     # the Proposer defaults
])
def test_refine_fused_matches_oracle(gpu, case):
    """The one-launch refinement against the oracle's every-round schedule, on
    two consecutive calls on one workspace (the second with other proposals,
    then the first again: the workspace's epochs carry over)."""
    H, W, R = case["H"], case["W"], case["R"]
    if case["kind"] == "tiles":
        cur, ref, _ = workloads.me_pair(H, W, min(R, 16), seed=9)
    else:
        cur, ref = workloads.smooth_motion_pair(H, W, R, seed=H + W)
    n = (H // 8) * (W // 8)
    g = workloads.rng(10)
    props = []
    for _ in range(2):
        prop = g.integers(-R - 3, R + 4, size=(n, 2)).astype(np.int32)   # out-of-range proposals are clamped
        prop[::3] = 0
        props.append(prop)
    props.append(props[0])
    res = _refine_fused_gpu(cur, ref, props, R, case["iters"], case["local"], case["rounds"], case["slack"],
                            case.get("pitch"))
    for i, (got, fb, changed) in enumerate(res[:2]):
        want, want_fb, want_changed = PR.refine(cur, ref, props[i], R, case["iters"], case["local"], case["rounds"],
                                                case["slack"], all_rounds=True)
        np.testing.assert_array_equal(got, want, err_msg=f"call {i}")
        assert len(fb) == len(set(fb)) and set(fb) == want_fb
        assert changed == want_changed
    np.testing.assert_array_equal(res[2][0], res[0][0])
    assert res[2][1] == res[0][1]


def test_refine_fused_full_frame_equals_multi_launch(gpu):
    """1920x1080 (1020 tiles, the whole grid resident), +-32, the Proposer's
    settings: the one-launch refinement equals the multi-launch one whenever
    the latter ran every round (no idle round before the last), which the
    test checks; the oracle of this frame takes ~20 s, so the small cases
    above carry the oracle comparison."""
    H, W, R = 1080, 1920, 32
    cur, ref = workloads.smooth_motion_pair(H, W, R, seed=H + W)
    n = (H // 8) * (W // 8)
    prop = workloads.rng(10).integers(-R - 3, R + 4, size=(n, 2)).astype(np.int32)
    prop[::3] = 0
    multi, fb_m, ch_m, _ = _refine_gpu(cur, ref, prop, R, 4, 4, 4, 1.1)
    assert False not in ch_m[:-1], f"an idle round before the last: {ch_m}"
    res = _refine_fused_gpu(cur, ref, [prop, prop], R, 4, 4, 4, 1.1)
    for got, fb, changed in res:
        np.testing.assert_array_equal(got, multi)
        assert sorted(fb) == sorted(fb_m) and changed == ch_m


def test_fused_refinement_refusal_falls_back(gpu):
    """A frame of more than 1024 tiles (3840x2160: 4080) is refused by the
    one-launch refinement with UMI_ERR_UNSUPPORTED before any launch; a
    Proposer that chose the one-launch form by itself then refines with one
    launch per stage, with the same results as one built for it."""
    from umi_amd._lib import UMI_ERR_UNSUPPORTED, UmiError
    if not proposer.CHECKPOINT.exists():
        pytest.skip("no trained checkpoint committed yet")
    net = proposer.Network.load(proposer.CHECKPOINT)
    H, W = 2160, 3840
    auto = proposer.Proposer(net, H=H, W=W, fused=True)
    auto._auto = True  # as if chosen automatically
    multi = proposer.Proposer(net, H=H, W=W, fused=False)
    cur, ref = workloads.smooth_motion_pair(H, W, 32, seed=5)
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    ma, mb = DeviceBuffer(auto.nblk * 12), DeviceBuffer(auto.nblk * 12)
    with pytest.raises(UmiError) as e:
        call("umi_me_refine_fused", dc.ptr, dr.ptr, H, W, W, 32, auto.prop.ptr, 4, 4, 4, C.c_float(1.1), ma.ptr,
             auto.fb_list.ptr, auto.fb_count.ptr, auto.ws.ptr, auto.ws.nbytes, None)
    assert e.value.status == UMI_ERR_UNSUPPORTED
    print("[refine fused] 4080 tiles: refused, falling back", flush=True)
    auto.run(dc, dr, ma, proposals=False)
    multi.run(dc, dr, mb, proposals=False)
    assert not auto.fused
    np.testing.assert_array_equal(ma.download((auto.nblk, 3), np.int32), mb.download((multi.nblk, 3), np.int32))


def test_listed_search_completes_fallbacks(gpu):
    H, W, R = 96, 136, 32
    cur, ref, _ = workloads.me_pair(H, W, 16, seed=9)
    n = (H // 8) * (W // 8)
    prop = workloads.rng(11).integers(-R, R + 1, size=(n, 2)).astype(np.int32)
    before, fb, _, (dc, dr, fl, fc, mv) = _refine_gpu(cur, ref, prop, R, 2, 0, 1, 1.1)
    assert 0 < len(fb) < n
    full, _ = ops.me_full_search_np(cur, ref, 8, R)
    full = full.reshape(n, 3)
    listed = np.zeros(n, bool)
    listed[fb] = True
    np.testing.assert_array_equal(before[listed], full[listed])  # the refinement's own full search
    call("umi_me_full_search_listed_u8", dc.ptr, dr.ptr, H, W, W, 8, R, fl.ptr, fc.ptr, n, mv.ptr, None)
    after = mv.download((n, 3), np.int32)
    np.testing.assert_array_equal(after[listed], full[listed])
    np.testing.assert_array_equal(after[~listed], before[~listed])


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
    """Held-out 1920x1080 pairs, +-32 search: with the committed network >= 99 %
    of blocks reach the full search's minimum SAD; the same refinement from
    zero proposals (the method without the network), at the same round
    budget, reaches fewer."""
    if not proposer.CHECKPOINT.exists():
        pytest.skip("no trained checkpoint committed yet")
    net = proposer.Network.load(proposer.CHECKPOINT)
    prop = proposer.Proposer(net)
    with_net = proposer.evaluate(prop, 3, 2_000_000, proposals=True, reps=3)
    without = proposer.evaluate(prop, 3, 2_000_000, proposals=False, reps=3)
    print(f"[proposer] with network {with_net}\n[proposer] without {without}", flush=True)
    assert with_net["at_min_sad"] >= 0.99
    assert without["at_min_sad"] < with_net["at_min_sad"]


def test_pipeline_graph_replays_equal_eager(gpu):
    """The one-pair pipeline (network + one-launch refinement) captured as one
    graph: three replays, with eager work between them, reproduce the eager
    vectors bit for bit (evaluate raises otherwise) and the graph timings are
    reported."""
    if not proposer.CHECKPOINT.exists():
        pytest.skip("no trained checkpoint committed yet")
    net = proposer.Network.load(proposer.CHECKPOINT)
    prop = proposer.Proposer(net)
    print("[proposer] graph replays of the one-pair pipeline", flush=True)
    r = proposer.evaluate(prop, 1, 2_000_000, proposals=True, reps=2, graphs=True)
    assert r["ms_pipeline_graph"] > 0 and r["ms_full_search_graph"] > 0 and r["speedup_graph"] > 0
    assert r["at_min_sad"] >= 0.99


def test_concurrent_graph_pipelines(gpu):
    """Two held-out pairs in flight on two streams, each pipeline captured in a
    graph: the first replay reproduces the eager pipeline bit for bit (checked
    inside `throughput`) and both pairs stay >= 99 % at the minimum SAD."""
    if not proposer.CHECKPOINT.exists():
        pytest.skip("no trained checkpoint committed yet")
    net = proposer.Network.load(proposer.CHECKPOINT)
    r = proposer.throughput(net, 2, 2_000_000, reps=2)
    print(f"[proposer] two streams {r}", flush=True)
    assert r["min_at_min_sad"] >= 0.99
    assert r["us_per_pair_pipeline"] > 0 and r["us_per_pair_full_search"] > 0
