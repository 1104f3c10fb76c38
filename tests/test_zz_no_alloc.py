"""The allocation-free integral image (umi_iimg_u8_u32_noalloc) through the C
ABI: bit-identical to umi_iimg_u8_u32 with a caller workspace, on two calls in
a row (the second smaller, on the same library scratch), in a graph replayed
with eager work between replays, and refusing an image larger than its
scratch before any launch."""
import ctypes

import numpy as np
import pytest

import oracle
from umi_amd import DeviceBuffer, Graph, Stream, ops, workloads
from umi_amd._lib import lib

pytestmark = pytest.mark.gpu


def _both(img, st):
    H, W = img.shape
    L = lib()
    d = DeviceBuffer.from_array(img)
    o_ref, o_na = DeviceBuffer(H * W * 4), DeviceBuffer(H * W * 4)
    o_na.fill(0xAB)
    ws = ops.iimg_workspace(H, W, np.uint8)
    assert L.umi_iimg_u8_u32(d.ptr, H, W, W, o_ref.ptr, W, ws.ptr, ws.nbytes, st.handle) == 0
    assert L.umi_iimg_u8_u32_noalloc(d.ptr, H, W, W, o_na.ptr, W, st.handle) == 0
    st.synchronize()
    return o_ref.download((H, W), np.uint32), o_na.download((H, W), np.uint32)


def test_noalloc_equals_workspace_form_twice(gpu):
    st = Stream()
    g = workloads.rng(2024)
    for H, W in ((1080, 1920), (37, 300)):  # the second call smaller, same scratch
        img = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
        print(f"[noalloc] {H}x{W}", flush=True)
        ref, got = _both(img, st)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, oracle.iimg(img))


def test_noalloc_graph_replays(gpu):
    st = Stream()
    g = workloads.rng(7)
    img = g.integers(0, 256, size=(300, 700), dtype=np.int64).astype(np.uint8)
    d, o = DeviceBuffer.from_array(img), DeviceBuffer(300 * 700 * 4)
    L = lib()
    assert L.umi_iimg_u8_u32_noalloc(d.ptr, 300, 700, 700, o.ptr, 700, st.handle) == 0  # eager first call
    gr = Graph(st)
    gr.begin()
    assert L.umi_iimg_u8_u32_noalloc(d.ptr, 300, 700, 700, o.ptr, 700, st.handle) == 0
    gr.end()
    ref = oracle.iimg(img)
    for rep in range(3):
        o.fill(0xAB)
        _both(img[:50, :90].copy(), st)  # other calls on the same scratch between replays
        gr.launch()
        st.synchronize()
        np.testing.assert_array_equal(o.download((300, 700), np.uint32), ref, err_msg=f"replay {rep}")


def test_noalloc_refuses_oversize(gpu):
    L = lib()
    d, o = DeviceBuffer(64), DeviceBuffer(64)
    # 8192 x 8192 needs about 36 MB of scratch: refused on the host, nothing launched
    assert L.umi_iimg_u8_u32_noalloc(d.ptr, 8192, 8192, 8192, o.ptr, 8192, None) == 1
    assert b"scratch" in L.umi_last_error()
