"""Host-side operator calls over the C ABI (device buffers in, device buffers out).

Each function validates shapes on the host, then calls the matching
``umi_*`` entry point of libumi_hip.  The array-level helpers (``*_np``)
upload numpy inputs, run the operator and download the result; they are what
the tests use.
"""
from __future__ import annotations

import numpy as np

from ._lib import call
from .device import DeviceBuffer


def _stream(stream):
    return None if stream is None else stream.handle


# ---------------------------------------------------------------------------
# Full-search motion estimation (main/full_me.cu:34-87, johnjohnlin/UMI)
# ---------------------------------------------------------------------------
def me_full_search(cur: DeviceBuffer, ref: DeviceBuffer, H: int, W: int, pitch: int, B: int, R: int,
                   mv_sad: DeviceBuffer | None = None, volume: DeviceBuffer | None = None, stream=None) -> None:
    nby, nbx, S = H // B, W // B, 2 * R + 1
    if mv_sad is not None and mv_sad.nbytes < nby * nbx * 3 * 4:
        raise ValueError("mv_sad buffer too small")
    if volume is not None and volume.nbytes < nby * nbx * S * S * 4:
        raise ValueError("volume buffer too small")
    if cur.nbytes < (H - 1) * pitch + W or ref.nbytes < (H - 1) * pitch + W:
        raise ValueError("frame buffer too small")
    call("umi_me_full_search_u8", cur.ptr, ref.ptr, H, W, pitch, B, R,
         mv_sad.ptr if mv_sad is not None else None, volume.ptr if volume is not None else None,
         _stream(stream))


def me_full_search_np(cur: np.ndarray, ref: np.ndarray, B: int = 8, R: int = 16, volume: bool = False):
    """Returns (mv_sad int32 [nby][nbx][3], volume int32 [nby][nbx][S][S] or None)."""
    assert cur.dtype == np.uint8 and ref.dtype == np.uint8 and cur.shape == ref.shape
    H, W = cur.shape
    nby, nbx, S = H // B, W // B, 2 * R + 1
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    mv = DeviceBuffer(max(nby * nbx * 3 * 4, 4))
    vol = DeviceBuffer(max(nby * nbx * S * S * 4, 4)) if volume else None
    me_full_search(dc, dr, H, W, W, B, R, mv, vol)
    out_mv = mv.download((nby, nbx, 3), np.int32)
    out_vol = vol.download((nby, nbx, S, S), np.int32) if vol is not None else None
    return out_mv, out_vol


# ---------------------------------------------------------------------------
# SAD cost volume (main/cost_volume.cu:35-80, johnjohnlin/UMI)
# ---------------------------------------------------------------------------
def cost_volume(L: DeviceBuffer, R: DeviceBuffer, H: int, W: int, pitch: int, D: int, s0: int, step: int,
                KY: int, KX: int, out: DeviceBuffer, stream=None) -> None:
    if out.nbytes < H * W * D * 4:
        raise ValueError("cost volume buffer too small")
    if L.nbytes < (H - 1) * pitch + W or R.nbytes < (H - 1) * pitch + W:
        raise ValueError("image buffer too small")
    call("umi_cost_volume_u8", L.ptr, R.ptr, H, W, pitch, D, s0, step, KY, KX, out.ptr, _stream(stream))


def cost_volume_np(L: np.ndarray, R: np.ndarray, D: int, s0: int = 0, step: int = -1, KY: int = 2,
                   KX: int = 2) -> np.ndarray:
    assert L.dtype == np.uint8 and R.dtype == np.uint8 and L.shape == R.shape
    H, W = L.shape
    dl, dr = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
    out = DeviceBuffer(H * W * D * 4)
    cost_volume(dl, dr, H, W, W, D, s0, step, KY, KX, out)
    return out.download((H, W, D), np.int32)


# ---------------------------------------------------------------------------
# 3x3 filters (main/filter3x3.cu, main/wfilter3x3.cu) and bilateral
# (main/bilateral.cu), johnjohnlin/UMI
# ---------------------------------------------------------------------------
PAD, CLAMP = 0, 1


def filter3x3(inp: DeviceBuffer, H: int, W: int, in_pitch: int, init, mul, pad, mode: int, out: DeviceBuffer,
              out_pitch: int, dtype=np.float32, stream=None) -> None:
    name = "umi_filter3x3_f32" if dtype == np.float32 else "umi_filter3x3_i32"
    call(name, inp.ptr, H, W, in_pitch, init, mul, pad, mode, out.ptr, out_pitch, _stream(stream))


def wfilter3x3(inp: DeviceBuffer, H: int, W: int, in_pitch: int, weights: DeviceBuffer, init, pad, mode: int,
               out: DeviceBuffer, out_pitch: int, dtype=np.float32, stream=None) -> None:
    name = "umi_wfilter3x3_f32" if dtype == np.float32 else "umi_wfilter3x3_i32"
    call(name, inp.ptr, H, W, in_pitch, weights.ptr, init, pad, mode, out.ptr, out_pitch, _stream(stream))


def bilateral(inp: DeviceBuffer, H: int, W: int, in_pitch: int, radius: int, spatial: DeviceBuffer, norm: float,
              pad: float, mode: int, out: DeviceBuffer, out_pitch: int, stream=None) -> None:
    call("umi_bilateral_f32", inp.ptr, H, W, in_pitch, radius, spatial.ptr, norm, pad, mode, out.ptr, out_pitch,
         _stream(stream))


def filter3x3_np(img: np.ndarray, init, mul, pad, mode: int = PAD) -> np.ndarray:
    dt = np.int32 if img.dtype == np.int32 else np.float32
    img = np.ascontiguousarray(img, dt)
    H, W = img.shape
    d, o = DeviceBuffer.from_array(img), DeviceBuffer(max(img.nbytes, 4))
    filter3x3(d, H, W, W, init, mul, pad, mode, o, W, dt)
    return o.download((H, W), dt)


def wfilter3x3_np(img: np.ndarray, w: np.ndarray, init, pad, mode: int = PAD) -> np.ndarray:
    dt = np.int32 if img.dtype == np.int32 else np.float32
    img = np.ascontiguousarray(img, dt)
    H, W = img.shape
    d, o = DeviceBuffer.from_array(img), DeviceBuffer(max(img.nbytes, 4))
    dw = DeviceBuffer.from_array(np.ascontiguousarray(w, dt).reshape(9))
    wfilter3x3(d, H, W, W, dw, init, pad, mode, o, W, dt)
    return o.download((H, W), dt)


def spatial_weights(radius: int, sigma_s: float) -> np.ndarray:
    """ws[d] = exp(-d^2 / (2 sigma_s^2)), d = -r..r (main/bilateral.cu:45-48)."""
    d = np.arange(-radius, radius + 1, dtype=np.float64)
    return np.exp(-d * d / (2.0 * sigma_s * sigma_s)).astype(np.float32)


def bilateral_np(img: np.ndarray, radius: int, sigma_s: float, sigma_r: float, pad: float = 0.0,
                 mode: int = PAD) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    ws = spatial_weights(radius, sigma_s)
    norm = np.float32(-1.0 / (2.0 * sigma_r * sigma_r))
    d, o, dw = DeviceBuffer.from_array(img), DeviceBuffer(max(img.nbytes, 4)), DeviceBuffer.from_array(ws)
    bilateral(d, H, W, W, radius, dw, float(norm), pad, mode, o, W)
    return o.download((H, W), np.float32)


# ---------------------------------------------------------------------------
# Integral image (main/iimg.cu), johnjohnlin/UMI
# ---------------------------------------------------------------------------
_IIMG = {np.dtype(np.uint8): (0, "umi_iimg_u8_u32", np.uint32), np.dtype(np.int32): (1, "umi_iimg_i32", np.int32),
         np.dtype(np.float32): (2, "umi_iimg_f32", np.float32)}


def iimg_workspace_size(H: int, W: int, dtype) -> int:
    import ctypes

    kind = _IIMG[np.dtype(dtype)][0]
    n = ctypes.c_size_t(0)
    call("umi_iimg_workspace_size", H, W, kind, ctypes.byref(n))
    return n.value


def iimg(inp: DeviceBuffer, H: int, W: int, in_pitch: int, out: DeviceBuffer, out_pitch: int, dtype,
         workspace: DeviceBuffer | None = None, stream=None) -> None:
    _, name, _ = _IIMG[np.dtype(dtype)]
    call(name, inp.ptr, H, W, in_pitch, out.ptr, out_pitch, workspace.ptr if workspace is not None else None,
         workspace.nbytes if workspace is not None else 0, _stream(stream))


def iimg_np(img: np.ndarray, use_workspace: bool = True) -> np.ndarray:
    img = np.ascontiguousarray(img)
    H, W = img.shape
    _, _, odt = _IIMG[img.dtype]
    d, o = DeviceBuffer.from_array(img), DeviceBuffer(max(H * W * 4, 4))
    ws = DeviceBuffer(max(iimg_workspace_size(H, W, img.dtype), 4)) if use_workspace else None
    iimg(d, H, W, W, o, W, img.dtype, ws)
    return o.download((H, W), odt)
