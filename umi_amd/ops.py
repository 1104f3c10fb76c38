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
