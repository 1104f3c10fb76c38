"""Host-side operator calls over the C ABI (device buffers in, device buffers out).

Each function validates shapes on the host, then calls the matching
``umi_*`` entry point of libumi_hip.  The array-level helpers (``*_np``)
upload numpy inputs, run the operator and download the result; they are what
the tests use.
"""
from __future__ import annotations

import ctypes

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


def me_subpel_refine(cur: DeviceBuffer, ref: DeviceBuffer, H: int, W: int, pitch: int, mv_in: DeviceBuffer,
                     mv_out: DeviceBuffer, stream=None) -> None:
    """Quarter-pel refinement of integer 8x8 block vectors (dy, dx, SAD) ->
    (dy4, dx4, SAD) in quarter pels."""
    nb = (H // 8) * (W // 8)
    if mv_in.nbytes < nb * 12 or mv_out.nbytes < nb * 12:
        raise ValueError("motion vector buffers too small")
    call("umi_me_subpel_refine_u8", cur.ptr, ref.ptr, H, W, pitch, mv_in.ptr, mv_out.ptr, _stream(stream))


def me_subpel_refine_np(cur: np.ndarray, ref: np.ndarray, mv: np.ndarray) -> np.ndarray:
    H, W = cur.shape
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    di, do = DeviceBuffer.from_array(np.ascontiguousarray(mv, np.int32)), DeviceBuffer(max(mv.size * 4, 4))
    me_subpel_refine(dc, dr, H, W, W, di, do)
    return do.download(mv.shape, np.int32)


# ---------------------------------------------------------------------------
# NV12 frames: full search over luma + chroma, and codec-facing MV formats
# ---------------------------------------------------------------------------
MV_Q4_INT16, MV_AVMOTIONVECTOR = 0, 1

# FFmpeg's AVMotionVector (libavutil/motion_vector.h), 40 bytes per record
AVMOTIONVECTOR_DTYPE = np.dtype({
    "names": ["source", "w", "h", "src_x", "src_y", "dst_x", "dst_y", "flags", "motion_x", "motion_y",
              "motion_scale"],
    "formats": [np.int32, np.uint8, np.uint8, np.int16, np.int16, np.int16, np.int16, np.uint64, np.int32, np.int32,
                np.uint16],
    "offsets": [0, 4, 5, 6, 8, 10, 12, 16, 24, 28, 32],
    "itemsize": 40})


def me_full_search_nv12(cur_y: DeviceBuffer, cur_uv: DeviceBuffer, ref_y: DeviceBuffer, ref_uv: DeviceBuffer,
                        H: int, W: int, pitch_y: int, pitch_uv: int, R: int, mv_cost: DeviceBuffer,
                        stream=None) -> None:
    """Full search over NV12 frames: cost = luma 8x8 SAD + chroma 4x4 U and V
    SAD (H.264 chroma interpolation at half-pel); mv_cost int32 [H/8][W/8][3]."""
    if mv_cost.nbytes < (H // 8) * (W // 8) * 12:
        raise ValueError("mv_cost buffer too small")
    for b, rows, pitch in ((cur_y, H, pitch_y), (ref_y, H, pitch_y), (cur_uv, H // 2, pitch_uv),
                           (ref_uv, H // 2, pitch_uv)):
        if b.nbytes < (rows - 1) * pitch + W:
            raise ValueError("plane buffer too small")
    call("umi_me_full_search_nv12", cur_y.ptr, cur_uv.ptr, ref_y.ptr, ref_uv.ptr, H, W, pitch_y, pitch_uv, R,
         mv_cost.ptr, _stream(stream))


def me_full_search_nv12_np(cur_y: np.ndarray, cur_uv: np.ndarray, ref_y: np.ndarray, ref_uv: np.ndarray,
                           R: int = 16) -> np.ndarray:
    H, W = cur_y.shape
    assert cur_uv.shape == (H // 2, W) and ref_uv.shape == cur_uv.shape and ref_y.shape == cur_y.shape
    bufs = [DeviceBuffer.from_array(np.ascontiguousarray(a, np.uint8)) for a in (cur_y, cur_uv, ref_y, ref_uv)]
    out = DeviceBuffer(max((H // 8) * (W // 8) * 12, 4))
    me_full_search_nv12(*bufs, H, W, W, W, R, out)
    return out.download((H // 8, W // 8, 3), np.int32)


def export_mvs(mv: DeviceBuffer, nby: int, nbx: int, B: int, units: int, fmt: int, out: DeviceBuffer,
               stream=None) -> None:
    """Motion vectors (dy, dx, cost) in pels (units 1) or quarter pels (units 4)
    to int16 quarter-pel (mvx, mvy) pairs (fmt 0) or FFmpeg AVMotionVector
    records (fmt 1)."""
    need = nby * nbx * (4 if fmt == MV_Q4_INT16 else 40)
    if out.nbytes < need or mv.nbytes < nby * nbx * 12:
        raise ValueError("motion vector buffers too small")
    call("umi_me_export_mvs", mv.ptr, nby, nbx, B, units, fmt, out.ptr, _stream(stream))


def export_mvs_np(mv: np.ndarray, B: int, units: int, fmt: int) -> np.ndarray:
    nby, nbx = mv.shape[:2]
    dm = DeviceBuffer.from_array(np.ascontiguousarray(mv, np.int32))
    if fmt == MV_Q4_INT16:
        out = DeviceBuffer(max(nby * nbx * 4, 4))
        export_mvs(dm, nby, nbx, B, units, fmt, out)
        return out.download((nby, nbx, 2), np.int16)
    out = DeviceBuffer(max(nby * nbx * 40, 4))
    export_mvs(dm, nby, nbx, B, units, fmt, out)
    return out.download((nby * nbx * 40,), np.uint8).view(AVMOTIONVECTOR_DTYPE).reshape(nby, nbx)


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


def stereo_wta(L: DeviceBuffer, R: DeviceBuffer, H: int, W: int, pitch: int, D: int, s0: int, step: int,
               KY: int, KX: int, out: DeviceBuffer, stream=None) -> None:
    """Winner-take-all shift per pixel over the cost volume of `cost_volume`
    (first minimum over i), without storing the volume: out int32 [H][W][2]
    = (i, cost)."""
    if out.nbytes < H * W * 8:
        raise ValueError("winner-take-all buffer too small")
    if L.nbytes < (H - 1) * pitch + W or R.nbytes < (H - 1) * pitch + W:
        raise ValueError("image buffer too small")
    call("umi_stereo_wta_u8", L.ptr, R.ptr, H, W, pitch, D, s0, step, KY, KX, out.ptr, _stream(stream))


def stereo_wta_np(L: np.ndarray, R: np.ndarray, D: int, s0: int = 0, step: int = -1, KY: int = 2,
                  KX: int = 2) -> np.ndarray:
    assert L.dtype == np.uint8 and R.dtype == np.uint8 and L.shape == R.shape
    H, W = L.shape
    dl, dr = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
    out = DeviceBuffer(H * W * 8)
    stereo_wta(dl, dr, H, W, W, D, s0, step, KY, KX, out)
    return out.download((H, W, 2), np.int32)


def sgm_workspace_bytes(H: int, W: int, D: int, s0: int = 0, step: int = -1, KY: int = 2, KX: int = 2) -> int:
    n = ctypes.c_size_t()
    call("umi_sgm_workspace_size", H, W, D, s0, step, KY, KX, ctypes.byref(n))
    return n.value


def sgm(L: DeviceBuffer, R: DeviceBuffer, H: int, W: int, pitch: int, D: int, s0: int, step: int, KY: int,
        KX: int, P1: int, P2: int, workspace: DeviceBuffer, out: DeviceBuffer, stream=None) -> None:
    """Semi-global matching (four paths, penalties P1/P2) over the cost volume
    of `cost_volume`, without storing it: out int32 [H][W][2] = (i, aggregated
    cost) with i the first minimum; workspace of `sgm_workspace_bytes` bytes."""
    if out.nbytes < H * W * 8:
        raise ValueError("SGM output buffer too small")
    if L.nbytes < (H - 1) * pitch + W or R.nbytes < (H - 1) * pitch + W:
        raise ValueError("image buffer too small")
    call("umi_sgm_u8", L.ptr, R.ptr, H, W, pitch, D, s0, step, KY, KX, P1, P2, workspace.ptr, workspace.nbytes,
         out.ptr, _stream(stream))


def sgm_np(L: np.ndarray, R: np.ndarray, D: int, s0: int = 0, step: int = -1, KY: int = 2, KX: int = 2,
           P1: int = 10, P2: int = 120) -> np.ndarray:
    assert L.dtype == np.uint8 and R.dtype == np.uint8 and L.shape == R.shape
    H, W = L.shape
    dl, dr = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
    ws = DeviceBuffer(sgm_workspace_bytes(H, W, D, s0, step, KY, KX))
    out = DeviceBuffer(H * W * 8)
    sgm(dl, dr, H, W, W, D, s0, step, KY, KX, P1, P2, ws, out)
    return out.download((H, W, 2), np.int32)


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


def iimg_set_config(rows_per_lane: int) -> None:
    """Tile height of the one-image integral image: rows per lane 1, 2, 4 (TH
    16, 32, 64) or 0 = chosen by size.  Results do not depend on it."""
    call("umi_iimg_set_config", rows_per_lane)


def iimg_workspace(H: int, W: int, dtype) -> DeviceBuffer:
    """A zero-filled workspace for iimg() (zeroed once; calls keep it ready)."""
    return DeviceBuffer.zeros(max(iimg_workspace_size(H, W, dtype), 16))


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
    ws = iimg_workspace(H, W, img.dtype) if use_workspace else None
    iimg(d, H, W, W, o, W, img.dtype, ws)
    return o.download((H, W), odt)


def box_sum(S: DeviceBuffer, H: int, W: int, in_pitch: int, ry: int, rx: int, out: DeviceBuffer, out_pitch: int,
            stream=None) -> None:
    """Box sums over (2 ry + 1) x (2 rx + 1) windows clipped to the image, from
    a uint32 integral image (mod 2^32)."""
    if S.nbytes < ((H - 1) * in_pitch + W) * 4 or out.nbytes < ((H - 1) * out_pitch + W) * 4:
        raise ValueError("box sum buffers too small")
    call("umi_box_sum_iimg_u32", S.ptr, H, W, in_pitch, ry, rx, out.ptr, out_pitch, _stream(stream))


def box_sum_np(S: np.ndarray, ry: int, rx: int) -> np.ndarray:
    S = np.ascontiguousarray(S, np.uint32)
    H, W = S.shape
    d, o = DeviceBuffer.from_array(S), DeviceBuffer(H * W * 4)
    box_sum(d, H, W, W, ry, rx, o, W)
    return o.download((H, W), np.uint32)


def haar_features(S: DeviceBuffer, H: int, W: int, pitch: int, feats: np.ndarray, win_h: int, win_w: int,
                  out: DeviceBuffer, stream=None) -> None:
    """Rectangle features (feats int32 [nfeat][3][5] = x, y, w, h, weight per
    rectangle) at every position of a win_h x win_w window, from a uint32
    integral image; out int32 [nfeat][H - win_h + 1][W - win_w + 1]."""
    feats = np.ascontiguousarray(feats, np.int32).reshape(-1, 3, 5)
    n = feats.shape[0]
    if out.nbytes < max(n * (H - win_h + 1) * (W - win_w + 1) * 4, 0):
        raise ValueError("feature buffer too small")
    call("umi_haar_features_u32", S.ptr, H, W, pitch, feats.ctypes.data, n, win_h, win_w, out.ptr,
         _stream(stream))


def haar_features_np(S: np.ndarray, feats: np.ndarray, win_h: int, win_w: int) -> np.ndarray:
    S = np.ascontiguousarray(S, np.uint32)
    H, W = S.shape
    feats = np.ascontiguousarray(feats, np.int32).reshape(-1, 3, 5)
    o = DeviceBuffer(max(feats.shape[0] * (H - win_h + 1) * (W - win_w + 1) * 4, 4))
    haar_features(DeviceBuffer.from_array(S), H, W, W, feats, win_h, win_w, o)
    return o.download((feats.shape[0], H - win_h + 1, W - win_w + 1), np.int32)


# ---------------------------------------------------------------------------
# MFMA GEMM / convolution (main/gemm.cu, main/fprop.cu), johnjohnlin/UMI
# ---------------------------------------------------------------------------
F32, BF16, BF16X3, FP8, FP8X3 = 0, 1, 2, 3, 4
MODE_NAMES = {F32: "f32", BF16: "bf16", BF16X3: "bf16x3", FP8: "fp8", FP8X3: "fp8x3"}
SPLIT = {BF16X3, FP8X3}
FP8_MODES = {FP8, FP8X3}


def elem_bytes(mode: int) -> int:
    import ctypes

    n = ctypes.c_int(0)
    call("umi_gemm_elem_bytes", mode, ctypes.byref(n))
    return n.value


class Operand:
    """A packed contraction operand: hi (and lo for split modes) plus its scale q."""

    def __init__(self, rows: int, k: int, mode: int):
        self.rows, self.k, self.mode = rows, k, mode
        eb = elem_bytes(mode)
        self.hi = DeviceBuffer(max(rows * k * eb, 16))
        self.lo = DeviceBuffer(max(rows * k * eb, 16)) if mode in SPLIT else None
        self.amax = DeviceBuffer(4)
        self.q = DeviceBuffer(4) if mode in FP8_MODES else None

    @property
    def lo_ptr(self):
        return self.lo.ptr if self.lo is not None else None

    @property
    def q_ptr(self):
        return self.q.ptr if self.q is not None else None


def gemm_pack(src: DeviceBuffer, R: int, K: int, ld_src: int, transpose: bool, mode: int, stream=None) -> Operand:
    op = Operand(R, K, mode)
    if mode in FP8_MODES:
        call("umi_amax_f32", src.ptr, R * K, op.amax.ptr, _stream(stream))
    call("umi_gemm_pack", src.ptr, R, K, ld_src, int(transpose), mode, op.amax.ptr, op.hi.ptr, op.lo_ptr, K,
         op.q_ptr, _stream(stream))
    return op


def gemm(mode: int, M: int, N: int, K: int, A: Operand, Bt: Operand, C: DeviceBuffer, stream=None) -> None:
    call("umi_gemm", mode, M, N, K, A.hi.ptr, A.lo_ptr, K, Bt.hi.ptr, Bt.lo_ptr, K, C.ptr, N, A.q_ptr, Bt.q_ptr,
         _stream(stream))


def gemm_np(A: np.ndarray, B: np.ndarray, mode: int, return_operands: bool = False):
    """C = A @ B for fp32 A (M x K) and B (K x N, the reference's row-major layout)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    dA, dB = DeviceBuffer.from_array(A), DeviceBuffer.from_array(B)
    oa = gemm_pack(dA, M, K, K, False, mode)
    ob = gemm_pack(dB, N, K, N, True, mode)
    dC = DeviceBuffer(M * N * 4)
    gemm(mode, M, N, K, oa, ob, dC)
    C = dC.download((M, N), np.float32)
    if return_operands:
        return C, download_operand(oa), download_operand(ob)
    return C


def download_operand(op: Operand):
    """(hi bytes, lo bytes or None, q or None) of a packed operand."""
    eb = elem_bytes(op.mode)
    dt = {4: np.float32, 2: np.uint16, 1: np.uint8}[eb]
    hi = op.hi.download((op.rows, op.k), dt)
    lo = op.lo.download((op.rows, op.k), dt) if op.lo is not None else None
    q = float(op.q.download((1,), np.float32)[0]) if op.q is not None else None
    return hi, lo, q


def conv_pack_input(src: DeviceBuffer, N: int, C: int, H: int, W: int, mode: int, stream=None) -> Operand:
    op = Operand(N * H * W, C, mode)
    if mode in FP8_MODES:
        call("umi_amax_f32", src.ptr, N * C * H * W, op.amax.ptr, _stream(stream))
    call("umi_conv_pack_input", src.ptr, N, C, H, W, mode, op.amax.ptr, op.hi.ptr, op.lo_ptr, op.q_ptr,
         _stream(stream))
    return op


def conv_pack_filter(src: DeviceBuffer, K: int, C: int, R: int, S: int, mode: int, stream=None) -> Operand:
    op = Operand(K, R * S * C, mode)
    if mode in FP8_MODES:
        call("umi_amax_f32", src.ptr, K * C * R * S, op.amax.ptr, _stream(stream))
    call("umi_conv_pack_filter", src.ptr, K, C, R, S, mode, op.amax.ptr, op.hi.ptr, op.lo_ptr, op.q_ptr,
         _stream(stream))
    return op


def conv_fprop(mode: int, N: int, C: int, H: int, W: int, K: int, R: int, S: int, stride: int, pad: int,
               x: Operand, w: Operand, out: DeviceBuffer, stream=None) -> None:
    call("umi_conv_fprop", mode, N, C, H, W, K, R, S, stride, pad, x.hi.ptr, x.lo_ptr, w.hi.ptr, w.lo_ptr, out.ptr,
         x.q_ptr, w.q_ptr, _stream(stream))


def conv_out_shape(N, K, H, W, R, S, stride, pad):
    return N, K, (H + 2 * pad - R) // stride + 1, (W + 2 * pad - S) // stride + 1


def conv_fprop_np(x: np.ndarray, w: np.ndarray, stride: int, pad: int, mode: int):
    """NCHW fp32 input, KCRS fp32 filter -> NCHW fp32 output."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    N, C, H, W = x.shape
    K, _, R, S = w.shape
    shp = conv_out_shape(N, K, H, W, R, S, stride, pad)
    dx, dw = DeviceBuffer.from_array(x), DeviceBuffer.from_array(w)
    ox = conv_pack_input(dx, N, C, H, W, mode)
    ow = conv_pack_filter(dw, K, C, R, S, mode)
    out = DeviceBuffer(int(np.prod(shp)) * 4)
    conv_fprop(mode, N, C, H, W, K, R, S, stride, pad, ox, ow, out)
    return out.download(shp, np.float32)
