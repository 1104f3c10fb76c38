"""ctypes loader of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The library never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_PATH = Path(__file__).resolve().parent / "lib" / "liboracle.so"
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_SIGS = {
    "oracle_me_full_search_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P],
    "oracle_sad_volume_i32": [_P, _P, _I, _I, _I, _I, _I, _P],
    "oracle_cost_volume_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P],
    "oracle_stereo_wta_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P],
    "oracle_box_sum_u8": [_P, _I, _I, _I, _I, _I, _P],
    "oracle_me_subpel_u8": [_P, _P, _I, _I, _I, _P, _P],
    "oracle_haar_u8": [_P, _I, _I, _I, _P, _I, _I, _I, _P],
    "oracle_filter3x3_i32": [_P, _I, _I, _I, _I, _I, _I, _I, _P, _I],
    "oracle_filter3x3_f32": [_P, _I, _I, _I, _F, _F, _F, _I, _P, _I],
    "oracle_wfilter3x3_i32": [_P, _I, _I, _I, _P, _I, _I, _I, _P, _I],
    "oracle_wfilter3x3_f32": [_P, _I, _I, _I, _P, _F, _F, _I, _P, _I],
    "oracle_wfilter3x3_f64": [_P, _I, _I, _I, _P, _F, _F, _I, _P, _P],
    "oracle_bilateral_f32": [_P, _I, _I, _I, _I, _P, _F, _F, _I, _P, _I],
    "oracle_bilateral_f64": [_P, _I, _I, _I, _I, _P, _F, _F, _I, _P],
    "oracle_iimg_u8_u32": [_P, _I, _I, _I, _P],
    "oracle_iimg_i32": [_P, _I, _I, _I, _P],
    "oracle_iimg_f64": [_P, _I, _I, _I, _P],
    "oracle_set_threads": [_I],
    "oracle_sgm_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P],
    "oracle_me_nv12_u8": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P],
    "oracle_nv12_chroma_pred": [_P, _I, _I, _I, _I, _I, _I, _I, _I],
    "oracle_gemm_f32_rows": [_P, _P, _P, _I, _I, _I, _I, _I],
    "oracle_conv_f32": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I],
}
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} is not built; run `make -C oracle`")
        L = C.CDLL(str(ORACLE_PATH))
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the row-parallel loops (1 = the reference's single thread)."""
    lib().oracle_set_threads(int(n))


def gemm_f32_rows(A, B, r0, r1):
    """Rows [r0, r1) of A @ B by the loop of main/gemm.cu:69-78 in float (fmaf)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), np.float32)
    lib().oracle_gemm_f32_rows(_ptr(A), _ptr(B), _ptr(C), M, N, K, r0, r1)
    return C


def conv_f32(x, w, stride, pad, n0=0, n1=None):
    """Images [n0, n1) of the direct convolution of main/fprop.cu:84-108 in float (fmaf)."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    N, C, H, W = x.shape
    K, _, R, S = w.shape
    OH, OW = (H + 2 * pad - R) // stride + 1, (W + 2 * pad - S) // stride + 1
    out = np.zeros((N, K, OH, OW), np.float32)
    lib().oracle_conv_f32(_ptr(x), _ptr(w), _ptr(out), N, C, H, W, K, R, S, stride, pad, n0, N if n1 is None else n1)
    return out


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def me_full_search(cur, ref, B=8, R=16, by_range=None, volume=False):
    """Oracle ME: returns (mv_sad [nby][nbx][3], volume or None); rows by_range only."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    H, W = cur.shape
    nby, nbx, S = H // B, W // B, 2 * R + 1
    b0, b1 = by_range if by_range is not None else (0, nby)
    mv = np.zeros((nby, nbx, 3), np.int32)
    vol = np.zeros((nby, nbx, S, S), np.int32) if volume else None
    lib().oracle_me_full_search_u8(_ptr(cur), _ptr(ref), H, W, W, B, R, b0, b1, _ptr(mv), _ptr(vol))
    return mv, vol


def sad_volume_i32(cur, ref, B, S, ofs):
    cur = np.ascontiguousarray(cur, np.int32)
    ref = np.ascontiguousarray(ref, np.int32)
    H, W = cur.shape
    out = np.zeros((H // B, W // B, S, S), np.int32)
    lib().oracle_sad_volume_i32(_ptr(cur), _ptr(ref), H, W, B, S, ofs, _ptr(out))
    return out


def cost_volume(L, R, D, s0=0, step=-1, KY=2, KX=2, y_range=None):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    y0, y1 = y_range if y_range is not None else (0, H)
    out = np.zeros((H, W, D), np.int32)
    lib().oracle_cost_volume_u8(_ptr(L), _ptr(R), H, W, W, D, s0, step, KY, KX, y0, y1, _ptr(out))
    return out


def stereo_wta(L, R, D, s0=0, step=-1, KY=2, KX=2, y_range=None):
    """Per pixel (first minimising shift index, its cost) over the cost volume."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    y0, y1 = y_range if y_range is not None else (0, H)
    out = np.zeros((H, W, 2), np.int32)
    lib().oracle_stereo_wta_u8(_ptr(L), _ptr(R), H, W, W, D, s0, step, KY, KX, y0, y1, _ptr(out))
    return out


def me_nv12(cur_y, cur_uv, ref_y, ref_uv, R=16):
    """Full-search ME over NV12 frames (luma 8x8 + chroma 4x4 U and V, H.264
    chroma interpolation at half-pel): int32 [H/8][W/8][3] = (dy, dx, cost)."""
    cur_y, ref_y = (np.ascontiguousarray(a, np.uint8) for a in (cur_y, ref_y))
    cur_uv, ref_uv = (np.ascontiguousarray(a, np.uint8) for a in (cur_uv, ref_uv))
    H, W = cur_y.shape
    assert cur_uv.shape == (H // 2, W) and ref_uv.shape == cur_uv.shape and ref_y.shape == cur_y.shape
    out = np.zeros((H // 8, W // 8, 3), np.int32)
    lib().oracle_me_nv12_u8(_ptr(cur_y), _ptr(cur_uv), _ptr(ref_y), _ptr(ref_uv), H, W, W, W, R, _ptr(out))
    return out


def nv12_chroma_pred(uv, ch, Y, X, fy, fx):
    """The chroma predictor of me_nv12: eighth-pel fractions fy, fx at integer
    chroma position (Y, X) of plane ch (0 = U, 1 = V) of an NV12 uv plane."""
    uv = np.ascontiguousarray(uv, np.uint8)
    f = lib().oracle_nv12_chroma_pred
    f.restype = C.c_int
    return f(_ptr(uv), uv.shape[0], uv.shape[1] // 2, uv.shape[1], ch, Y, X, fy, fx)


def sgm(L, R, D, s0=0, step=-1, KY=2, KX=2, P1=10, P2=120):
    """Semi-global matching over the cost volume (four paths): per pixel (first
    index with the least aggregated cost, that cost), int32 [H][W][2]."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    out = np.zeros((H, W, 2), np.int32)
    lib().oracle_sgm_u8(_ptr(L), _ptr(R), H, W, W, D, s0, step, KY, KX, P1, P2, _ptr(out))
    return out


def me_subpel(cur, ref, mv):
    """Quarter-pel refinement of integer 8x8 block vectors mv [H/8][W/8][3];
    returns (dy, dx, SAD) in quarter pels."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    mv = np.ascontiguousarray(mv, np.int32)
    H, W = cur.shape
    out = np.zeros((H // 8, W // 8, 3), np.int32)
    lib().oracle_me_subpel_u8(_ptr(cur), _ptr(ref), H, W, W, _ptr(mv), _ptr(out))
    return out

def haar(img, feats, win_h, win_w):
    """Brute-force rectangle features (feats int32 [nfeat][3][5]) at every window position, int32."""
    img = np.ascontiguousarray(img, np.uint8)
    feats = np.ascontiguousarray(feats, np.int32).reshape(-1, 3, 5)
    H, W = img.shape
    out = np.zeros((feats.shape[0], H - win_h + 1, W - win_w + 1), np.int32)
    lib().oracle_haar_u8(_ptr(img), H, W, W, _ptr(feats), feats.shape[0], win_h, win_w, _ptr(out))
    return out

def box_sum(img, ry, rx):
    """Brute-force box sums (window clipped to the image) of a uint8 image, uint32."""
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    out = np.zeros((H, W), np.uint32)
    lib().oracle_box_sum_u8(_ptr(img), H, W, W, ry, rx, _ptr(out))
    return out

def filter3x3(img, init, mul, pad, mode=0):
    dt = np.int32 if img.dtype == np.int32 else np.float32
    img = np.ascontiguousarray(img, dt)
    H, W = img.shape
    out = np.zeros_like(img)
    f = lib().oracle_filter3x3_i32 if dt == np.int32 else lib().oracle_filter3x3_f32
    f(_ptr(img), H, W, W, init, mul, pad, mode, _ptr(out), W)
    return out


def wfilter3x3(img, w, init, pad, mode=0):
    dt = np.int32 if img.dtype == np.int32 else np.float32
    img = np.ascontiguousarray(img, dt)
    w = np.ascontiguousarray(w, dt).reshape(9)
    H, W = img.shape
    out = np.zeros_like(img)
    f = lib().oracle_wfilter3x3_i32 if dt == np.int32 else lib().oracle_wfilter3x3_f32
    f(_ptr(img), H, W, W, _ptr(w), init, pad, mode, _ptr(out), W)
    return out


def wfilter3x3_f64(img, w, init, pad, mode=0):
    """(value, magnitude sum |w x| + |init|) in double."""
    img = np.ascontiguousarray(img, np.float32)
    w = np.ascontiguousarray(w, np.float32).reshape(9)
    H, W = img.shape
    out = np.zeros((H, W), np.float64)
    mag = np.zeros((H, W), np.float64)
    lib().oracle_wfilter3x3_f64(_ptr(img), H, W, W, _ptr(w), init, pad, mode, _ptr(out), _ptr(mag))
    return out, mag


def bilateral(img, r, ws, norm, pad=0.0, mode=0, f64=False):
    img = np.ascontiguousarray(img, np.float32)
    ws = np.ascontiguousarray(ws, np.float32)
    H, W = img.shape
    if f64:
        out = np.zeros((H, W), np.float64)
        lib().oracle_bilateral_f64(_ptr(img), H, W, W, r, _ptr(ws), norm, pad, mode, _ptr(out))
    else:
        out = np.zeros((H, W), np.float32)
        lib().oracle_bilateral_f32(_ptr(img), H, W, W, r, _ptr(ws), norm, pad, mode, _ptr(out), W)
    return out


def iimg(img):
    """u8 -> u32, i32 -> i32 (wrapping), f32 -> f64 reference."""
    H, W = img.shape
    img = np.ascontiguousarray(img)
    if img.dtype == np.uint8:
        out = np.zeros((H, W), np.uint32)
        lib().oracle_iimg_u8_u32(_ptr(img), H, W, W, _ptr(out))
    elif img.dtype == np.int32:
        out = np.zeros((H, W), np.int32)
        lib().oracle_iimg_i32(_ptr(img), H, W, W, _ptr(out))
    else:
        img = np.ascontiguousarray(img, np.float32)
        out = np.zeros((H, W), np.float64)
        lib().oracle_iimg_f64(_ptr(img), H, W, W, _ptr(out))
    return out


# ---------------------------------------------------------------------------
# FP64 oracles of the dense contractions (main/gemm.cu:69-78 and
# main/fprop.cu:84-108 restated in float64 numpy) and decoders of the packed
# operand formats, for error statistics.
# ---------------------------------------------------------------------------
def decode_bf16(u16: np.ndarray) -> np.ndarray:
    return (u16.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _e4m3_table() -> np.ndarray:
    t = np.zeros(256, np.float64)
    for b in range(256):
        s = -1.0 if b & 0x80 else 1.0
        e = (b >> 3) & 0xF
        m = b & 7
        if e == 15 and m == 7:
            t[b] = np.nan
        elif e == 0:
            t[b] = s * (m / 8.0) * 2.0 ** -6
        else:
            t[b] = s * (1 + m / 8.0) * 2.0 ** (e - 7)
    return t


E4M3 = _e4m3_table()


def decode_e4m3(u8: np.ndarray) -> np.ndarray:
    return E4M3[u8]


def gemm_f64(A, B):
    """C = A @ B in float64 (the row-major A[m][k], B[k][n] of main/gemm.cu)."""
    return np.asarray(A, np.float64) @ np.asarray(B, np.float64)


def im2col_f64(x, R, S, stride, pad):
    N, C, H, W = x.shape
    OH, OW = (H + 2 * pad - R) // stride + 1, (W + 2 * pad - S) // stride + 1
    xp = np.pad(np.asarray(x, np.float64), ((0, 0), (0, 0), (pad, pad), (pad, pad)))
    cols = np.empty((N, OH, OW, C, R, S), np.float64)
    for r in range(R):
        for s in range(S):
            cols[:, :, :, :, r, s] = xp[:, :, r:r + stride * OH:stride, s:s + stride * OW:stride].transpose(0, 2, 3, 1)
    return cols.reshape(N * OH * OW, C * R * S), (N, OH, OW)


def conv_f64(x, w, stride, pad):
    """out[n][k][oy][ox] = sum_{c,r,s} w[k][c][r][s] x[n][c][oy*s+r-p][ox*s+s-p], zero padding, float64."""
    K, C, R, S = w.shape
    cols, (N, OH, OW) = im2col_f64(x, R, S, stride, pad)
    out = cols @ np.asarray(w, np.float64).reshape(K, C * R * S).T
    return out.reshape(N, OH, OW, K).transpose(0, 3, 1, 2)


def conv_direct_i32(x, w, stride, pad):
    """The loop of main/fprop.cu:84-108 (int32 wrap), for small shapes."""
    N, C, H, W = x.shape
    K, _, R, S = w.shape
    OH, OW = (H + 2 * pad - R) // stride + 1, (W + 2 * pad - S) // stride + 1
    out = np.zeros((N, K, OH, OW), np.int64)
    for n in range(N):
        for k in range(K):
            for oy in range(OH):
                for ox in range(OW):
                    acc = 0
                    for c in range(C):
                        for r in range(R):
                            for s in range(S):
                                yy, xx = oy * stride + r - pad, ox * stride + s - pad
                                if 0 <= yy < H and 0 <= xx < W:
                                    acc += int(w[k, c, r, s]) * int(x[n, c, yy, xx])
                    out[n, k, oy, ox] = acc
    return (out & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def contraction_error(C, ref, A_rows_norm, B_cols_norm):
    """max / mean of |C - ref| / (||a_m|| ||b_n||), and max |C - ref| / max |ref|."""
    err = np.abs(np.asarray(C, np.float64) - ref)
    scale = np.outer(A_rows_norm, B_cols_norm)
    rel = err / np.maximum(scale, 1e-300)
    return {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
            "max_abs_over_max_ref": float(err.max() / np.abs(ref).max())}
