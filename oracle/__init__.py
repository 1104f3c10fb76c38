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
_SIGS = {
    "oracle_me_full_search_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P],
    "oracle_sad_volume_i32": [_P, _P, _I, _I, _I, _I, _I, _P],
    "oracle_cost_volume_u8": [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P],
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
