"""The C ABI: libumi_hip.so loads and exports every function include/*.h declares.

No compute call is made here (no GPU needed): every call below is refused by
the library's host-side checks.  Pointer arguments that are not the subject of
a null check are real device buffers of the requested size when a device is
visible, so a check that regressed would launch on valid memory and show up as
a wrong status code, never as a fault.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(umi_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "umi_me_full_search_u8" in names and "umi_malloc" in names


def test_library_exports_every_declared_symbol():
    from umi_amd._lib import LIB_PATH, SIGNATURES

    assert LIB_PATH.exists(), "libumi_hip.so not built"
    lib = ctypes.CDLL(str(LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in declared_functions() if n not in SIGNATURES and n != "umi_last_error"]
    assert not unbound, f"declared but not bound in umi_amd._lib: {unbound}"


def _buffers(*sizes):
    """Device buffers of the given byte sizes when a device is visible, else
    distinct non-null placeholders (with no device nothing can be launched)."""
    import umi_amd
    from umi_amd import DeviceBuffer

    if umi_amd.device_count() > 0:
        bufs = [DeviceBuffer(max(n, 4)) for n in sizes]
        return bufs, [b.ptr for b in bufs]
    return [], [0x1000 * (i + 1) for i in range(len(sizes))]


def test_errors_are_status_codes_not_aborts():
    # argument validation happens before any device call
    from umi_amd._lib import lib

    L = lib()
    assert L.umi_me_full_search_u8(None, None, 8, 8, 8, 8, 4, None, None, None) == 1
    assert b"null" in L.umi_last_error()
    keep, (cur, ref, mv) = _buffers(64 * 64, 64 * 64, 64 * 3 * 4)
    assert L.umi_me_full_search_u8(cur, ref, 64, 64, 64, 12, 4, mv, None, None) == 3  # B = 12 unsupported
    keep2, (l, r, cv) = _buffers(8 * 8, 8 * 8, 8 * 8 * 4 * 4)
    assert L.umi_cost_volume_u8(l, r, 8, 8, 8, 4, 0, 2, 1, 1, cv, None) == 1  # step 2 unsupported
    # a feature rectangle outside its window is refused on the host
    import numpy as np

    feats = np.zeros((1, 3, 5), np.int32)
    feats[0, 0] = (20, 0, 8, 4, 1)  # x + w = 28 > 24
    keep3, (S, out) = _buffers(64 * 64 * 4, 41 * 41 * 4)
    assert L.umi_haar_features_u32(S, 64, 64, 64, feats.ctypes.data, 1, 24, 24, out, None) == 1
    assert b"outside" in L.umi_last_error()


def test_gemm_form_switch_rejects_unknown_values():
    # the measurement switch of the 256 x 256 GEMM takes 0, 1, 2, 4 or 5 only
    from umi_amd._lib import lib

    L = lib()
    for bad in (-1, 3, 6, 8, 9):
        assert L.umi_gemm_set_lds_direct(bad) == 1
        assert b"umi_gemm_set_lds_direct" in L.umi_last_error()
    assert L.umi_gemm_set_lds_direct(1) == 0


def _pack(pd, ad):
    import numpy as np

    return np.array(pd, np.int32).reshape(16), np.array(ad, np.int32).reshape(8)


def _validate(pd, ad, expand=0x1111, views=None, phys=None, planes=None, n_in=0, n_out=0, zso=0):
    import numpy as np

    from umi_amd._lib import lib

    p, a = _pack(pd, ad)
    v = np.zeros(max(1, 24 * (n_in + n_out)), np.int32) if views is None else np.asarray(views, np.int32)
    ph = np.ones(max(1, 8 * (n_in + n_out)), np.int32) if phys is None else np.asarray(phys, np.int32)
    if phys is None:
        ph[1::2] = 0
    pl = np.ones(max(1, n_in + n_out), np.int32) if planes is None else np.asarray(planes, np.int32)
    return lib().umi_validate_execute(p.ctypes.data, a.ctypes.data, n_in, n_out, v.ctypes.data, ph.ctypes.data,
                                      pl.ctypes.data, expand, zso)


GOOD_P = [[1, 1, 1, 1], [1, 1, 1, 1], [300, 16, 4, 1], [500, 16, 8, 1]]
GOOD_A = [[1, 1], [1, 1], [3, 3], [3, 3]]


def test_validate_execute_accepts_reference_dims():
    assert _validate(GOOD_P, GOOD_A, n_in=1, n_out=1) == 0
    # main/iimg.cu:60 passes 3-element ParallelDims (warp_shuf 0): accepted as 1
    assert _validate([[1, 1, 1, 0], [1, 1, 1, 0], [1, 1, 1, 0], [500, 128, 32, 0]], GOOD_A) == 0


@pytest.mark.parametrize("pd,why", [
    ([[1, 1, 1, 1], [1, 1, 1, 1], [300, 16, 2, 1], [500, 16, 8, 1]], b"not 32"),
    ([[1, 1, 1, 1], [1, 1, 1, 1], [300, 16, 4, 3], [500, 16, 8, 1]], b"power of two"),
    ([[1, 1, 1, 1], [1, 1, 1, 1], [300, 12, 8, 1], [500, 16, 4, 1]], b"multiple"),
    ([[1, 1, 1, 1], [1, 1, 1, 1], [-3, 16, 4, 1], [500, 16, 8, 1]], b"total_size"),
])
def test_validate_execute_rejects(pd, why):
    from umi_amd._lib import lib

    assert _validate(pd, GOOD_A) == 1
    assert why in lib().umi_last_error()


def test_validate_execute_rejects_bad_views_and_sizes():
    import numpy as np

    from umi_amd._lib import lib

    views = np.zeros(24, np.int32)
    views[3] = 7  # p view 1 -> physical dim 7
    assert _validate(GOOD_P, GOOD_A, views=views, n_in=1) == 1
    phys = np.array([1, 0, 0, 0, 5, 0, 5, 0], np.int32)  # input with a zero-size dim
    assert _validate(GOOD_P, GOOD_A, phys=phys, n_in=1) == 1
    assert b"size" in lib().umi_last_error()
    assert _validate(GOOD_P, GOOD_A, planes=[5], n_in=1) == 1
    assert _validate(GOOD_P, GOOD_A, expand=0x1101) == 1
    assert _validate(GOOD_P, [[1, 1], [1, 0], [3, 3], [3, 3]]) == 1


def test_validate_zero_stride_opt_against_views():
    # main/wfilter3x3.cu:39: weights (input 1) do not move with the expanded dims 2, 3 -> 0xc0 is sound
    from umi_amd._lib import lib

    views = np.zeros(48, np.int32)
    views[3 * 2:3 * 2 + 3] = (2, 0, 1)  # input 0, p view dim 2 -> physical 2, stride 1
    views[3 * 3:3 * 3 + 3] = (3, 0, 1)  # input 0, p view dim 3 -> physical 3, stride 1
    assert _validate(GOOD_P, GOOD_A, expand=0x1122, views=views, n_in=2, zso=0xc0) == 0
    # the same bit pattern on input 0, whose view moves along dims 2 and 3: refused, not silently wrong
    assert _validate(GOOD_P, GOOD_A, expand=0x1122, views=views, n_in=2, zso=0x0c) == 1
    assert b"ZeroStrideOpt" in lib().umi_last_error()
    # without expansion along the dim the bit shares nothing and is accepted
    assert _validate(GOOD_P, GOOD_A, expand=0x1111, views=views, n_in=2, zso=0x0c) == 0


def test_validate_block_size_limit():
    from umi_amd._lib import lib

    big = [[1, 1, 1, 1], [1, 1, 1, 1], [300, 64, 4, 1], [500, 32, 8, 1]]  # 64 x 32 = 2048 threads
    assert _validate(big, GOOD_A) == 1
    assert b"1024" in lib().umi_last_error()


def test_batch_create_validates_on_host():
    import ctypes

    import numpy as np

    from umi_amd._lib import lib

    H = np.array([10, 0], np.int32)
    W = np.array([10, 10], np.int32)
    off = np.array([0, 100], np.int64)
    h = ctypes.c_void_p()
    assert lib().umi_batch_create(2, H.ctypes.data, W.ctypes.data, off.ctypes.data, W.ctypes.data, off.ctypes.data,
                                  W.ctypes.data, ctypes.byref(h)) == 1
    assert b"image 1" in lib().umi_last_error()
    n = ctypes.c_int()
    assert lib().umi_batch_configs(0, ctypes.byref(n)) == 0 and n.value >= 1
    assert lib().umi_batch_configs(7, ctypes.byref(n)) == 1


def test_sgm_validates_on_host():
    from umi_amd._lib import lib

    L = lib()
    ws_n = ctypes.c_size_t()
    assert L.umi_sgm_workspace_size(10, 20, 30, 0, -1, 2, 2, ctypes.byref(ws_n)) == 0
    assert ws_n.value >= 10 * 20 * 30 * 2  # at least the uint16 partial sums
    assert L.umi_sgm_workspace_size(10, 20, 30, 0, 3, 2, 2, ctypes.byref(ws_n)) == 1  # step 3
    assert L.umi_sgm_workspace_size(16, 16, 8, 0, -1, 2, 2, ctypes.byref(ws_n)) == 0
    need = ws_n.value
    keep, (l, r, ws, out) = _buffers(16 * 16, 16 * 16, need, 16 * 16 * 8)
    ok = (l, r, 16, 16, 16, 8, 0, -1, 2, 2)
    big = need
    assert L.umi_sgm_u8(*ok, 10, 120, None, big, out, None) == 1  # no workspace
    assert L.umi_sgm_u8(l, r, 16, 16, 16, 129, 0, -1, 2, 2, 10, 120, ws, big, out, None) == 3  # D > 128
    assert L.umi_sgm_u8(l, r, 16, 16, 16, 8, 0, 2, 2, 2, 10, 120, ws, big, out, None) == 1  # step 2
    assert L.umi_sgm_u8(l, r, 16, 16, 16, 8, 0, -1, 4, 2, 10, 120, ws, big, out, None) == 3  # KY = 4
    assert L.umi_sgm_u8(*ok, 130, 120, ws, big, out, None) == 1  # P1 > P2
    assert L.umi_sgm_u8(*ok, 10, 12000, ws, big, out, None) == 1  # partial sums would overflow uint16
    assert b"65536" in L.umi_last_error()
    assert L.umi_sgm_u8(*ok, 10, 120, ws, need - 1, out, None) == 1  # workspace one byte short
    assert b"workspace" in L.umi_last_error()


def test_nv12_and_export_validate_on_host():
    from umi_amd._lib import lib

    L = lib()
    keep, (y, uv, mv, out) = _buffers(64 * 64, 32 * 64, 64 * 12, 64 * 40)
    assert L.umi_me_full_search_nv12(y, uv, y, uv, 64, 64, 64, 64, 17, mv, None) == 3  # R > 16
    assert L.umi_me_full_search_nv12(y, uv, y, uv, 63, 64, 64, 64, 8, mv, None) == 1  # odd height
    assert L.umi_me_full_search_nv12(y, uv, y, uv, 64, 64, 64, 32, 8, mv, None) == 1  # uv pitch < W
    assert L.umi_me_full_search_nv12(y, None, y, uv, 64, 64, 64, 64, 8, mv, None) == 1
    assert L.umi_me_export_mvs(mv, 8, 8, 8, 2, 0, out, None) == 1  # units 2
    assert L.umi_me_export_mvs(mv, 8, 8, 8, 1, 2, out, None) == 1  # format 2
    assert L.umi_me_export_mvs(mv, 8, 8, 0, 1, 0, out, None) == 1  # block 0
    assert L.umi_me_export_mvs(mv, 4096, 8, 16, 1, 1, out, None) == 3  # positions past int16
