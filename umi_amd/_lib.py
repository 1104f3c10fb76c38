"""ctypes binding of libumi_hip.so (the C ABI declared in include/umi_hip.h).

The library must have been built (``python -m umi_amd.build`` or
``__graft_entry__.build()``).  There is no fallback: if it is missing or a
call fails, an exception says so.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libumi_hip.so"

c_int, c_size_t, c_void_p, c_float = C.c_int, C.c_size_t, C.c_void_p, C.c_float
P = C.c_void_p  # device pointers travel as integers

# name -> argument types (return type is always umi_status = int unless noted)
SIGNATURES = {
    "umi_device_count": [C.POINTER(c_int)],
    "umi_set_device": [c_int],
    "umi_device_name": [C.c_char_p, c_size_t],
    "umi_malloc": [C.POINTER(c_void_p), c_size_t],
    "umi_free": [P],
    "umi_host_alloc": [C.POINTER(c_void_p), c_size_t],
    "umi_host_free": [P],
    "umi_memcpy": [P, P, c_size_t, c_int],
    "umi_memcpy_async": [P, P, c_size_t, c_int, P],
    "umi_memset": [P, c_int, c_size_t],
    "umi_fill_u32": [P, C.c_uint32, c_size_t, P],
    "umi_synchronize": [],
    "umi_copy_d2d": [P, P, c_size_t, P],
    "umi_stream_create": [C.POINTER(c_void_p)],
    "umi_stream_destroy": [P],
    "umi_stream_synchronize": [P],
    "umi_stream_wait": [P, P],
    "umi_launch_kernel": [P, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, c_size_t,
                          C.POINTER(c_void_p), P],
    "umi_timer_create": [C.POINTER(c_void_p)],
    "umi_timer_destroy": [P],
    "umi_timer_start": [P, P],
    "umi_timer_stop": [P, P],
    "umi_timer_elapsed_ms": [P, C.POINTER(c_float)],
    "umi_graph_begin": [P],
    "umi_graph_end": [P, C.POINTER(c_void_p)],
    "umi_graph_launch": [P, P],
    "umi_graph_destroy": [P],
    "umi_validate_execute": [P, P, c_int, c_int, P, P, P, C.c_uint64, C.c_uint64],
    "umi_me_set_variant": [c_int],
    "umi_me_full_search_u8": [P, P, c_int, c_int, c_int, c_int, c_int, P, P, P],
    "umi_cost_volume_u8": [P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, P],
    "umi_stereo_wta_u8": [P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, P],
    "umi_me_full_search_nv12": [P, P, P, P, c_int, c_int, c_int, c_int, c_int, P, P],
    "umi_me_export_mvs": [P, c_int, c_int, c_int, c_int, c_int, P, P],
    "umi_sgm_workspace_size": [c_int, c_int, c_int, c_int, c_int, c_int, c_int, P],
    "umi_sgm_u8": [P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, c_size_t, P, P],
    "umi_box_sum_iimg_u32": [P, c_int, c_int, c_int, c_int, c_int, P, c_int, P],
    "umi_haar_features_u32": [P, c_int, c_int, c_int, P, c_int, c_int, c_int, P, P],
    "umi_me_subpel_refine_u8": [P, P, c_int, c_int, c_int, P, P, P],
    "umi_conv_set_direct": [c_int],
    "umi_gemm_set_lds_direct": [c_int],
    "umi_batch_box_sum_u32": [P, P, c_int, c_int, P, P],
    "umi_filter3x3_f32": [P, c_int, c_int, c_int, c_float, c_float, c_float, c_int, P, c_int, P],
    "umi_filter3x3_i32": [P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, c_int, P],
    "umi_wfilter3x3_f32": [P, c_int, c_int, c_int, P, c_float, c_float, c_int, P, c_int, P],
    "umi_wfilter3x3_i32": [P, c_int, c_int, c_int, P, c_int, c_int, c_int, P, c_int, P],
    "umi_bilateral_f32": [P, c_int, c_int, c_int, c_int, P, c_float, c_float, c_int, P, c_int, P],
    "umi_iimg_workspace_size": [c_int, c_int, c_int, C.POINTER(c_size_t)],
    "umi_iimg_set_config": [c_int],
    "umi_iimg_set_grid": [c_int],
    "umi_iimg_workspace_status": [P, C.POINTER(c_int)],
    "umi_iimg_u8_u32": [P, c_int, c_int, c_int, P, c_int, P, c_size_t, P],
    "umi_iimg_u8_u32_noalloc": [P, c_int, c_int, c_int, P, c_int, P],
    "umi_iimg_i32": [P, c_int, c_int, c_int, P, c_int, P, c_size_t, P],
    "umi_iimg_f32": [P, c_int, c_int, c_int, P, c_int, P, c_size_t, P],
    "umi_gemm_elem_bytes": [c_int, C.POINTER(c_int)],
    "umi_amax_f32": [P, c_size_t, P, P],
    "umi_gemm_pack": [P, c_int, c_int, c_int, c_int, c_int, P, P, P, c_int, P, P],
    "umi_gemm": [c_int, c_int, c_int, c_int, P, P, c_int, P, P, c_int, P, c_int, P, P, P],
    "umi_conv_pack_input": [P, c_int, c_int, c_int, c_int, c_int, P, P, P, P, P],
    "umi_conv_pack_filter": [P, c_int, c_int, c_int, c_int, c_int, P, P, P, P, P],
    "umi_batch_create": [c_int, P, P, P, P, P, P, C.POINTER(c_void_p)],
    "umi_batch_destroy": [P],
    "umi_batch_configs": [c_int, C.POINTER(c_int)],
    "umi_batch_wfilter3x3_f32": [P, P, P, c_float, c_float, c_int, P, c_int, P],
    "umi_batch_bilateral_f32": [P, P, c_int, P, c_float, c_float, c_int, P, c_int, P],
    "umi_batch_iimg_workspace_size": [P, c_int, C.POINTER(c_size_t)],
    "umi_batch_iimg_u8_u32": [P, P, P, P, c_size_t, c_int, P],
    "umi_batch_iimg_f32": [P, P, P, P, c_size_t, c_int, P],
    "umi_me_full_search_listed_u8": [P, P, c_int, c_int, c_int, c_int, c_int, P, P, c_int, P, P],
    "umi_synth_frame0": [P, c_int, c_int, c_int, C.c_uint64, P],
    "umi_synth_warp": [P, P, c_int, c_int, c_int, P, c_float, C.c_uint64, P],
    "umi_cnn_downsample_pair": [P, P, c_int, c_int, c_int, P, c_int, P],
    "umi_cnn_conv_fwd": [P, c_int, c_int, c_int, c_int, c_int, c_int, P, P, c_int, P, P],
    "umi_cnn_conv_bwd_data": [P, P, c_int, c_int, c_int, c_int, c_int, c_int, P, c_int, P, P],
    "umi_cnn_conv_bwd_weight_workspace": [c_int, c_int, c_int, c_int, c_int, C.POINTER(c_size_t)],
    "umi_cnn_conv_bwd_weight": [P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, P, c_size_t, P],
    "umi_cnn_huber": [P, P, c_int, c_int, c_int, P, P, C.POINTER(c_int), P],
    "umi_adam": [P, P, P, P, c_int, c_float, c_float, c_float, c_float, c_int, P],
    "umi_cnn_packed_bytes": [c_int, P],
    "umi_cnn_pack_layer": [c_int, P, P, P],
    "umi_cnn_infer_layer": [c_int, P, c_int, c_int, c_int, c_int, P, P, P, P],
    "umi_cnn_first_layer_frames": [P, P, c_int, c_int, c_int, P, P, P, P],
    "umi_cnn_first_two_layers_frames": [P, P, c_int, c_int, c_int, P, P, c_int, P, P, P, P],
    "umi_cnn_last_layer_proposals": [P, c_int, c_int, c_int, P, P, c_int, P, P, P],
    "umi_me_proposals": [P, c_int, c_int, c_int, P, P],
    "umi_me_refine_workspace": [c_int, c_int, c_int, P],
    "umi_me_refine": [P, P, c_int, c_int, c_int, c_int, P, c_int, c_int, c_int, c_float, P, P, P, P, c_size_t, P],
    "umi_me_refine_fused_workspace": [c_int, c_int, c_int, C.POINTER(c_size_t)],
    "umi_me_refine_fused": [P, P, c_int, c_int, c_int, c_int, P, c_int, c_int, c_int, c_float, P, P, P, P, c_size_t,
                            P],
    "umi_me_refine_fused_status": [P, c_int, P, C.POINTER(c_int)],
    "umi_conv_fprop": [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P, P, P, P, P, P, P,
                       P],
}


class UmiError(RuntimeError):
    """A libumi_hip call returned a non-zero status (`status`; 0 when the
    error did not come from a call)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


UMI_ERR_UNSUPPORTED = 3


_lib = None


def lib() -> C.CDLL:
    """Load libumi_hip.so once and attach the argument types."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise UmiError(f"{LIB_PATH} is not built; run `python -m umi_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, argtypes in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = c_int
        L.umi_last_error.argtypes = []
        L.umi_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().umi_last_error().decode(errors="replace")
        raise UmiError(f"{what or 'libumi_hip'} failed (status {status}): {msg}", status)


def call(name: str, *args) -> None:
    """Call entry point `name`; only entry points with declared argument
    types (without them ctypes would pass 64-bit pointers as 32-bit ints)."""
    if name not in SIGNATURES:
        raise UmiError(f"{name}: no argument types declared in umi_amd._lib.SIGNATURES")
    check(getattr(lib(), name)(*args), name)
