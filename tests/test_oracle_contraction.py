"""Pins the FP64 contraction oracles and the operand decoders (CPU only).

* conv_f64 equals the reference's direct int32 loop (main/fprop.cu:84-108,
  restated as conv_direct_i32) on small integer tensors, for stride 1 and 2;
* gemm_f64 equals the wrapping int32 loop of main/gemm.cu:69-78 on small
  iota operands (no wrap at that size) and closed forms (identity, rank 1);
* e4m3 / bf16 decoders on known codes (1.0, 448, the smallest subnormal, NaN).
"""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("stride,pad,R", [(1, 1, 3), (2, 3, 7), (1, 0, 1)])
def test_conv_f64_matches_direct_loop(stride, pad, R):
    g = np.random.Generator(np.random.PCG64(R + stride))
    x = g.integers(0, 11, size=(2, 3, 9, 11))
    w = g.integers(0, 11, size=(4, 3, R, R))
    ref = oracle.conv_direct_i32(x, w, stride, pad)
    np.testing.assert_array_equal(oracle.conv_f64(x, w, stride, pad), ref.astype(np.float64))


def test_gemm_f64_closed_forms():
    m, k, n = 10, 12, 20
    A = np.arange(m * k).reshape(m, k)
    B = np.arange(k * n).reshape(k, n)
    loop = np.array([[sum(A[y, kk] * B[kk, x] for kk in range(k)) for x in range(n)] for y in range(m)])
    np.testing.assert_array_equal(oracle.gemm_f64(A, B), loop)
    np.testing.assert_array_equal(oracle.gemm_f64(np.eye(5), B[:5]), B[:5])
    u, v = np.arange(1, 7.0), np.arange(1, 9.0)
    np.testing.assert_array_equal(oracle.gemm_f64(u[:, None], v[None, :]), np.outer(u, v))


def test_decoders():
    assert oracle.decode_e4m3(np.array([0x38]))[0] == 1.0
    assert oracle.decode_e4m3(np.array([0x7E]))[0] == 448.0
    assert oracle.decode_e4m3(np.array([0xFE]))[0] == -448.0
    assert oracle.decode_e4m3(np.array([0x01]))[0] == 2.0 ** -9
    assert np.isnan(oracle.decode_e4m3(np.array([0x7F]))[0])
    assert oracle.decode_bf16(np.array([0x3F80], np.uint16))[0] == 1.0
    assert oracle.decode_bf16(np.array([0xC040], np.uint16))[0] == -3.0
    # every finite e4m3 value appears once per sign; 448 is the largest
    fin = oracle.E4M3[np.isfinite(oracle.E4M3)]
    assert fin.max() == 448.0 and len(np.unique(np.abs(fin))) == 127
