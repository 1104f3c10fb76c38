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


def test_gemm_and_conv_f32_loops():
    # the float restatements of main/gemm.cu:69-78 and main/fprop.cu:84-108 used as CPU baselines:
    # exact on small integers (equal to the int32 loops), close to FP64 on N(0,1), and the same
    # for any thread count and row range
    g = np.random.default_rng(11)
    A = g.integers(-3, 4, size=(40, 33)).astype(np.float32)
    B = g.integers(-3, 4, size=(33, 27)).astype(np.float32)
    np.testing.assert_array_equal(oracle.gemm_f32_rows(A, B, 0, 40), (A.astype(np.int64) @ B.astype(np.int64)))
    part = oracle.gemm_f32_rows(A, B, 5, 9)
    np.testing.assert_array_equal(part[5:9], (A[5:9].astype(np.int64) @ B.astype(np.int64)))
    assert not part[:5].any() and not part[9:].any()
    x = g.integers(0, 11, size=(2, 5, 9, 8))
    w = g.integers(0, 11, size=(4, 5, 3, 3))
    for stride, pad in ((1, 1), (2, 1), (1, 0)):
        np.testing.assert_array_equal(oracle.conv_f32(x, w, stride, pad).astype(np.int64),
                                      oracle.conv_direct_i32(x, w, stride, pad).astype(np.int64))
    xf = g.standard_normal((2, 6, 10, 7)).astype(np.float32)
    wf = g.standard_normal((3, 6, 3, 3)).astype(np.float32)
    one = oracle.conv_f32(xf, wf, 1, 1)
    assert np.abs(one - oracle.conv_f64(xf, wf, 1, 1)).max() < 1e-4
    try:
        oracle.set_threads(4)
        np.testing.assert_array_equal(oracle.conv_f32(xf, wf, 1, 1), one)
        np.testing.assert_array_equal(oracle.gemm_f32_rows(A, B, 0, 40), oracle.gemm_f32_rows(A, B, 0, 40))
    finally:
        oracle.set_threads(1)


def test_me_oracle_threads_do_not_change_results():
    from umi_amd import workloads

    cur, ref, _ = workloads.me_pair(96, 160, 8, seed=4)
    one, _ = oracle.me_full_search(cur, ref, 8, 8)
    try:
        oracle.set_threads(4)
        many, _ = oracle.me_full_search(cur, ref, 8, 8)
    finally:
        oracle.set_threads(1)
    np.testing.assert_array_equal(one, many)
