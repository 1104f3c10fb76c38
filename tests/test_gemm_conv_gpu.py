"""MFMA GEMM and implicit-GEMM convolution on the GPU.

Kernel correctness: the output equals the FP64 contraction of the operands
the kernel actually received (decoded from the packed bf16 / e4m3 bytes; for
the split modes the three products hi*hi + lo*hi + hi*lo, with the fp8 lo
weight 2^-4), to accumulation rounding: 2e-6 of ||a|| ||b|| for the fp32
and bf16 instructions, 5e-5 for the fp8 instruction, whose 64-term dot
product is not summed in IEEE fp32 inside the matrix core (measured max
2.2e-5, mean 1e-6).  Precision: against the FP64 oracle of the unrounded inputs,
each mode meets its stated bound; the measured max / mean are printed.
"""
import numpy as np
import pytest

import oracle
from umi_amd import ops, workloads

pytestmark = pytest.mark.gpu
MODES = [ops.F32, ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3]
# bound on max |C - C64| / (||a_m|| ||b_n||) against the FP64 oracle of the fp32 inputs
PRECISION = {ops.F32: 1e-6, ops.BF16X3: 2e-6, ops.BF16: 1e-3, ops.FP8: 4e-2, ops.FP8X3: 4e-3}


OWN_TOL = {ops.F32: 2e-6, ops.BF16: 2e-6, ops.BF16X3: 2e-6, ops.FP8: 5e-5, ops.FP8X3: 5e-5}


def parts(op, mode):
    """(hi, lo or None) as float64, already divided by q and with the lo weight applied."""
    hi, lo, q = op
    if mode == ops.F32:
        return hi.astype(np.float64), None
    if mode in (ops.BF16, ops.BF16X3):
        return oracle.decode_bf16(hi), (oracle.decode_bf16(lo) if lo is not None else None)
    h = oracle.decode_e4m3(hi) / q
    return h, (oracle.decode_e4m3(lo) / (16.0 * q) if lo is not None else None)


def dequant(op, mode):
    h, l = parts(op, mode)
    return h if l is None else h + l


def three_products(contract, a, b, mode):
    """hi*hi (+ lo*hi + hi*lo): exactly the products the kernel forms."""
    (ah, al), (bh, bl) = parts(a, mode), parts(b, mode)
    out = contract(ah, bh)
    if al is not None:
        out = out + contract(al, bh) + contract(ah, bl)
    return out


def normed_err(C, ref, A, Bt):
    na = np.linalg.norm(A, axis=1)
    nb = np.linalg.norm(Bt, axis=1)
    return oracle.contraction_error(C, ref, na, nb)


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
@pytest.mark.parametrize("M,N,K", [(256, 256, 64), (300, 200, 128), (64, 520, 256), (512, 512, 1024)])
def test_gemm_kernel_exact_on_its_operands(mode, M, N, K, gpu):
    g = workloads.rng(M + N + K + mode)
    A = g.standard_normal((M, K)).astype(np.float32)
    B = g.standard_normal((K, N)).astype(np.float32)
    print(f"[gemm] {ops.MODE_NAMES[mode]} {M}x{N}x{K}", flush=True)
    C, oa, ob = ops.gemm_np(A, B, mode, return_operands=True)
    ref = three_products(lambda a, b: a @ b.T, oa, ob, mode)
    e = normed_err(C, ref, dequant(oa, mode), dequant(ob, mode))
    print(f"[gemm]   vs own operands {e}", flush=True)
    assert e["max_rel"] <= OWN_TOL[mode]


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
def test_gemm_precision_vs_fp64(mode, gpu):
    # N(0,1) generated in FP64 and cast down; 1024^3 keeps the FP64 oracle quick
    g = workloads.rng(3)
    A64 = g.standard_normal((1024, 1024))
    B64 = g.standard_normal((1024, 1024))
    C = ops.gemm_np(A64.astype(np.float32), B64.astype(np.float32), mode)
    e = normed_err(C, A64 @ B64, A64, B64.T)
    print(f"[gemm] precision {ops.MODE_NAMES[mode]} 1024^3: {e}", flush=True)
    assert e["max_rel"] <= PRECISION[mode]


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
def test_gemm_4096_precision_vs_fp64(mode, gpu):
    # the precision config: M = N = K = 4096, N(0,1) in FP64 cast to fp32; every 64th row of C
    # (64 rows, all 4096 columns) against the FP64 oracle, with the same bounds as 1024^3
    g = workloads.rng(4096)
    A64 = g.standard_normal((4096, 4096))
    B64 = g.standard_normal((4096, 4096))
    print(f"[gemm] precision {ops.MODE_NAMES[mode]} 4096^3", flush=True)
    C = ops.gemm_np(A64.astype(np.float32), B64.astype(np.float32), mode)
    rows = np.arange(0, 4096, 64)
    e = normed_err(C[rows], oracle.gemm_f64(A64[rows], B64), A64[rows], B64.T)
    print(f"[gemm] precision {ops.MODE_NAMES[mode]} 4096^3 (64 rows): {e}", flush=True)
    assert e["max_rel"] <= PRECISION[mode]
    assert np.isfinite(C).all()


def test_gemm_identity_and_int_exact(gpu):
    # small integers are exact in every mode's operand format (fp8: {-2..2} with q = 224)
    g = workloads.rng(9)
    A = g.integers(-2, 3, size=(256, 128)).astype(np.float32)
    A[0, 0] = 2.0
    B = g.integers(-2, 3, size=(128, 256)).astype(np.float32)
    B[0, 0] = -2.0
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for mode in MODES:
        C = ops.gemm_np(A, B, mode)
        np.testing.assert_allclose(C, exact, rtol=1e-6, atol=1e-4, err_msg=ops.MODE_NAMES[mode])
    I = np.eye(256, dtype=np.float32)
    np.testing.assert_array_equal(ops.gemm_np(I, I, ops.BF16), I)


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
@pytest.mark.parametrize("shape", [(2, 64, 12, 14, 32, 3, 1, 1), (1, 32, 9, 7, 64, 3, 2, 1), (3, 64, 10, 10, 16, 1, 1, 0),
                                   (2, 128, 6, 8, 40, 3, 1, 2)])
def test_conv_kernel_exact_on_its_operands(mode, shape, gpu):
    N, C, H, W, K, R, stride, pad = shape
    if (R * R * C * ops.elem_bytes(mode)) % 64:
        pytest.skip("R*S*C not a whole number of 64-byte stages for this mode")
    g = workloads.rng(sum(shape) + mode)
    x = g.standard_normal((N, C, H, W)).astype(np.float32)
    w = g.standard_normal((K, C, R, R)).astype(np.float32)
    print(f"[conv] {ops.MODE_NAMES[mode]} {shape}", flush=True)
    out = ops.conv_fprop_np(x, w, stride, pad, mode)
    ref = oracle.conv_f64(x, w, stride, pad)
    # operands as rounded by the mode, via the same packers on the GPU
    from umi_amd import DeviceBuffer
    ox = ops.conv_pack_input(DeviceBuffer.from_array(x), N, C, H, W, mode)
    ow = ops.conv_pack_filter(DeviceBuffer.from_array(w), K, C, R, R, mode)
    to_x = lambda v: v.reshape(N, H, W, C).transpose(0, 3, 1, 2)
    to_w = lambda v: v.reshape(K, R, R, C).transpose(0, 3, 1, 2)
    refq = three_products(lambda a, b: oracle.conv_f64(to_x(b), to_w(a), stride, pad),
                          ops.download_operand(ow), ops.download_operand(ox), mode)
    xq, wq = to_x(dequant(ops.download_operand(ox), mode)), to_w(dequant(ops.download_operand(ow), mode))
    scale = np.sqrt(C * R * R) * np.abs(xq).max() * np.abs(wq).max()
    assert np.abs(out - refq).max() <= OWN_TOL[mode] * scale
    assert np.abs(out - ref).max() <= max(PRECISION[mode] * 40, 1e-5) * scale


def test_conv_integer_matches_reference_loop(gpu):
    # main/fprop.cu's data: integers 0..10; the fp32 MFMA path is exact on them
    g = workloads.rng(4)
    x = g.integers(0, 11, size=(2, 16, 12, 10))
    w = g.integers(0, 11, size=(8, 16, 3, 3))
    out = ops.conv_fprop_np(x.astype(np.float32), w.astype(np.float32), 1, 1, ops.F32)
    np.testing.assert_array_equal(out.astype(np.int64), oracle.conv_direct_i32(x, w, 1, 1).astype(np.int64))


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
def test_conv_config_precision(mode, gpu):
    # N=16, C=64, 56x56, K=64, 3x3, stride 1, pad 1; N(0,1) in FP64 cast to fp32
    g = workloads.rng(16)
    x64 = g.standard_normal((16, 64, 56, 56))
    w64 = g.standard_normal((64, 64, 3, 3))
    out = ops.conv_fprop_np(x64.astype(np.float32), w64.astype(np.float32), 1, 1, mode)
    cols, _ = oracle.im2col_f64(x64, 3, 3, 1, 1)
    ref = oracle.conv_f64(x64, w64, 1, 1)
    na = np.linalg.norm(w64.reshape(64, -1), axis=1)            # per out channel
    nb = np.linalg.norm(cols, axis=1).reshape(16, 56, 56)        # per output pixel
    rel = np.abs(out - ref) / (na[None, :, None, None] * nb[:, None, :, :])
    print(f"[conv] precision {ops.MODE_NAMES[mode]}: max {rel.max():.3e} mean {rel.mean():.3e}", flush=True)
    assert rel.max() <= PRECISION[mode]


@pytest.mark.parametrize("shape", [(16, 56, 56), (2, 56, 56), (3, 7, 9), (1, 64, 64), (2, 13, 8), (1, 3, 63), (2, 40, 33),
                                   (2, 13, 1), (1, 1, 5)])
@pytest.mark.parametrize("mode", [ops.BF16, ops.BF16X3, ops.FP8, ops.FP8X3], ids=["bf16", "bf16x3", "fp8", "fp8x3"])
def test_conv_direct_equals_implicit_gemm(shape, mode, gpu):
    # the direct 3x3 form (C = K = 64, stride 1, pad 1, bf16 / bf16x3 / fp8 / fp8x3) uses
    # the implicit GEMM's operands, MFMA order and scaling: identical bits, and both
    # within the mode's bound of the FP64 oracle (bf16x3 takes the direct form for
    # W <= 57, where both input parts and two filter taps fit the LDS)
    from umi_amd._lib import call

    N, H, W = shape
    g = workloads.rng(N * 131 + H * 7 + W)
    x = g.standard_normal((N, 64, H, W)).astype(np.float32)
    w = g.standard_normal((64, 64, 3, 3)).astype(np.float32)
    print(f"[conv] direct vs implicit GEMM {ops.MODE_NAMES[mode]} {shape}", flush=True)
    try:
        call("umi_conv_set_direct", 0)
        ref = ops.conv_fprop_np(x, w, 1, 1, mode)
    finally:
        call("umi_conv_set_direct", 1)
    out = ops.conv_fprop_np(x, w, 1, 1, mode)
    np.testing.assert_array_equal(out, ref)
    f64 = oracle.conv_f64(x, w, 1, 1)
    scale = np.sqrt(64 * 9) * np.abs(x).max() * np.abs(w).max()
    assert np.abs(out - f64).max() <= PRECISION[mode] * 40 * scale


@pytest.mark.parametrize("mode", MODES, ids=[ops.MODE_NAMES[m] for m in MODES])
@pytest.mark.parametrize("M,N,K", [(4096, 4096, 256), (4000, 4100, 320)])
def test_gemm_direct_lds_equals_register_staging(mode, M, N, K, gpu):
    # 256 x 256 tiles staged by direct-to-LDS loads use the register-staged kernel's
    # operands, LDS layout and MFMA order: identical bits (ragged M, N included)
    from umi_amd._lib import call

    g = workloads.rng(M + N + K + mode)
    A = g.standard_normal((M, K)).astype(np.float32)
    B = g.standard_normal((K, N)).astype(np.float32)
    print(f"[gemm] direct-LDS vs register staging {ops.MODE_NAMES[mode]} {M}x{N}x{K}", flush=True)
    try:
        call("umi_gemm_set_lds_direct", 0)
        ref = ops.gemm_np(A, B, mode)
        # the LDS ring with and without the register pipeline, 8 and 4 waves
        for form in (2, 4, 5):
            call("umi_gemm_set_lds_direct", form)
            np.testing.assert_array_equal(ops.gemm_np(A, B, mode), ref, err_msg=f"form {form}")
    finally:
        call("umi_gemm_set_lds_direct", 1)
    np.testing.assert_array_equal(ops.gemm_np(A, B, mode), ref)
