"""GPU parity of full-search motion estimation against the oracle (bit-exact)."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from umi_amd import ops, workloads
from umi_amd._lib import call

GOLD = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu


# bit 0: v_qsad_pk_u16_u8 (1) or v_sad_u8 (0); bit 1: four-column groups only,
# without the light lane that takes the last column when S = 4G + 1; bit 2:
# dword staging even where frames allow 16-byte staging; bit 3: runtime skips
# in the last row chunk even where 8x8 blocks with R % 4 == 0 make them static
@pytest.fixture(params=[1, 0, 3, 2, 5, 9], ids=["qsad", "sad", "qsad_groups", "sad_groups", "qsad_dword",
                                                "qsad_runtime"])
def variant(request, gpu):
    call("umi_me_set_variant", request.param)
    yield request.param
    call("umi_me_set_variant", 1)


@pytest.mark.parametrize("name", ["me_ramp", "me_shift", "me_random8", "me_random16"])
def test_me_golden(variant, name):
    f = np.load(GOLD / f"{name}.npz")
    print(f"[me] golden {name} variant={variant}", flush=True)
    mv, vol = ops.me_full_search_np(f["cur"], f["ref"], int(f["B"]), int(f["R"]), volume=True)
    np.testing.assert_array_equal(vol, f["vol"])
    np.testing.assert_array_equal(mv, f["mv"])


@pytest.mark.parametrize("H,W,B,R", [(64, 96, 8, 16), (72, 80, 16, 7), (100, 130, 8, 32), (37, 53, 8, 1),
                                     (96, 160, 16, 16), (80, 80, 8, 64), (72, 112, 8, 8)])
def test_me_random_vs_oracle(variant, H, W, B, R):
    g = workloads.rng(H * 1000 + W + R)
    cur = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    ref = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    print(f"[me] random H={H} W={W} B={B} R={R} variant={variant}", flush=True)
    mv, vol = ops.me_full_search_np(cur, ref, B, R, volume=True)
    omv, ovol = oracle.me_full_search(cur, ref, B, R, volume=True)
    np.testing.assert_array_equal(vol, ovol)
    np.testing.assert_array_equal(mv, omv)


@pytest.mark.parametrize("H,W,B,R", [(64, 96, 8, 16), (100, 130, 8, 32), (37, 53, 8, 1), (80, 80, 8, 64),
                                     (72, 88, 8, 7), (48, 64, 8, 2), (72, 80, 16, 7), (72, 112, 8, 8),
                                     (56, 144, 8, 12), (64, 208, 8, 4), (48, 176, 16, 12)])
def test_me_mv_only_vs_oracle(variant, H, W, B, R):
    """Motion vectors without the volume (the kernel form the bench and the
    proposer's judge run), on frames with many equal SADs."""
    g = workloads.rng(H * 31 + W + R)
    cur = (g.integers(0, 4, size=(H, W)) * 60).astype(np.uint8)   # 4 levels: frequent ties
    ref = (g.integers(0, 4, size=(H, W)) * 60).astype(np.uint8)
    print(f"[me] mv-only H={H} W={W} B={B} R={R} variant={variant}", flush=True)
    mv, _ = ops.me_full_search_np(cur, ref, B, R)
    omv, _ = oracle.me_full_search(cur, ref, B, R)
    np.testing.assert_array_equal(mv, omv)


def test_me_ties_first_in_raster_order(gpu):
    f = np.full((64, 64), 9, np.uint8)
    mv, _ = ops.me_full_search_np(f, f, 8, 5)
    assert (mv[..., 0] == -5).all() and (mv[..., 1] == -5).all() and (mv[..., 2] == 0).all()


def test_me_full_hd_r16_vs_oracle(gpu):
    # 1920x1080 pair, 8x8 blocks, +-16: every block bit-exact against the oracle
    cur, ref, true_mv = workloads.me_pair(1080, 1920, 16, seed=1)
    print("[me] 1080p R=16 full frame", flush=True)
    mv, _ = ops.me_full_search_np(cur, ref, 8, 16)
    omv, _ = oracle.me_full_search(cur, ref, 8, 16)
    np.testing.assert_array_equal(mv, omv)
    # most interior blocks recover the planted motion
    t = true_mv[np.arange(135)[:, None] // 2, np.arange(240)[None, :] // 2]
    hit = (mv[..., :2] == t).all(axis=2)[2:-2, 2:-2].mean()
    assert hit > 0.9, hit


def test_me_full_hd_r32_rows_vs_oracle(gpu):
    # the +-32 judge search: every 4th block row (34 of 135, both frame edges included) bit-exact
    cur, ref, _ = workloads.me_pair(1080, 1920, 32, seed=5)
    print("[me] 1080p R=32", flush=True)
    mv, vol = ops.me_full_search_np(cur, ref, 8, 32)
    rows = sorted(set(range(0, 135, 4)) | {134})
    for r in rows:
        omv, _ = oracle.me_full_search(cur, ref, 8, 32, by_range=(r, r + 1))
        np.testing.assert_array_equal(mv[r], omv[r], err_msg=f"block row {r}")


def test_me_volume_argmin_consistent(gpu):
    # size-independent property: the MV output is the first minimum of the volume output
    cur, ref, _ = workloads.me_pair(264, 480, 16, seed=3)
    mv, vol = ops.me_full_search_np(cur, ref, 8, 16, volume=True)
    S = 33
    flat = vol.reshape(vol.shape[0], vol.shape[1], S * S)
    idx = flat.argmin(axis=2)
    np.testing.assert_array_equal(mv[..., 0], idx // S - 16)
    np.testing.assert_array_equal(mv[..., 1], idx % S - 16)
    np.testing.assert_array_equal(mv[..., 2], flat.min(axis=2))


def test_me_pitch_and_offsets(gpu):
    # frames inside a wider allocation (pitch > W)
    from umi_amd import DeviceBuffer

    H, W, P = 48, 56, 80
    g = workloads.rng(11)
    big_c = g.integers(0, 256, size=(H, P), dtype=np.int64).astype(np.uint8)
    big_r = g.integers(0, 256, size=(H, P), dtype=np.int64).astype(np.uint8)
    dc, dr = DeviceBuffer.from_array(big_c), DeviceBuffer.from_array(big_r)
    mv = DeviceBuffer(6 * 7 * 3 * 4)
    ops.me_full_search(dc, dr, H, W, P, 8, 4, mv)
    got = mv.download((6, 7, 3), np.int32)
    omv, _ = oracle.me_full_search(np.ascontiguousarray(big_c[:, :W]), np.ascontiguousarray(big_r[:, :W]), 8, 4)
    np.testing.assert_array_equal(got, omv)


@pytest.mark.parametrize("H,W,B,R,nlist", [(64, 96, 8, 16, 20), (72, 80, 16, 7, 5), (100, 130, 8, 32, 40),
                                           (37, 53, 8, 1, 3), (96, 160, 16, 16, 17), (80, 88, 8, 64, 9),
                                           (64, 64, 8, 8, 0)])
def test_me_listed_vs_oracle(gpu, H, W, B, R, nlist):
    """umi_me_full_search_listed_u8 writes the oracle's (dy, dx, SAD) for the
    listed blocks and leaves every other entry untouched."""
    from umi_amd import DeviceBuffer
    g = workloads.rng(H * 7 + W + R)
    cur = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    ref = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    nby, nbx = H // B, W // B
    n = nby * nbx
    lst = g.permutation(n)[:nlist].astype(np.int32)
    print(f"[me] listed H={H} W={W} B={B} R={R} n={nlist}", flush=True)
    dc, dr = DeviceBuffer.from_array(cur), DeviceBuffer.from_array(ref)
    dl, dn = DeviceBuffer.from_array(np.concatenate([lst, [0]]).astype(np.int32)), DeviceBuffer.from_array(
        np.array([nlist], np.int32))
    init = np.full((n, 3), -7, np.int32)
    mv = DeviceBuffer.from_array(init)
    call("umi_me_full_search_listed_u8", dc.ptr, dr.ptr, H, W, W, B, R, dl.ptr, dn.ptr, n, mv.ptr, None)
    got = mv.download((n, 3), np.int32)
    want, _ = oracle.me_full_search(cur, ref, B, R)
    want = want.reshape(n, 3)
    listed = np.zeros(n, bool)
    listed[lst] = True
    np.testing.assert_array_equal(got[listed], want[listed])
    np.testing.assert_array_equal(got[~listed], init[~listed])


# ---------------------------------------------------------------------------
# Quarter-pel refinement of the integer vectors (umi_me_subpel_refine_u8)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("H,W,R", [(64, 96, 8), (120, 200, 16)])
def test_subpel_after_full_search(gpu, H, W, R):
    g = workloads.rng(H + W)
    ref = workloads.smooth_noise_u8(H, W, g)
    cur = np.roll(ref, (3, -5), axis=(0, 1))
    cur = np.clip(cur.astype(int) + g.normal(0, 2, cur.shape).round().astype(int), 0, 255).astype(np.uint8)
    print(f"[subpel] {H}x{W} R={R}", flush=True)
    mv, _ = ops.me_full_search_np(cur, ref, 8, R)
    got = ops.me_subpel_refine_np(cur, ref, mv)
    np.testing.assert_array_equal(got, oracle.me_subpel(cur, ref, mv))
    assert (got[..., 2] <= mv[..., 2]).all()


def test_subpel_arbitrary_vectors_and_edges(gpu):
    # vectors pointing past every frame edge: the clamp inside the interpolation
    g = workloads.rng(31)
    H, W = 56, 72
    ref = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    cur = g.integers(0, 256, size=(H, W), dtype=np.int64).astype(np.uint8)
    mv = np.zeros((H // 8, W // 8, 3), np.int32)
    mv[..., :2] = g.integers(-20, 21, size=(H // 8, W // 8, 2))
    print("[subpel] arbitrary vectors", flush=True)
    np.testing.assert_array_equal(ops.me_subpel_refine_np(cur, ref, mv), oracle.me_subpel(cur, ref, mv))


def test_subpel_quarter_shift_known_answer(gpu):
    from test_oracle_subpel import shifted  # the test directory is on sys.path

    ref = workloads.smooth_noise_u8(64, 96, workloads.rng(6))
    cur = shifted(ref, 1, -3)
    mv = np.zeros((8, 12, 3), np.int32)
    mv[..., 1] = -1
    print("[subpel] known answer", flush=True)
    out = ops.me_subpel_refine_np(cur, ref, mv)
    assert (out[..., 0] == 1).all() and (out[..., 1] == -3).all() and (out[..., 2] == 0).all()


def test_subpel_1080p_and_nv12_luma(gpu):
    # full 1080p pair; the same search on the luma plane of NV12 buffers (Y
    # plane, then the interleaved UV plane) through their pitch gives the same
    from umi_amd import DeviceBuffer

    H, W = 1080, 1920
    cur, ref, _ = workloads.me_pair(H, W, seed=3)
    print("[subpel] 1080p", flush=True)
    mv, _ = ops.me_full_search_np(cur, ref, 8, 16)
    got = ops.me_subpel_refine_np(cur, ref, mv)
    np.testing.assert_array_equal(got, oracle.me_subpel(cur, ref, mv))
    P = W + 64
    nv = []
    for f in (cur, ref):
        buf = np.zeros((H + H // 2, P), np.uint8)
        buf[:H, :W] = f
        buf[H:, :W] = 128
        nv.append(DeviceBuffer.from_array(buf))
    mv_d, out_d = DeviceBuffer(mv.size * 4), DeviceBuffer(mv.size * 4)
    print("[subpel] NV12 luma planes", flush=True)
    ops.me_full_search(nv[0], nv[1], H, W, P, 8, 16, mv_d)
    ops.me_subpel_refine(nv[0], nv[1], H, W, P, mv_d, out_d)
    np.testing.assert_array_equal(mv_d.download(mv.shape, np.int32), mv)
    np.testing.assert_array_equal(out_d.download(mv.shape, np.int32), got)
