// Quarter-pel refinement of block motion vectors (gfx950).
//
// No reference counterpart: johnjohnlin/UMI's motion search (main/full_me.cu)
// is integer-pel.  This refines the integer vectors of the full search the
// way a video encoder does before coding them, in quarter-pel units:
//
//   refq(qy, qx) = ((4-fx)(4-fy) A + fx (4-fy) B + (4-fx) fy C + fx fy D + 8) >> 4,
//   Y = qy >> 2, fy = qy & 3 (X, fx likewise), A = ref[cl(Y)][cl(X)],
//   B = ref[cl(Y)][cl(X+1)], C = ref[cl(Y+1)][cl(X)], D = ref[cl(Y+1)][cl(X+1)],
//
// clamp-to-edge as in the integer search.  SAD(v) of an 8x8 block at the
// quarter-pel vector v = (vy, vx) sums |cur - refq(4 y + vy, 4 x + vx)| over
// its pixels.  From c = 4 x (integer MV): the 8 half-pel neighbours c + 2 o,
// then the 8 quarter-pel neighbours of the best, o in {-1,0,1}^2 \ {0} in
// raster order, each replacing the centre only when strictly smaller.
//
// Work layout: 8 lanes per block, lane = block row.  A lane loads the 3 x 12
// clamped reference bytes around its row once (every candidate lies within
// one integer pel of the integer vector), then per candidate forms its 8
// interpolated pixels from registers (two v_dot4_u32_u8 each); the block's
// SAD is a DPP sum over its 8 lanes.
#include "common.h"

namespace {
using namespace umi;

constexpr int SP_THREADS = 256;  // 32 blocks x 8 lanes

// Sum over each aligned group of 8 lanes (quad swaps, then the half-row mirror).
__device__ inline uint32_t sum8(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    return v;
}

// SAD of the lane's row at quarter-pel offset (oy, ox) from the integer vector,
// oy, ox in [-3, 3].  win[r] = 12 reference bytes of row (y + dy - 1 + r) from
// column (x0 + dx - 1), cw = the lane's 8 current pixels.  A pixel's
// interpolation wa A + wb B (+ wc C + wd D) is a v_dot4_u32_u8 of the row dword
// that starts at A (B is its next byte) with the weights (wa, wb, 0, 0); the 8
// rounded values are packed into bytes and v_sad_u8 takes 4 differences at once.
__device__ inline uint32_t row_sad(const uint32_t (&win)[3][3], const uint32_t (&cw)[2], int oy, int ox) {
    const int r0 = (oy >> 2) + 1, c0 = (ox >> 2) + 1;  // 0 or 1
    const uint32_t fy = (uint32_t)(oy & 3), fx = (uint32_t)(ox & 3);
    const uint32_t wab = (4 - fx) * (4 - fy) | (fx * (4 - fy)) << 8;
    const uint32_t wcd = (4 - fx) * fy | (fx * fy) << 8;
    // a, c: bytes c0 .. c0 + 8 of the A row and the C row (rows r0 and r0 + 1)
    uint32_t a[3], c[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint32_t hi0 = j < 2 ? win[0][j + 1] : win[0][2], hi1 = j < 2 ? win[1][j + 1] : win[1][2],
                       hi2 = j < 2 ? win[2][j + 1] : win[2][2];
        const uint32_t a0 = __builtin_amdgcn_alignbyte(hi0, win[0][j], c0);
        const uint32_t m0 = __builtin_amdgcn_alignbyte(hi1, win[1][j], c0);
        const uint32_t c1 = __builtin_amdgcn_alignbyte(hi2, win[2][j], c0);
        a[j] = r0 ? m0 : a0;
        c[j] = r0 ? c1 : m0;
    }
    uint32_t s = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        uint32_t pk = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t ad = q ? __builtin_amdgcn_alignbyte(a[h + 1], a[h], q) : a[h];
            const uint32_t cd = q ? __builtin_amdgcn_alignbyte(c[h + 1], c[h], q) : c[h];
            const uint32_t v = __builtin_amdgcn_udot4(ad, wab, __builtin_amdgcn_udot4(cd, wcd, 8u, false), false) >> 4;
            pk |= v << (8 * q);
        }
        s = __builtin_amdgcn_sad_u8(pk, cw[h], s);
    }
    return s;
}

__global__ __launch_bounds__(SP_THREADS) void subpel_kernel(const uint8_t *__restrict__ cur,
                                                            const uint8_t *__restrict__ ref, int H, int W, int pitch,
                                                            const int32_t *__restrict__ mv_in, int nblocks, int nbx,
                                                            int32_t *__restrict__ mv_out) {
    const int blk = blockIdx.x * (SP_THREADS / 8) + (int)(threadIdx.x >> 3);
    const int py = threadIdx.x & 7;
    const bool valid = blk < nblocks;
    const int b = valid ? blk : nblocks - 1;  // lanes past the end work on the last block, write nothing
    const int by = b / nbx, bx = b - by * nbx;
    const int dy = mv_in[3 * b], dx = mv_in[3 * b + 1];
    const int y = 8 * by + py, x0 = 8 * bx;
    // Bytes by aligned dword loads when a row span lies inside its row (the
    // dwords around it hold at least one byte of the span, so they stay in
    // the row's mapped memory), then v_alignbyte; a window reaching past a
    // frame edge takes clamped byte loads.
    const auto load_span = [](const uint8_t *row, int x, int nd, uint32_t *dst) {
        const uintptr_t p = reinterpret_cast<uintptr_t>(row + x), a0 = p & ~(uintptr_t)3;
        const uint32_t sh = (uint32_t)(p & 3);
        const uintptr_t last = (p + 4 * nd - 1) & ~(uintptr_t)3;  // dword holding the span's last byte
        uint32_t d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = *reinterpret_cast<const uint32_t *>(min(a0 + 4 * k, last));
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (k < nd) dst[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh);
    };
    uint32_t cw[2];
    {
        uint32_t t[3];
        load_span(cur + (size_t)y * pitch, x0, 2, t);
        cw[0] = t[0];
        cw[1] = t[1];
    }
    uint32_t win[3][3];
    const int xs = x0 + dx - 1;
    const bool inside = xs >= 0 && xs + 12 <= W;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const uint8_t *row = ref + (size_t)clampi(y + dy - 1 + r, 0, H - 1) * pitch;
        if (inside) {
            load_span(row, xs, 3, win[r]);
        } else {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) v |= (uint32_t)row[clampi(xs + 4 * j + k, 0, W - 1)] << (8 * k);
                win[r][j] = v;
            }
        }
    }
    uint32_t best = sum8(row_sad(win, cw, 0, 0));
    int cy = 0, cx = 0;
#pragma unroll
    for (int step = 2; step >= 1; step >>= 1) {
        const int oy0 = cy, ox0 = cx;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == 4) continue;
            const int oy = oy0 + step * (k / 3 - 1), ox = ox0 + step * (k % 3 - 1);
            const uint32_t s = sum8(row_sad(win, cw, oy, ox));
            if (s < best) {
                best = s;
                cy = oy;
                cx = ox;
            }
        }
    }
    if (valid && py == 0) {
        mv_out[3 * b] = 4 * dy + cy;
        mv_out[3 * b + 1] = 4 * dx + cx;
        mv_out[3 * b + 2] = (int32_t)best;
    }
}

}  // namespace

extern "C" umi_status umi_me_subpel_refine_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                              const int32_t *mv_in, int32_t *mv_out, umi_stream s) {
    if (!cur || !ref || !mv_in || !mv_out) return set_error(UMI_ERR_INVALID, "umi_me_subpel_refine_u8: null pointer");
    if (H < 8 || W < 8 || pitch < W)
        return set_error(UMI_ERR_INVALID, "umi_me_subpel_refine_u8: H=%d W=%d pitch=%d", H, W, pitch);
    const int nbx = W / 8, nb = (H / 8) * nbx;
    hipLaunchKernelGGL(subpel_kernel, dim3(ceil_div(nb, SP_THREADS / 8)), dim3(SP_THREADS), 0, as_stream(s), cur, ref,
                       H, W, pitch, mv_in, nb, nbx, mv_out);
    return launch_check("subpel_kernel");
}
