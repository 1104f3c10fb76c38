// SAD cost volume over a (2KY+1) x (2KX+1) window, uint8 pair (gfx950).
//
//   cost[y][x][i] = sum_{|dy|<=KY,|dx|<=KX} | L[cy][cl(x+dx)] - R[cy][cl(x+dx+s)] |,
//   s = s0 + step*i, cy = cl(y+dy).
//
// Generalises the MotionEst launch of main/cost_volume.cu:35-65 and its check
// loop main/cost_volume.cu:67-80 (johnjohnlin/UMI): one image there, 1x7
// window, 15 shifts, clamp on every final coordinate.
//
// Design: the volume (H*W*D int32) dwarfs the inputs, so the kernel is bound
// by the output write and everything else is arranged to stay under it.
//  * workgroup = 4 waves x 64 lanes; lane = shift index i (consecutive lanes
//    write consecutive int32 of one pixel -> 256 B per store instruction);
//    wave w owns a run of 8 output columns; the workgroup owns 32 columns x
//    TY rows.  L and R tiles (rows clamped, columns clamped, R widened by the
//    shift span of the workgroup's 64 lanes) are staged once in LDS;
//  * separable window: per tile row, the horizontal window sum of each of the
//    8 columns comes from v_sad_u8 over byte-aligned 4-byte chunks (chunks
//    built with v_alignbyte_b32, the last partial chunk masked); the vertical
//    sum is a running sum over a ring of 2KY+1 rows kept in registers.
#include "common.h"

namespace {
using namespace umi;

constexpr int CV_TX = 32;   // columns per workgroup (4 waves x 8)
constexpr int CV_TY = 32;   // output rows per workgroup
constexpr int CV_LANES = 64;

template <int KX, int KY>
__global__ __launch_bounds__(256) void cost_volume_kernel(const uint8_t *__restrict__ L,
                                                          const uint8_t *__restrict__ R, int H, int W,
                                                          int pitch, int D, int s0, int step, int ls4,
                                                          int rs4, int32_t *__restrict__ out) {
    constexpr int NT = 2 * KX + 1;          // horizontal taps
    constexpr int NC = (NT + 3) / 4;        // 4-byte chunks per window
    constexpr int NR = 2 * KY + 1;          // ring rows
    constexpr int TR = CV_TY + 2 * KY;      // tile rows
    extern __shared__ uint32_t lds[];
    uint32_t *lt = lds;                     // TR x ls4
    uint32_t *rt = lds + TR * ls4;          // TR x rs4

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * CV_TX, y0 = blockIdx.y * CV_TY;
    const int ibase = blockIdx.z * CV_LANES;
    const int nvalid = min(CV_LANES, D - ibase);
    // shift range of this workgroup's lanes
    const int sa = s0 + step * ibase, sb = s0 + step * (ibase + nvalid - 1);
    const int smin = min(sa, sb);

    for (int e = tid; e < TR * ls4; e += 256) {
        const int r = e / ls4, c4 = e - r * ls4;
        const uint8_t *row = L + (size_t)clampi(y0 - KY + r, 0, H - 1) * pitch;
        const int gx = x0 - KX + 4 * c4;
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) v |= (uint32_t)row[clampi(gx + b, 0, W - 1)] << (8 * b);
        lt[e] = v;
    }
    for (int e = tid; e < TR * rs4; e += 256) {
        const int r = e / rs4, c4 = e - r * rs4;
        const uint8_t *row = R + (size_t)clampi(y0 - KY + r, 0, H - 1) * pitch;
        const int gx = x0 - KX + smin + 4 * c4;
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) v |= (uint32_t)row[clampi(gx + b, 0, W - 1)] << (8 * b);
        rt[e] = v;
    }
    __syncthreads();

    const int i = ibase + lane;
    const bool lane_ok = lane < nvalid;
    const int s = s0 + step * (lane_ok ? i : ibase);
    const int xr = 8 * wv;                             // run start within the tile
    const int roff = xr + s - smin;                    // R tile byte offset of tap 0 of column xr
    const int rdw = roff >> 2, rsh = roff & 3;

    uint32_t ring[NR][8];
    uint32_t V[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        V[t] = 0;
#pragma unroll
        for (int q = 0; q < NR; ++q) ring[q][t] = 0;
    }
    const int ylim = min(CV_TY, H - y0);

    for (int yy0 = 0; yy0 < TR; yy0 += NR) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int yy = yy0 + q;
            if (yy < TR) {
                // 16 aligned bytes of L (uniform) and R (per lane) starting at tap 0.
                uint32_t lw[4], rw[5], ra[4];
                const uint32_t *lrow = lt + yy * ls4 + (xr >> 2);
                const uint32_t *rrow = rt + yy * rs4 + rdw;
#pragma unroll
                for (int k = 0; k < 4; ++k) lw[k] = lrow[k];
#pragma unroll
                for (int k = 0; k < 5; ++k) rw[k] = rrow[k];
#pragma unroll
                for (int k = 0; k < 4; ++k) ra[k] = __builtin_amdgcn_alignbyte(rw[k + 1], rw[k], rsh);
                // chunk(o) = bytes o..o+3 of the 16-byte window
                auto chunk = [](const uint32_t *w, int o) -> uint32_t {
                    return (o & 3) ? __builtin_amdgcn_alignbyte(w[(o >> 2) + 1], w[o >> 2], o & 3) : w[o >> 2];
                };
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    uint32_t h = 0;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const int nb = (NT - 4 * c) < 4 ? (NT - 4 * c) : 4;
                        const uint32_t mask = nb == 4 ? 0xffffffffu : ((1u << (8 * nb)) - 1u);
                        h = __builtin_amdgcn_sad_u8(chunk(lw, t + 4 * c) & mask, chunk(ra, t + 4 * c) & mask, h);
                    }
                    V[t] += h - ring[q][t];
                    ring[q][t] = h;
                }
                const int yo = yy - 2 * KY;   // output row within the tile
                if (yo >= 0 && yo < ylim && lane_ok) {
                    const int y = y0 + yo;
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const int x = x0 + xr + t;
                        if (x < W) out[((size_t)y * W + x) * D + i] = (int32_t)V[t];
                    }
                }
            }
        }
    }
}

using CvKernel = void (*)(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int, int32_t *);

template <int KX, int KY>
CvKernel cv_pick() {
    return cost_volume_kernel<KX, KY>;
}

CvKernel cv_table(int kx, int ky) {
    static const CvKernel table[4][4] = {
        {cv_pick<0, 0>(), cv_pick<0, 1>(), cv_pick<0, 2>(), cv_pick<0, 3>()},
        {cv_pick<1, 0>(), cv_pick<1, 1>(), cv_pick<1, 2>(), cv_pick<1, 3>()},
        {cv_pick<2, 0>(), cv_pick<2, 1>(), cv_pick<2, 2>(), cv_pick<2, 3>()},
        {cv_pick<3, 0>(), cv_pick<3, 1>(), cv_pick<3, 2>(), cv_pick<3, 3>()},
    };
    return table[kx][ky];
}

}  // namespace

extern "C" umi_status umi_cost_volume_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D,
                                         int s0, int step, int KY, int KX, int32_t *out, umi_stream s) {
    using umi::set_error;
    if (!L || !R || !out) return set_error(UMI_ERR_INVALID, "umi_cost_volume_u8: null pointer");
    if (H < 1 || W < 1 || pitch < W || D < 1)
        return set_error(UMI_ERR_INVALID, "umi_cost_volume_u8: H=%d W=%d pitch=%d D=%d", H, W, pitch, D);
    if (step != 1 && step != -1) return set_error(UMI_ERR_INVALID, "umi_cost_volume_u8: step %d", step);
    if (KX < 0 || KX > 3 || KY < 0 || KY > 3)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_cost_volume_u8: window radius KY=%d KX=%d (0..3)", KY, KX);
    const int ls4 = umi::ceil_div(CV_TX + 2 * KX, 4) + 2;
    const int rs4 = umi::ceil_div(CV_TX + 2 * KX + (CV_LANES - 1) + 8, 4) + 1;
    const size_t lds = 4 * (size_t)(CV_TY + 2 * KY) * (ls4 + rs4);
    dim3 grid(umi::ceil_div(W, CV_TX), umi::ceil_div(H, CV_TY), umi::ceil_div(D, CV_LANES));
    hipLaunchKernelGGL(cv_table(KX, KY), grid, dim3(256), lds, umi::as_stream(s), L, R, H, W, pitch, D, s0, step,
                       ls4, rs4, out);
    return umi::launch_check("cost_volume_kernel");
}
