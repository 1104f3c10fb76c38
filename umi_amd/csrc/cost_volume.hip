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
#include "stage.h"
#include "wave.h"

namespace {
using namespace umi;

constexpr int CV_TX = 32;   // columns per workgroup (4 waves x 8)
constexpr int CV_TY = 32;   // output rows per workgroup
constexpr int CV_LANES = 64;
constexpr int WTA_TY = 16;  // output rows per workgroup of the winner-take-all kernel

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

// ---------------------------------------------------------------------------
// Winner-take-all over the cost volume without storing it: per pixel the
// first minimum over i of cost[y][x][i] (same costs as cost_volume_kernel),
// as the key (cost << ib) | i.  The volume never reaches device memory.
//  * tile and LDS staging as cost_volume_kernel, on 32 columns x WTA_TY rows
//    (wave w owns columns 8 w .. 8 w + 7), except that the R tile is
//    widened by the span of all D shifts and starts at a dword boundary.
//    Lane = shift: a pass holds NCH chunks of 64 shifts in registers and
//    merges their keys in-lane before the cross-lane step; lanes past D
//    repeat the chunk's first shift (a chunk past D repeats shift D - 1),
//    whose key they then duplicate;
//  * the L window is the same for all lanes: its words are made wave-uniform
//    so that its byte alignments run on the scalar unit;
//  * the vertical window is a running column sum C (accumulated straight
//    into the v_sad_u8 chains) minus its value 2KY+1 rows earlier (a ring of
//    2KY+1 rows in registers); differences are exact mod 2^32.  The tile's
//    rows are unrolled (static ring slots and output rows);
//  * per output row a reduce-scatter over the 64 lanes (lane swaps at
//    distances 32, 16, 8 halve the columns per lane, then the minimum over
//    groups of 8 lanes) leaves the minimum of column (lane >> 3) & 7 in
//    every lane, with no LDS traffic; one lane per column merges it into
//    an LDS tile of keys by an LDS atomic minimum;
//  * the tile is written as int32 pairs (i, cost).
template <int KX, int KY, int NCH>
__global__ __launch_bounds__(256) void stereo_wta_kernel(const uint8_t *__restrict__ L,
                                                         const uint8_t *__restrict__ R, int H, int W, int pitch,
                                                         int D, int s0, int step, int ls4, int rs4, int ib,
                                                         int32_t *__restrict__ out) {
    constexpr int NT = 2 * KX + 1;
    constexpr int NC = (NT + 3) / 4;
    constexpr int NR = 2 * KY + 1;
    constexpr int TR = WTA_TY + 2 * KY;
    constexpr int NTHR = 256;
    extern __shared__ uint32_t lds[];
    uint32_t *lt = lds;                     // TR x ls4
    uint32_t *rt = lt + TR * ls4;           // TR x rs4
    uint32_t *best = rt + TR * rs4;         // WTA_TY x CV_TX keys

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * CV_TX, y0 = blockIdx.y * WTA_TY;
    const int smin = min(s0, s0 + step * (D - 1));
    // the R tile starts at the dword boundary at or below x0 - KX + smin
    const int rx0 = (x0 - KX + smin) & ~3, rskew = x0 - KX + smin - rx0;
    const auto lrow = [&](int r) { return L + (size_t)clampi(y0 - KY + r, 0, H - 1) * pitch; };
    const auto rrow = [&](int r) { return R + (size_t)clampi(y0 - KY + r, 0, H - 1) * pitch; };
    if (W >= 8) {
        stage_clamped<2>(lt, TR * ls4, tid, NTHR, W, [&](int e, const uint8_t *&row, int &gx) {
            const int r = e / ls4;
            row = lrow(r);
            gx = x0 - KX + 4 * (e - r * ls4);
        });
        stage_clamped<4>(rt, TR * rs4, tid, NTHR, W, [&](int e, const uint8_t *&row, int &gx) {
            const int r = e / rs4;
            row = rrow(r);
            gx = rx0 + 4 * (e - r * rs4);
        });
    } else {
        for (int e = tid; e < TR * ls4; e += NTHR) {
            const int r = e / ls4;
            lt[e] = load4_clamped(lrow(r), x0 - KX + 4 * (e - r * ls4), W);
        }
        for (int e = tid; e < TR * rs4; e += NTHR) {
            const int r = e / rs4;
            rt[e] = load4_clamped(rrow(r), rx0 + 4 * (e - r * rs4), W);
        }
    }
    for (int e = tid; e < WTA_TY * CV_TX; e += NTHR) best[e] = 0xffffffffu;
    __syncthreads();

    const int xr = 8 * wv;
    const int ylim = min(WTA_TY, H - y0);
    const int tcol = (lane >> 3) & 7;  // column of this lane after the reduce-scatter
    const bool hi8 = lane & 8;

    for (int pbase = 0; pbase < D; pbase += 64 * NCH) {
        int idx[NCH], rdw[NCH], rsh[NCH];
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
            const int cb = min(pbase + 64 * k, D - 1);  // a chunk past D repeats shift D - 1
            idx[k] = cb + lane < D ? cb + lane : cb;
            const int roff = xr + s0 + step * idx[k] - smin + rskew;
            rdw[k] = roff >> 2;
            rsh[k] = roff & 3;
        }
        uint32_t C[NCH][8], ring[NR][NCH][8];
#pragma unroll
        for (int k = 0; k < NCH; ++k)
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                C[k][t] = 0;
#pragma unroll
                for (int q = 0; q < NR; ++q) ring[q][k][t] = 0;
            }
#pragma unroll
        for (int yy = 0; yy < TR; ++yy) {  // all rows unrolled: ring slots and output rows are static
            const int q = yy % NR;
            uint32_t lw[4];
            const uint32_t *lp = lt + yy * ls4 + (xr >> 2);
#pragma unroll
            for (int j = 0; j < 4; ++j) lw[j] = __builtin_amdgcn_readfirstlane(lp[j]);
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
                uint32_t rw[5], ra[4];
                const uint32_t *rp = rt + yy * rs4 + rdw[k];
#pragma unroll
                for (int j = 0; j < 5; ++j) rw[j] = rp[j];
#pragma unroll
                for (int j = 0; j < 4; ++j) ra[j] = __builtin_amdgcn_alignbyte(rw[j + 1], rw[j], rsh[k]);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const int o = t + 4 * c;
                        const int nb = (NT - 4 * c) < 4 ? (NT - 4 * c) : 4;
                        const uint32_t mask = nb == 4 ? 0xffffffffu : ((1u << (8 * nb)) - 1u);
                        // L bytes o .. o + 3 (uniform: scalar shifts)
                        const uint64_t lpair = ((uint64_t)lw[(o >> 2) + ((o & 3) ? 1 : 0)] << 32) | lw[o >> 2];
                        const uint32_t lc = (uint32_t)(lpair >> (8 * (o & 3))) & mask;
                        uint32_t rc;
                        if (nb == 1)
                            rc = __builtin_amdgcn_ubfe(ra[o >> 2], 8 * (o & 3), 8);
                        else
                            rc = ((o & 3) ? __builtin_amdgcn_alignbyte(ra[(o >> 2) + 1], ra[o >> 2], o & 3)
                                          : ra[o >> 2]) & mask;
                        C[k][t] = __builtin_amdgcn_sad_u8(lc, rc, C[k][t]);
                    }
                }
            }
            const int yo = yy - 2 * KY;  // output row (rows past the image are merged but never written)
            if (yo >= 0) {
                uint32_t kb[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    kb[t] = ((C[0][t] - ring[q][0][t]) << ib) | (uint32_t)idx[0];
#pragma unroll
                    for (int k = 1; k < NCH; ++k)
                        kb[t] = min(kb[t], ((C[k][t] - ring[q][k][t]) << ib) | (uint32_t)idx[k]);
                }
                // reduce-scatter: v_permlane32_swap pairs lane k with k + 32 (columns
                // t and t + 4), v_permlane16_swap lane k with k + 16 within each
                // half (t and t + 2), DPP row_ror 8 lane k with k ^ 8 (t and t + 1);
                // then the minimum over each group of 8 lanes
                uint32_t m4[4], m2[2];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const auto r = __builtin_amdgcn_permlane32_swap(kb[j], kb[j + 4], false, false);
                    m4[j] = min((uint32_t)r[0], (uint32_t)r[1]);
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const auto r = __builtin_amdgcn_permlane16_swap(m4[j], m4[j + 2], false, false);
                    m2[j] = min((uint32_t)r[0], (uint32_t)r[1]);
                }
                const uint32_t r0 = min(m2[0], dpp_u32<DPP_ROW_ROR8, 0xf, false>(m2[0]));
                const uint32_t r1 = min(m2[1], dpp_u32<DPP_ROW_ROR8, 0xf, false>(m2[1]));
                uint32_t m1 = hi8 ? r1 : r0;
                m1 = min(m1, dpp_u32<DPP_QUAD_XOR1, 0xf, false>(m1));
                m1 = min(m1, dpp_u32<DPP_QUAD_XOR2, 0xf, false>(m1));
                m1 = min(m1, dpp_u32<DPP_ROW_HALF_MIRROR, 0xf, false>(m1));
                if ((lane & 7) == 0) atomicMin(&best[yo * CV_TX + xr + tcol], m1);
            }
#pragma unroll
            for (int k = 0; k < NCH; ++k)
#pragma unroll
                for (int t = 0; t < 8; ++t) ring[q][k][t] = C[k][t];
        }
    }
    __syncthreads();
    const uint32_t imask = (1u << ib) - 1u;  // ib <= 10
#pragma unroll
    for (int j = 0; j < WTA_TY / 8; ++j) {
        const int yo = (lane >> 3) + 8 * j, t = lane & 7, x = x0 + xr + t;
        if (yo < ylim && x < W) {
            const uint32_t key = best[yo * CV_TX + xr + t];
            int2 v;
            v.x = (int)(key & imask);
            v.y = (int)(key >> ib);
            reinterpret_cast<int2 *>(out)[(size_t)(y0 + yo) * W + x] = v;
        }
    }
}

using WtaKernel = void (*)(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int, int, int32_t *);

template <int NCH>
WtaKernel wta_table(int kx, int ky) {
#define WTA_ROW(X) {stereo_wta_kernel<X, 0, NCH>, stereo_wta_kernel<X, 1, NCH>, stereo_wta_kernel<X, 2, NCH>, \
                     stereo_wta_kernel<X, 3, NCH>}
    static const WtaKernel table[4][4] = {WTA_ROW(0), WTA_ROW(1), WTA_ROW(2), WTA_ROW(3)};
#undef WTA_ROW
    return table[kx][ky];
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

extern "C" umi_status umi_stereo_wta_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0,
                                        int step, int KY, int KX, int32_t *out, umi_stream s) {
    using umi::set_error;
    if (!L || !R || !out) return set_error(UMI_ERR_INVALID, "umi_stereo_wta_u8: null pointer");
    if (H < 1 || W < 1 || pitch < W || D < 1)
        return set_error(UMI_ERR_INVALID, "umi_stereo_wta_u8: H=%d W=%d pitch=%d D=%d", H, W, pitch, D);
    if (step != 1 && step != -1) return set_error(UMI_ERR_INVALID, "umi_stereo_wta_u8: step %d", step);
    if (KX < 0 || KX > 3 || KY < 0 || KY > 3)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_stereo_wta_u8: window radius KY=%d KX=%d (0..3)", KY, KX);
    if (D > 1024) return set_error(UMI_ERR_UNSUPPORTED, "umi_stereo_wta_u8: D=%d (at most 1024)", D);
    const int ib = umi::bits_for((uint32_t)(D - 1));
    const uint32_t maxcost = 255u * (2 * KX + 1) * (2 * KY + 1);
    if (umi::bits_for(maxcost) + ib > 32) return set_error(UMI_ERR_UNSUPPORTED, "umi_stereo_wta_u8: key overflow");
    const int ls4 = umi::ceil_div(CV_TX + 2 * KX, 4) + 2;
    const int rs4 = umi::ceil_div(CV_TX + 2 * KX + (D - 1) + 8 + 3, 4) + 1;  // + 3: the aligned start
    const size_t lds = 4 * ((size_t)(WTA_TY + 2 * KY) * (ls4 + rs4) + WTA_TY * CV_TX);
    dim3 grid(umi::ceil_div(W, CV_TX), umi::ceil_div(H, WTA_TY));
    // two 64-shift chunks per lane when D > 64
    const WtaKernel k = D > 64 ? wta_table<2>(KX, KY) : wta_table<1>(KX, KY);
    hipLaunchKernelGGL(k, grid, dim3(256), lds, umi::as_stream(s), L, R, H, W, pitch, D, s0, step, ls4, rs4, ib, out);
    return umi::launch_check("stereo_wta_kernel");
}
