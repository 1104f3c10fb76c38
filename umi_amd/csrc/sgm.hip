// Semi-global matching (SGM) over the SAD cost volume of
// umi_cost_volume_u8, gfx950.  No reference counterpart: the aggregation step
// downstream of the volume that the MotionEst launch of
// main/cost_volume.cu:35-65 (johnjohnlin/UMI) writes.  With C[y][x][i] that
// volume (window (2KY+1) x (2KX+1), shift s = s0 + step*i, clamped
// coordinates) and four paths r (from the left, right, top, bottom):
//
//   L_r(p, i) = C(p, i) + min(L_r(q, i), L_r(q, i-1) + P1, L_r(q, i+1) + P1,
//                            M_r(q) + P2) - M_r(q),     q = p - r,
//   M_r(q) = min_k L_r(q, k), L_r = C on the first pixel of a path,
//   S(p, i) = sum_r L_r(p, i),  out(p) = (first i with S minimal, that S).
//
// The cost volume is never stored: every sweep recomputes its costs from the
// image pair with running window sums.  Only the partial sum S lives in
// device memory, as uint16 [H][W][D] (every L_r <= C + P2, and the host
// requires 4 (Cmax + P2) < 2^16).  Two launches:
//  * vertical: a wave per image column walks it down (S = L_down) and up
//    (S += L_up); the column's L bytes and the R bytes of every shift are
//    staged in LDS in chunks of rows;
//  * horizontal: a wave per image row walks it left to right (S += L_right,
//    loads of S prefetched 8 steps ahead) and back, adding L_left in
//    registers and taking the argmin; (i, S) per pixel is the output.
// Lanes hold disparities i and i + 64 (D <= 128); the d +- 1 neighbours come
// from DPP wave shifts, the minimum over d from DPP row reductions and four
// lane reads.  Integer arithmetic: bit-exact against the oracle.
#include "common.h"

namespace {
using namespace umi;

constexpr uint32_t SGM_INF = 0x3fffffffu;
constexpr int SGM_KD = 2;    // disparities per lane
constexpr int SGM_CR = 64;   // output rows per staged chunk of the vertical kernel
constexpr int SGM_PF = 8;    // prefetch distance of the horizontal kernel (steps)

// lane i <- lane i - 1 (lane 0 <- fill); lane i <- lane i + 1 (lane 63 <- fill)
__device__ inline uint32_t wave_shr1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ inline uint32_t wave_shl1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x130, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline uint32_t dpp_same(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xf, 0xf, false);
}
// Minimum over the 64 lanes (quad swaps, half-row and row mirrors, then the
// four row results), the same in every lane.
__device__ inline uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp_same<0xb1>(v));   // quad_perm [1,0,3,2]
    v = min(v, dpp_same<0x4e>(v));   // quad_perm [2,3,0,1]
    v = min(v, dpp_same<0x141>(v));  // row_half_mirror
    v = min(v, dpp_same<0x140>(v));  // row_mirror
    const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}

// One step of a path: Lp (previous pixel's values, INF where d >= D) and its
// minimum M become the current pixel's values for costs c.
__device__ inline void sgm_step(uint32_t (&Lp)[SGM_KD], uint32_t &M, const uint32_t (&c)[SGM_KD],
                                const bool (&valid)[SGM_KD], bool first, uint32_t P1, uint32_t P2) {
    uint32_t Ln[SGM_KD];
    if (first) {
#pragma unroll
        for (int k = 0; k < SGM_KD; ++k) Ln[k] = valid[k] ? c[k] : SGM_INF;
    } else {
        const uint32_t top0 = __builtin_amdgcn_readlane(Lp[0], 63), bot1 = __builtin_amdgcn_readlane(Lp[1], 0);
#pragma unroll
        for (int k = 0; k < SGM_KD; ++k) {
            const uint32_t lo = wave_shr1(Lp[k], k == 0 ? SGM_INF : top0);          // d - 1
            const uint32_t hi = wave_shl1(Lp[k], k + 1 < SGM_KD ? bot1 : SGM_INF);  // d + 1
            const uint32_t best = min(Lp[k], min(min(lo, hi) + P1, M + P2));
            Ln[k] = valid[k] ? c[k] + best - M : SGM_INF;
        }
    }
    uint32_t m = SGM_INF;
#pragma unroll
    for (int k = 0; k < SGM_KD; ++k) {
        Lp[k] = Ln[k];
        m = min(m, Ln[k]);
    }
    M = wave_min_u32(m);
}

// ---------------------------------------------------------------- vertical
// Workgroup = 4 waves = 4 consecutive columns x0 .. x0 + 3.  A chunk stages
// rows ya .. ya + NROWS - 1 (ya = c0 - KY - 1, rows clamped) of L columns
// x0 - KX .. x0 + 3 + KX and of R columns x0 - KX + smin .. x0 + 3 + KX + smax
// (clamped), so every term of the window sums of rows c0 - KY - 1 ..
// c0 + CR + KY is an LDS byte.
template <int KY, int KX>
__global__ __launch_bounds__(256) void sgm_vertical_kernel(const uint8_t *__restrict__ L, const uint8_t *__restrict__ R,
                                                           int H, int W, int pitch, int D, int s0, int step, int smin,
                                                           int rw, int P1, int P2, uint16_t *__restrict__ S) {
    constexpr int NROWS = SGM_CR + 2 * KY + 2, LW = 4 + 2 * KX;
    extern __shared__ uint8_t sgm_lds[];
    const int lrow = LW + rw;  // bytes per staged row: L part, then R part
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * 4, x = x0 + wv;
    const bool active = x < W;
    const int lbase = x0 - KX, rbase = x0 - KX + smin;
    int sft[SGM_KD], lcol[2 * KX + 1], rcol[SGM_KD][2 * KX + 1];
    bool valid[SGM_KD];
#pragma unroll
    for (int k = 0; k < SGM_KD; ++k) {
        const int d = lane + 64 * k;
        valid[k] = active && d < D;
        sft[k] = s0 + step * d;
#pragma unroll
        for (int j = 0; j < 2 * KX + 1; ++j) rcol[k][j] = valid[k] ? x + j - KX + sft[k] - rbase : 0;
    }
#pragma unroll
    for (int j = 0; j < 2 * KX + 1; ++j) lcol[j] = x + j - KX - lbase;
    const int nch = ceil_div(H, SGM_CR);
    auto stage = [&](int c0) {
        const int ya = c0 - KY - 1;
        __syncthreads();
        for (int e = tid; e < NROWS * lrow; e += 256) {
            const int r = e / lrow, b = e - r * lrow;
            const size_t row = (size_t)clampi(ya + r, 0, H - 1) * pitch;
            sgm_lds[e] = b < LW ? L[row + clampi(lbase + b, 0, W - 1)] : R[row + clampi(rbase + b - LW, 0, W - 1)];
        }
        __syncthreads();
        return ya;
    };
    // horizontal window sum of image row y (LDS row y - ya) for the lane's disparities
    auto hsum = [&](int ya, int y, uint32_t (&h)[SGM_KD]) {
        const uint8_t *rowp = sgm_lds + (y - ya) * lrow;
#pragma unroll
        for (int k = 0; k < SGM_KD; ++k) {
            uint32_t a = 0;
#pragma unroll
            for (int j = 0; j < 2 * KX + 1; ++j) a += (uint32_t)abs((int)rowp[lcol[j]] - (int)rowp[LW + rcol[k][j]]);
            h[k] = a;
        }
    };
    uint32_t Lp[SGM_KD] = {SGM_INF, SGM_INF}, M = 0, c[SGM_KD] = {0, 0};
    uint16_t *col = S + (size_t)x * D;
    const size_t rs = (size_t)W * D;
    // down: S = L_down
    for (int ch = 0; ch < nch; ++ch) {
        const int c0 = ch * SGM_CR, c1 = min(H, c0 + SGM_CR);
        const int ya = stage(c0);
        for (int y = c0; y < c1; ++y) {
            uint32_t h[SGM_KD];
            if (y == 0) {
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k) c[k] = 0;
                for (int j = -KY; j <= KY; ++j) {
                    hsum(ya, j, h);
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] += h[k];
                }
            } else {
                uint32_t g[SGM_KD];
                hsum(ya, y - 1 - KY, h);
                hsum(ya, y + KY, g);
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k) c[k] = c[k] - h[k] + g[k];
            }
            sgm_step(Lp, M, c, valid, y == 0, P1, P2);
#pragma unroll
            for (int k = 0; k < SGM_KD; ++k)
                if (valid[k]) col[(size_t)y * rs + lane + 64 * k] = (uint16_t)Lp[k];
        }
    }
    // up: S += L_up
    for (int ch = nch - 1; ch >= 0; --ch) {
        const int c0 = ch * SGM_CR, c1 = min(H, c0 + SGM_CR);
        const int ya = stage(c0);
        for (int y = c1 - 1; y >= c0; --y) {
            uint32_t h[SGM_KD];
            uint16_t sv[SGM_KD];
#pragma unroll
            for (int k = 0; k < SGM_KD; ++k) sv[k] = valid[k] ? col[(size_t)y * rs + lane + 64 * k] : (uint16_t)0;
            if (y == H - 1) {
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k) c[k] = 0;
                for (int j = -KY; j <= KY; ++j) {
                    hsum(ya, H - 1 + j, h);
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] += h[k];
                }
            } else {
                uint32_t g[SGM_KD];
                hsum(ya, y + 1 + KY, h);
                hsum(ya, y - KY, g);
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k) c[k] = c[k] - h[k] + g[k];
            }
            sgm_step(Lp, M, c, valid, y == H - 1, P1, P2);
#pragma unroll
            for (int k = 0; k < SGM_KD; ++k)
                if (valid[k]) col[(size_t)y * rs + lane + 64 * k] = (uint16_t)(sv[k] + Lp[k]);
        }
    }
}

// -------------------------------------------------------------- horizontal
// Workgroup = 4 waves = 4 consecutive rows y0 .. y0 + 3; LDS holds image rows
// y0 - KY .. y0 + 3 + KY (clamped) of L and of R, whole width.
template <int KY, int KX>
__global__ __launch_bounds__(256) void sgm_horizontal_kernel(const uint8_t *__restrict__ L,
                                                             const uint8_t *__restrict__ R, int H, int W, int pitch,
                                                             int D, int s0, int step, int P1, int P2,
                                                             uint16_t *__restrict__ S, int32_t *__restrict__ out) {
    constexpr int NR = 4 + 2 * KY;
    extern __shared__ uint8_t sgm_lds[];
    uint8_t *ls = sgm_lds, *rs_ = sgm_lds + NR * W;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int y0 = blockIdx.x * 4, y = y0 + wv;
    for (int e = tid; e < NR * W; e += 256) {
        const int r = e / W, xx = e - r * W;
        const size_t src = (size_t)clampi(y0 - KY + r, 0, H - 1) * pitch + xx;
        ls[e] = L[src];
        rs_[e] = R[src];
    }
    __syncthreads();
    if (y >= H) return;  // no barrier follows
    int sft[SGM_KD];
    bool valid[SGM_KD];
#pragma unroll
    for (int k = 0; k < SGM_KD; ++k) {
        const int d = lane + 64 * k;
        valid[k] = d < D;
        sft[k] = s0 + step * d;
    }
    const uint8_t *lrow = ls + (y - y0) * W, *rrow = rs_ + (y - y0) * W;  // row y - KY; rows step by W
    // vertical window sum at column xx (clamped inside) for the lane's disparities
    auto vsum = [&](int xx, uint32_t (&v)[SGM_KD]) {
        const int cx = clampi(xx, 0, W - 1);
#pragma unroll
        for (int k = 0; k < SGM_KD; ++k) {
            const int rx = clampi(xx + sft[k], 0, W - 1);
            uint32_t a = 0;
#pragma unroll
            for (int j = 0; j < 2 * KY + 1; ++j) a += (uint32_t)abs((int)lrow[j * W + cx] - (int)rrow[j * W + rx]);
            v[k] = a;
        }
    };
    uint16_t *srow = S + (size_t)y * W * D;
    auto sload = [&](int xx, uint16_t (&sv)[SGM_KD]) {
        const int cx = clampi(xx, 0, W - 1);
#pragma unroll
        for (int k = 0; k < SGM_KD; ++k) sv[k] = valid[k] ? srow[(size_t)cx * D + lane + 64 * k] : (uint16_t)0;
    };
    uint32_t Lp[SGM_KD] = {SGM_INF, SGM_INF}, M = 0, c[SGM_KD] = {0, 0};
    uint16_t cur[SGM_PF][SGM_KD], nxt[SGM_PF][SGM_KD];
    // left to right: S += L_right
#pragma unroll
    for (int u = 0; u < SGM_PF; ++u) sload(u, cur[u]);
    for (int xb = 0; xb < W; xb += SGM_PF) {
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u) sload(xb + SGM_PF + u, nxt[u]);
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u) {
            const int xx = xb + u;
            if (xx < W) {
                uint32_t h[SGM_KD], g[SGM_KD];
                if (xx == 0) {
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] = 0;
                    for (int j = -KX; j <= KX; ++j) {
                        vsum(j, h);
#pragma unroll
                        for (int k = 0; k < SGM_KD; ++k) c[k] += h[k];
                    }
                } else {
                    vsum(xx - 1 - KX, h);
                    vsum(xx + KX, g);
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] = c[k] - h[k] + g[k];
                }
                sgm_step(Lp, M, c, valid, xx == 0, P1, P2);
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k)
                    if (valid[k]) srow[(size_t)xx * D + lane + 64 * k] = (uint16_t)(cur[u][k] + Lp[k]);
            }
        }
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u)
#pragma unroll
            for (int k = 0; k < SGM_KD; ++k) cur[u][k] = nxt[u][k];
    }
    // right to left: L_left in registers, S complete, argmin
    const int xlast = W - 1;
#pragma unroll
    for (int u = 0; u < SGM_PF; ++u) sload(xlast - u, cur[u]);
    for (int xb = xlast; xb >= 0; xb -= SGM_PF) {
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u) sload(xb - SGM_PF - u, nxt[u]);
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u) {
            const int xx = xb - u;
            if (xx >= 0) {
                uint32_t h[SGM_KD], g[SGM_KD];
                if (xx == xlast) {
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] = 0;
                    for (int j = -KX; j <= KX; ++j) {
                        vsum(xlast + j, h);
#pragma unroll
                        for (int k = 0; k < SGM_KD; ++k) c[k] += h[k];
                    }
                } else {
                    vsum(xx + 1 + KX, h);
                    vsum(xx - KX, g);
#pragma unroll
                    for (int k = 0; k < SGM_KD; ++k) c[k] = c[k] - h[k] + g[k];
                }
                sgm_step(Lp, M, c, valid, xx == xlast, P1, P2);
                uint32_t key = 0xffffffffu;
#pragma unroll
                for (int k = 0; k < SGM_KD; ++k)
                    if (valid[k]) key = min(key, ((uint32_t)cur[u][k] + Lp[k]) << 8 | (uint32_t)(lane + 64 * k));
                key = wave_min_u32(key);
                if (lane == 0) {
                    out[((size_t)y * W + xx) * 2] = (int32_t)(key & 0xff);
                    out[((size_t)y * W + xx) * 2 + 1] = (int32_t)(key >> 8);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SGM_PF; ++u)
#pragma unroll
            for (int k = 0; k < SGM_KD; ++k) cur[u][k] = nxt[u][k];
    }
}

template <int KY, int KX>
umi_status sgm_launch(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step, int P1,
                      int P2, uint16_t *S, int32_t *out, hipStream_t st) {
    const int smin = min(s0, s0 + step * (D - 1)), smax = max(s0, s0 + step * (D - 1));
    const int rw = 4 + 2 * KX + (smax - smin);
    const size_t lds_v = (size_t)(SGM_CR + 2 * KY + 2) * (4 + 2 * KX + rw);
    const size_t lds_h = (size_t)2 * (4 + 2 * KY) * W;
    static const hipError_t a1 = hipFuncSetAttribute((const void *)sgm_vertical_kernel<KY, KX>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    static const hipError_t a2 = hipFuncSetAttribute((const void *)sgm_horizontal_kernel<KY, KX>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    UMI_TRY(hip_check(a1, "hipFuncSetAttribute (sgm vertical)"));
    UMI_TRY(hip_check(a2, "hipFuncSetAttribute (sgm horizontal)"));
    hipLaunchKernelGGL((sgm_vertical_kernel<KY, KX>), dim3(ceil_div(W, 4)), dim3(256), lds_v, st, L, R, H, W, pitch, D,
                       s0, step, smin, rw, P1, P2, S);
    UMI_TRY(launch_check("sgm_vertical_kernel"));
    hipLaunchKernelGGL((sgm_horizontal_kernel<KY, KX>), dim3(ceil_div(H, 4)), dim3(256), lds_h, st, L, R, H, W, pitch,
                       D, s0, step, P1, P2, S, out);
    return launch_check("sgm_horizontal_kernel");
}

using SgmLaunch = umi_status (*)(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int,
                                 uint16_t *, int32_t *, hipStream_t);
#define SGM_ROW(Y) {sgm_launch<Y, 0>, sgm_launch<Y, 1>, sgm_launch<Y, 2>, sgm_launch<Y, 3>}
const SgmLaunch kSgm[4][4] = {SGM_ROW(0), SGM_ROW(1), SGM_ROW(2), SGM_ROW(3)};
#undef SGM_ROW

}  // namespace

extern "C" umi_status umi_sgm_workspace_size(int H, int W, int D, size_t *bytes) {
    if (!bytes || H < 0 || W < 0 || D < 0) return set_error(UMI_ERR_INVALID, "umi_sgm_workspace_size: bad arguments");
    *bytes = (size_t)H * W * D * sizeof(uint16_t);
    return UMI_OK;
}

extern "C" umi_status umi_sgm_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step,
                                 int KY, int KX, int P1, int P2, void *workspace, size_t workspace_bytes, int32_t *out,
                                 umi_stream s) {
    if (!L || !R || !out || !workspace) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: null pointer");
    if (H < 1 || W < 1 || pitch < W || D < 1)
        return set_error(UMI_ERR_INVALID, "umi_sgm_u8: H=%d W=%d pitch=%d D=%d", H, W, pitch, D);
    if (step != 1 && step != -1) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: step %d", step);
    if (KY < 0 || KY > 3 || KX < 0 || KX > 3)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: window radius KY=%d KX=%d (0..3)", KY, KX);
    if (D > 64 * SGM_KD) return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: D=%d (at most %d)", D, 64 * SGM_KD);
    if ((size_t)2 * (4 + 2 * KY) * W > 160 * 1024)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: W=%d too wide for the row staging", W);
    const long cmax = 255L * (2 * KY + 1) * (2 * KX + 1);
    if (P1 < 0 || P2 < P1 || 4 * (cmax + P2) >= 65536)
        return set_error(UMI_ERR_INVALID, "umi_sgm_u8: P1=%d P2=%d (0 <= P1 <= P2, 4 (%ld + P2) < 65536)", P1, P2, cmax);
    if (workspace_bytes < (size_t)H * W * D * sizeof(uint16_t))
        return set_error(UMI_ERR_INVALID, "umi_sgm_u8: workspace %zu B < %zu B", workspace_bytes,
                         (size_t)H * W * D * sizeof(uint16_t));
    return kSgm[KY][KX](L, R, H, W, pitch, D, s0, step, P1, P2, static_cast<uint16_t *>(workspace), out, as_stream(s));
}
