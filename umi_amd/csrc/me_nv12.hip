// Full-search motion estimation over NV12 frames (gfx950), and the export of
// motion vectors in formats video codecs consume.
//
// No reference counterpart: main/full_me.cu (johnjohnlin/UMI) searches one
// luma plane; this extends the same search (clamp boundary, first minimum in
// raster order over (dy, dx) in [-R, R]^2) to the chroma of a 4:2:0 frame.
// The cost of an integer luma vector (dy, dx) for an 8x8 block is
//   SAD_Y(8x8) + SAD_U(4x4) + SAD_V(4x4),
// with the chroma predictor of H.264 8.4.2.2.2: the vector in quarter-luma
// units (4dy, 4dx) is an eighth-pel chroma vector, integer part
// floor(d / 2), fraction 4 (d & 1), so that a chroma sample is one of
//   A, (A + B + 1) >> 1, (A + C + 1) >> 1, (A + B + C + D + 2) >> 2
// (the general ((8-fx)(8-fy) A + fx (8-fy) B + (8-fx) fy C + fx fy D + 32) >> 6
// at fx, fy in {0, 4}), A..D = ref chroma at (Y, X), (Y, X+1), (Y+1, X),
// (Y+1, X+1), clamped to the (H/2) x (W/2) chroma plane.  U is byte 2x and V
// byte 2x + 1 of a uv row.
//
// Design: a workgroup (1024 threads) takes 30 blocks of one block row, so
// 990 threads own one (block, dy) each for R = 16 and a 1080p row is 8 whole
// workgroups.  It stages the luma window (8 + 32 rows, clamped; one 16-byte
// load per chunk when the plane is 16-byte aligned) and, for the chroma, the
// four half-pel phases of the U and V windows: the clamped samples, then one
// thread per sample dword computes its three half-pel phases from one set of
// loads.  A thread sweeps all dx with v_qsad_pk_u16_u8: two per luma row and
// group of four dx (its 11 window dwords read as 64-bit pairs), and for each
// chroma row, plane and x phase five more covering 4 chroma offsets each
// (consecutive offsets of one phase are every other dx).  The 16-bit sums
// (luma <= 16320, chroma <= 8160) are added per dx; a thread keeps the
// smallest (cost << 6 | dx + 16), so ties go to the smaller dx, and turns it
// into (cost << 11 | raster index) for the per-block LDS atomic min.
#include "common.h"
#include "stage.h"

#include <cstddef>

namespace {
using namespace umi;

constexpr int NV_RMAX = 16;               // search radius limit
constexpr int NV_TBX = 30;                // blocks per workgroup (x0 = 240 bx - 16 stays 16-byte aligned)
constexpr int NV_THREADS = 1024;           // >= NV_TBX * (2 NV_RMAX + 1), whole waves
constexpr int NV_LR = 8 + 2 * NV_RMAX;    // luma window rows
// luma window bytes per row, whole 16-byte chunks (a lane reads 44 bytes from 8k)
constexpr int NV_LW = NV_TBX * 8 + 2 * NV_RMAX + 16;
constexpr int NV_CR = 4 + NV_RMAX;        // chroma window rows (floor(dy/2) in [-8, 8])
constexpr int NV_CW = NV_TBX * 4 + NV_RMAX + 8;      // chroma window bytes per row
constexpr int NV_PS = NV_CW / 4 + 1;      // dwords per chroma plane row (one more for the +1 column)
static_assert(NV_THREADS >= NV_TBX * (2 * NV_RMAX + 1) && NV_THREADS % 64 == 0, "workgroup too small");
static_assert(NV_LW % 16 == 0 && NV_CW % 4 == 0 && (NV_TBX * 8) % 16 == 0, "row sizes");
static_assert(NV_LR * NV_LW / 16 <= NV_THREADS, "one 16-byte luma chunk per thread");

struct NvSmem {
    alignas(16) uint32_t lum[NV_LR * NV_LW / 4];    // luma window
    uint32_t pl[4][2][(NV_CR + 1) * NV_PS];         // [2 fy/4 + fx/4][U,V][row][dword]; phase 0 = the samples
    alignas(8) uint32_t curl[NV_TBX][8][2];         // current luma, 8 bytes per row
    uint32_t curc[NV_TBX][2][4];                    // current chroma, 4 bytes per row and plane
    uint32_t key[NV_TBX];
};

// FULL: R == NV_RMAX, every dx of the sweep is a candidate.
template <bool FULL>
__global__ __launch_bounds__(NV_THREADS) void me_nv12_kernel(const uint8_t *__restrict__ cy,
                                                             const uint8_t *__restrict__ cuv,
                                                             const uint8_t *__restrict__ ry,
                                                             const uint8_t *__restrict__ ruv, int H, int W,
                                                             int pitch_y, int pitch_uv, int R,
                                                             int32_t *__restrict__ out) {
    __shared__ NvSmem sm;
    const int tid = threadIdx.x;
    const int nbx = W / 8, by = blockIdx.y, bx0 = blockIdx.x * NV_TBX;
    const int HC = H / 2, WC = W / 2;
    const int y0 = by * 8 - NV_RMAX, x0 = bx0 * 8 - NV_RMAX;   // luma window origin
    const int cy0 = by * 4 - NV_RMAX / 2, cx0 = bx0 * 4 - NV_RMAX / 2;  // chroma window origin

    // ---- staging (clamped coordinates)
    if ((((reinterpret_cast<uintptr_t>(ry) | (uintptr_t)pitch_y) & 15) == 0) && W >= 16) {
        // x0 = 128 blockIdx.x - 16 is a multiple of 16
        stage_rows16<1>(sm.lum, NV_LR, NV_LW / 16, NV_LW / 16, x0, W, tid, NV_THREADS,
                        [&](int r) { return ry + (size_t)clampi(y0 + r, 0, H - 1) * pitch_y; });
    } else {
        const bool y_aligned = ((reinterpret_cast<uintptr_t>(ry) | (uintptr_t)pitch_y) & 3) == 0;
        for (int e = tid; e < NV_LR * NV_LW / 4; e += NV_THREADS) {
            const int rr = e / (NV_LW / 4), c4 = (e - rr * (NV_LW / 4)) * 4;
            const uint8_t *row = ry + (size_t)clampi(y0 + rr, 0, H - 1) * pitch_y;
            const int x = x0 + c4;  // a multiple of 4
            uint32_t v = 0;
            if (y_aligned && x >= 0 && x + 3 < W) {
                v = *reinterpret_cast<const uint32_t *>(row + x);
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b) v |= (uint32_t)row[clampi(x + b, 0, W - 1)] << (8 * b);
            }
            sm.lum[e] = v;
        }
    }
    // clamped chroma samples, U and V apart, 4 per dword (8 interleaved bytes per load inside the plane)
    const bool uv_aligned = ((reinterpret_cast<uintptr_t>(ruv) | (uintptr_t)pitch_uv) & 7) == 0;
    for (int e = tid; e < (NV_CR + 1) * (NV_CW / 4 + 1); e += NV_THREADS) {
        const int rr = e / (NV_CW / 4 + 1), c = e - rr * (NV_CW / 4 + 1), c4 = c * 4;
        const uint8_t *row = ruv + (size_t)clampi(cy0 + rr, 0, HC - 1) * pitch_uv;
        const int x = cx0 + c4;
        uint32_t u = 0, v = 0;
        if (uv_aligned && x >= 0 && x + 3 < WC) {
            const uint2 q = *reinterpret_cast<const uint2 *>(row + 2 * x);
            u = __builtin_amdgcn_perm(q.y, q.x, 0x06040200u);  // bytes 0, 2, 4, 6
            v = __builtin_amdgcn_perm(q.y, q.x, 0x07050301u);  // bytes 1, 3, 5, 7
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint8_t *sp = row + 2 * clampi(x + b, 0, WC - 1);
                u |= (uint32_t)sp[0] << (8 * b);
                v |= (uint32_t)sp[1] << (8 * b);
            }
        }
        sm.pl[0][0][rr * NV_PS + c] = u;
        sm.pl[0][1][rr * NV_PS + c] = v;
    }
    for (int e = tid; e < NV_TBX * 16; e += NV_THREADS) {  // current luma: [k][r][w]
        const int k = e >> 4, r = (e >> 1) & 7, w = e & 1, bx = bx0 + k;
        uint32_t v = 0;
        if (bx < nbx) {
            const uint8_t *p = cy + (size_t)(by * 8 + r) * pitch_y + bx * 8 + 4 * w;
#pragma unroll
            for (int b = 0; b < 4; ++b) v |= (uint32_t)p[b] << (8 * b);
        }
        sm.curl[k][r][w] = v;
    }
    for (int e = tid; e < NV_TBX * 8; e += NV_THREADS) {  // current chroma: [k][ch][j]
        const int k = e >> 3, ch = (e >> 2) & 1, j = e & 3, bx = bx0 + k;
        uint32_t v = 0;
        if (bx < nbx) {
            const uint8_t *p = cuv + (size_t)(by * 4 + j) * pitch_uv + 2 * (bx * 4) + ch;
#pragma unroll
            for (int i = 0; i < 4; ++i) v |= (uint32_t)p[2 * i] << (8 * i);
        }
        sm.curc[k][ch][j] = v;
    }
    if (tid < NV_TBX) sm.key[tid] = 0xffffffffu;
    __syncthreads();
    // the three half-pel phases of U and V from the samples, 4 per dword, one
    // thread per (plane, row, dword): (A+B+1)>>1 and (A+C+1)>>1 bytewise as
    // (a|b) - (((a^b) & 0xfe..) >> 1), (A+B+C+D+2)>>2 in 16-bit lanes
    for (int e = tid; e < 2 * NV_CR * (NV_CW / 4); e += NV_THREADS) {
        const int ch = e / (NV_CR * (NV_CW / 4)), r1 = e - ch * (NV_CR * (NV_CW / 4));
        const int rr = r1 / (NV_CW / 4), c = r1 - rr * (NV_CW / 4);
        const int o = rr * NV_PS + c;
        const uint32_t *s0 = sm.pl[0][ch] + o, *s1 = s0 + NV_PS;
        const uint32_t A = s0[0], B = __builtin_amdgcn_alignbyte(s0[1], s0[0], 1);
        const uint32_t C = s1[0], D = __builtin_amdgcn_alignbyte(s1[1], s1[0], 1);
        sm.pl[1][ch][o] = (A | B) - (((A ^ B) & 0xfefefefeu) >> 1);
        sm.pl[2][ch][o] = (A | C) - (((A ^ C) & 0xfefefefeu) >> 1);
        const uint32_t m = 0x00ff00ffu;
        const uint32_t lo = (A & m) + (B & m) + (C & m) + (D & m) + 0x00020002u;
        const uint32_t hi = ((A >> 8) & m) + ((B >> 8) & m) + ((C >> 8) & m) + ((D >> 8) & m) + 0x00020002u;
        sm.pl[3][ch][o] = ((lo >> 2) & m) | (((hi >> 2) & m) << 8);
    }
    __syncthreads();

    // ---- search: thread = (dy, block k)
    const int S = 2 * R + 1;
    if (tid < NV_TBX * S) {
        const int dyi = tid / NV_TBX, k = tid - dyi * NV_TBX, dy = dyi - R;
        uint64_t acc[9];
#pragma unroll
        for (int g = 0; g < 9; ++g) acc[g] = 0;
#pragma unroll 1
        for (int r = 0; r < 8; ++r) {
            const uint32_t *wrow = sm.lum + (r + dy + NV_RMAX) * (NV_LW / 4) + 2 * k;
            uint32_t w[11];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const uint2 t = reinterpret_cast<const uint2 *>(wrow)[j];
                w[2 * j] = t.x;
                w[2 * j + 1] = t.y;
            }
            w[10] = wrow[10];
            const uint2 c = *reinterpret_cast<const uint2 *>(sm.curl[k][r]);
#pragma unroll
            for (int g = 0; g < 9; ++g) {
                acc[g] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w[g] | ((uint64_t)w[g + 1] << 32), c.x, acc[g]);
                acc[g] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w[g + 1] | ((uint64_t)w[g + 2] << 32), c.y, acc[g]);
            }
        }
        // chroma: cacc[fx][m] holds offsets floor(dx/2) + 8 = 4m .. 4m+3 of x phase fx
        const int fy = dy & 1, oy = dy >> 1;
        uint64_t cacc[2][5];
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int m = 0; m < 5; ++m) cacc[f][m] = 0;
#pragma unroll 1
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t cc = sm.curc[k][ch][j];
                const int rr = j + oy + NV_RMAX / 2;
#pragma unroll
                for (int fx = 0; fx < 2; ++fx) {
                    const uint32_t *prow = sm.pl[2 * fy + fx][ch] + rr * NV_PS + k;
                    uint32_t d[6];
#pragma unroll
                    for (int j2 = 0; j2 < 6; ++j2) d[j2] = prow[j2];
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        cacc[fx][m] =
                            __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)d[m] | ((uint64_t)d[m + 1] << 32), cc, cacc[fx][m]);
                }
            }
        // candidates: luma index l = dx + 16 (compile-time), chroma x phase
        // l & 1 at offset index l >> 1; cost <= 24480 < 2^15
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int l = 0; l <= 2 * NV_RMAX; ++l) {
            const uint32_t ls = (uint32_t)(acc[l >> 2] >> (16 * (l & 3))) & 0xffffu;
            const int ci = l >> 1;
            const uint32_t cs = (uint32_t)(cacc[l & 1][ci >> 2] >> (16 * (ci & 3))) & 0xffffu;
            const uint32_t key = (ls + cs) << 6 | (uint32_t)l;
            if (FULL || (l >= NV_RMAX - R && l <= NV_RMAX + R)) best = min(best, key);
        }
        const uint32_t key = (best >> 6) << 11 | (uint32_t)(dyi * S + (int)(best & 63u) - (NV_RMAX - R));
        if (bx0 + k < nbx) atomicMin(&sm.key[k], key);
    }
    __syncthreads();
    if (tid < NV_TBX && bx0 + tid < nbx) {
        const uint32_t key = sm.key[tid], idx = key & 0x7ffu;
        int32_t *o = out + ((size_t)by * nbx + bx0 + tid) * 3;
        o[0] = (int32_t)(idx / S) - R;
        o[1] = (int32_t)(idx % S) - R;
        o[2] = (int32_t)(key >> 11);
    }
}

// FFmpeg's exported motion vector record (libavutil/motion_vector.h,
// AVMotionVector), the side data a decoder or filter graph hands on.
struct AvMotionVector {
    int32_t source;          // -1: from the past reference
    uint8_t w, h;            // block size
    int16_t src_x, src_y;    // absolute source position (block centre in the reference)
    int16_t dst_x, dst_y;    // absolute destination position (block centre)
    uint64_t flags;
    int32_t motion_x, motion_y;  // src = dst + motion / motion_scale
    uint16_t motion_scale;
};
static_assert(sizeof(AvMotionVector) == 40 && offsetof(AvMotionVector, flags) == 16 &&
                  offsetof(AvMotionVector, motion_x) == 24 && offsetof(AvMotionVector, motion_scale) == 32,
              "AVMotionVector layout");

__global__ __launch_bounds__(256) void mv_export_kernel(const int32_t *__restrict__ mv, int nby, int nbx, int B,
                                                        int qshift, int format, void *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nby * nbx) return;
    const int by = i / nbx, bx = i - by * nbx;
    const int my = mv[3 * (size_t)i] << qshift, mx = mv[3 * (size_t)i + 1] << qshift;  // quarter pels
    if (format == 0) {
        reinterpret_cast<int16_t *>(out)[2 * (size_t)i] = (int16_t)mx;
        reinterpret_cast<int16_t *>(out)[2 * (size_t)i + 1] = (int16_t)my;
    } else {
        AvMotionVector r{};
        r.source = -1;
        r.w = r.h = (uint8_t)B;
        r.dst_x = (int16_t)(bx * B + B / 2);
        r.dst_y = (int16_t)(by * B + B / 2);
        r.src_x = (int16_t)(r.dst_x + (mx >> 2));
        r.src_y = (int16_t)(r.dst_y + (my >> 2));
        r.flags = 0;
        r.motion_x = mx;
        r.motion_y = my;
        r.motion_scale = 4;
        reinterpret_cast<AvMotionVector *>(out)[i] = r;
    }
}

}  // namespace

extern "C" umi_status umi_me_full_search_nv12(const uint8_t *cur_y, const uint8_t *cur_uv, const uint8_t *ref_y,
                                              const uint8_t *ref_uv, int H, int W, int pitch_y, int pitch_uv, int R,
                                              int32_t *mv_cost, umi_stream s) {
    if (!cur_y || !cur_uv || !ref_y || !ref_uv || !mv_cost)
        return set_error(UMI_ERR_INVALID, "umi_me_full_search_nv12: null pointer");
    if (H < 8 || W < 8 || (H & 1) || (W & 1) || pitch_y < W || pitch_uv < W)
        return set_error(UMI_ERR_INVALID, "umi_me_full_search_nv12: H=%d W=%d pitch_y=%d pitch_uv=%d", H, W, pitch_y,
                         pitch_uv);
    if (R < 1 || R > NV_RMAX) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_nv12: radius %d (1..16)", R);
    const int nby = H / 8, nbx = W / 8;
    if (R == NV_RMAX)
        hipLaunchKernelGGL(me_nv12_kernel<true>, dim3(ceil_div(nbx, NV_TBX), nby), dim3(NV_THREADS), 0, as_stream(s),
                           cur_y, cur_uv, ref_y, ref_uv, H, W, pitch_y, pitch_uv, R, mv_cost);
    else
        hipLaunchKernelGGL(me_nv12_kernel<false>, dim3(ceil_div(nbx, NV_TBX), nby), dim3(NV_THREADS), 0, as_stream(s),
                           cur_y, cur_uv, ref_y, ref_uv, H, W, pitch_y, pitch_uv, R, mv_cost);
    return launch_check("me_nv12_kernel");
}

extern "C" umi_status umi_me_export_mvs(const int32_t *mv, int nby, int nbx, int B, int units, int format, void *out,
                                        umi_stream s) {
    if (!mv || !out) return set_error(UMI_ERR_INVALID, "umi_me_export_mvs: null pointer");
    if (nby < 0 || nbx < 0 || B < 1 || B > 255 || (units != 1 && units != 4) || (format != 0 && format != 1))
        return set_error(UMI_ERR_INVALID, "umi_me_export_mvs: nby=%d nbx=%d B=%d units=%d format=%d", nby, nbx, B,
                         units, format);
    if ((size_t)nby * B + B > 32767 || (size_t)nbx * B + B > 32767)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_me_export_mvs: frame too large for 16-bit positions");
    const int n = nby * nbx;
    if (n == 0) return UMI_OK;
    hipLaunchKernelGGL(mv_export_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(s), mv, nby, nbx, B,
                       units == 1 ? 2 : 0, format, out);
    return launch_check("mv_export_kernel");
}
