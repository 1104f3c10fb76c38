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
// Three launches:
//  * pack: for every row and (padded) column, the 2KY+1 bytes of the column
//    window and the 2KX+1 bytes of the row window of L and of R, each as one
//    8-byte word, so that a window sum is two v_sad_u8;
//  * paths: every image row and column is walked both ways by a pair of
//    waves that meet in the middle; the two pairs of paths (rows, columns)
//    accumulate into their own partial sums Ph and Pv, so rows and columns
//    run concurrently in one launch;
//  * argmin: S = Pv + Ph, first minimum per pixel.
// The cost volume is never stored: each step updates a running window sum.
// The partial sums live in device memory as uint16 [H][W][Dp] (Dp = D
// rounded up to even; every sum of the four paths is < 2^16 by the host's
// check), one 32-bit word per lane.  Lanes hold disparities 2l and
// 2l + 1 (D <= 128); d - 1 and d + 1 across lanes come from DPP wave shifts,
// the minimum over d from DPP row reductions and the gfx950 permlane16/32
// swaps.  Loads run one block of 8 steps ahead of their use.  Integer
// arithmetic: bit-exact against the oracle.
#include "common.h"

namespace {
using namespace umi;

// Added to the running cost of lanes past the last disparity: their L_r stays
// in [2^29, 2^29 + Cmax + P2], above every real value (< 2^16), so it never
// wins a minimum and needs no select per step.
constexpr uint32_t SGM_BIAS = 0x20000000u;
constexpr int SGM_U = 8;      // steps per block; loads run one block ahead
constexpr int SGM_MAXK = 3;   // window radius limit (2K + 1 bytes fit one 8-byte word)

template <int CTRL>
__device__ inline uint32_t dpp_same(uint32_t v) {  // full-row permutations: every source lane exists
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}
// Minimum over the 64 lanes, in every lane: quad swaps and the half-row and
// row mirrors give each row's minimum; permlane16_swap pairs rows (0,1) and
// (2,3), permlane32_swap the two halves.  Needs all 64 lanes active.
__device__ inline uint32_t wave_min_all(uint32_t v) {
    v = min(v, dpp_same<0xb1>(v));   // quad_perm [1,0,3,2]
    v = min(v, dpp_same<0x4e>(v));   // quad_perm [2,3,0,1]
    v = min(v, dpp_same<0x141>(v));  // row_half_mirror
    v = min(v, dpp_same<0x140>(v));  // row_mirror
    const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    v = min((uint32_t)a[0], (uint32_t)a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return min((uint32_t)b[0], (uint32_t)b[1]);
}

// sum over the window of |L - R|: two words of up to 8 packed bytes each
// (unused bytes are zero in both)
__device__ inline uint32_t win_sad(uint2 l, uint2 r) {
    return __builtin_amdgcn_sad_u8(l.x, r.x, __builtin_amdgcn_sad_u8(l.y, r.y, 0u));
}

// ------------------------------------------------------------------- pack
// Packed words (byte j of a window in byte j of the word, unused bytes 0):
//  * for horizontal lines, Lv/Rv [H][Wh]: word i of row y = bytes of rows
//    clamp(y-KY .. y+KY) at column clamp(i - padh);
//  * for vertical lines, Lh/Rh [H + 2 SGM_PR][Wv]: word i of padded row r =
//    bytes of row clamp(r - SGM_PR) at columns clamp(i - KX - 1 - KX .. + KX).
// The padding lets every walk read one block of steps past either end of its
// line without clamping.
struct SgmGeom {
    int H, W, D, Dp, KY, KX;
    int smin, smax;  // shift range, capped to [-(W+KX), W+KX] (wider shifts read the same clamped words)
    int padh, Wh, Wv;
    size_t p_words, p_slack, pack_h, pack_v;  // P words (+ slack before and after), packed words per array
};

constexpr int SGM_SLACK = 2 * SGM_U;  // positions a walk may read past either end of its line
constexpr int SGM_PR = SGM_SLACK + SGM_MAXK + 1;  // padded rows above and below

SgmGeom sgm_geom(int H, int W, int D, int s0, int step, int KY, int KX) {
    SgmGeom g{};
    g.H = H, g.W = W, g.D = D, g.Dp = D + (D & 1), g.KY = KY, g.KX = KX;
    const int cap = W + KX, a = s0, b = s0 + step * (D - 1);
    g.smin = clampi(min(a, b), -cap, cap);
    g.smax = clampi(max(a, b), -cap, cap);
    g.padh = KX + SGM_SLACK + 1 + max(0, -g.smin);
    g.Wh = g.padh + W + KX + SGM_SLACK + 1 + max(0, g.smax);
    g.Wv = W + 2 * (KX + 1);
    g.p_slack = (size_t)SGM_SLACK * W * (g.Dp / 2) + 64;
    g.p_words = (size_t)H * W * (g.Dp / 2) + 2 * g.p_slack;
    g.pack_h = (size_t)H * g.Wh;
    g.pack_v = (size_t)(H + 2 * SGM_PR) * g.Wv;
    return g;
}
size_t sgm_p_bytes(const SgmGeom &g) { return (g.p_words * 4 + 255) & ~(size_t)255; }
size_t sgm_bytes(const SgmGeom &g) { return 2 * sgm_p_bytes(g) + 2 * (g.pack_h + g.pack_v) * sizeof(uint2); }

__device__ inline uint2 pack_col(const uint8_t *__restrict__ img, int pitch, int H, int y, int K, int cx) {
    uint32_t w[2] = {0, 0};
    for (int j = 0; j < 2 * K + 1; ++j)
        w[j >> 2] |= (uint32_t)img[(size_t)clampi(y - K + j, 0, H - 1) * pitch + cx] << (8 * (j & 3));
    return make_uint2(w[0], w[1]);
}
__device__ inline uint2 pack_row(const uint8_t *__restrict__ row, int W, int K, int c) {
    uint32_t w[2] = {0, 0};
    for (int j = 0; j < 2 * K + 1; ++j) w[j >> 2] |= (uint32_t)row[clampi(c - K + j, 0, W - 1)] << (8 * (j & 3));
    return make_uint2(w[0], w[1]);
}

// grid.y covers H + 2 SGM_PR rows; grid.x the wider of Wh and Wv
__global__ __launch_bounds__(256) void sgm_pack_kernel(const uint8_t *__restrict__ L, const uint8_t *__restrict__ R,
                                                       int pitch, SgmGeom g, uint2 *__restrict__ Lv,
                                                       uint2 *__restrict__ Rv, uint2 *__restrict__ Lh,
                                                       uint2 *__restrict__ Rh) {
    const int i = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
    const int y = r - SGM_PR;
    if (y >= 0 && y < g.H && i < g.Wh) {
        const int cx = clampi(i - g.padh, 0, g.W - 1);
        const size_t e = (size_t)y * g.Wh + i;
        Lv[e] = pack_col(L, pitch, g.H, y, g.KY, cx);
        Rv[e] = pack_col(R, pitch, g.H, y, g.KY, cx);
    }
    if (i < g.Wv) {
        const size_t row = (size_t)clampi(y, 0, g.H - 1) * pitch, e = (size_t)r * g.Wv + i;
        Lh[e] = pack_row(L + row, g.W, g.KX, i - g.KX - 1);
        Rh[e] = pack_row(R + row, g.W, g.KX, i - g.KX - 1);
    }
}

// -------------------------------------------------------------- line walks
struct SgmLines {
    const uint2 *Lv, *Rv, *Lh, *Rh;  // packed words (column windows, row windows)
    uint32_t *Pv, *Ph;  // partial sums of the vertical / horizontal pair of paths, uint16
                        // pairs: pixel n, lane l at P[p_slack + n * Dp/2 + l]
    SgmGeom g;
    int s0, step, P1, P2;
    int h_blocks;       // workgroups [0, h_blocks) walk rows, the rest walk columns
};

enum { SGM_STORE = 0, SGM_ADD = 1 };

// One wave's walk along one line in direction DIR.  VERT: the line is image
// column `line` (positions are rows), else image row `line` (positions are
// columns).  line and positions are wave-uniform; a load is a uniform base
// plus a constant per-lane offset (plus, within a block, an immediate for
// horizontal lines).
template <bool VERT, int DIR>
struct LineWalk {
    SgmLines a;
    int line, len, lane;
    bool valid[2];
    uint32_t roffb[2], laneb;  // lane's byte offsets: R word per k, P word
    __amdgpu_buffer_rsrc_t rL, rR, rP;  // buffer descriptors of the L and R words and of P
    int lq0;                   // byte offset in L of coordinate 0 (same stride as R)
    int rq0, rqs;              // byte offset in R of coordinate 0 (lane offset 0), per coordinate
    int pq0, pqs;              // byte offset in P of position 0 (lane 0), per position
    uint32_t c[2], Lp[2], M;

    int KA;  // window radius along the line

    __device__ LineWalk(const SgmLines &a_, int line_, int lane_) : a(a_), line(line_), lane(lane_) {
        const SgmGeom &g = a.g;
        len = VERT ? g.H : g.W;
        KA = VERT ? g.KY : g.KX;
        // vertical lines: row windows, running sum over rows; horizontal: column windows
        const uint2 *Lk = VERT ? a.Lh : a.Lv, *Rk = VERT ? a.Rh : a.Rv;
        uint32_t *P = VERT ? a.Pv : a.Ph;
        const int hw = g.Dp >> 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int d = 2 * lane + k, sk = clampi(a.s0 + a.step * d, g.smin, g.smax);
            valid[k] = d < g.D;
            // vertical: R column (clamped words repeat beyond [-KX-1, W+KX]); horizontal: shift
            roffb[k] = 8u * (VERT ? (uint32_t)(clampi(line + sk, -g.KX - 1, g.W + g.KX) + g.KX + 1)
                                  : (uint32_t)(sk - g.smin));
        }
        // lanes without a valid disparity address P at 2^31, outside the
        // buffer: their stores are dropped and their loads read 0
        laneb = valid[0] ? 4u * (uint32_t)lane : 0x80000000u;
        // the host keeps both arrays below 2^31 bytes
        const int kbytes = (int)(8 * (VERT ? g.pack_v : g.pack_h));
        rL = __builtin_amdgcn_make_buffer_rsrc((void *)Lk, 0, kbytes, 0x00020000);
        rR = __builtin_amdgcn_make_buffer_rsrc((void *)Rk, 0, kbytes, 0x00020000);
        rP = __builtin_amdgcn_make_buffer_rsrc((void *)P, 0, (int)(4 * g.p_words), 0x00020000);
        if (VERT) {
            lq0 = 8 * (SGM_PR * g.Wv + line + g.KX + 1);
            rq0 = 8 * SGM_PR * g.Wv, rqs = 8 * g.Wv;
            pq0 = 4 * ((int)g.p_slack + line * hw), pqs = 4 * g.W * hw;
        } else {
            lq0 = 8 * (line * g.Wh + g.padh);
            rq0 = 8 * (line * g.Wh + g.padh + g.smin), rqs = 8;
            pq0 = 4 * ((int)g.p_slack + line * g.W * hw), pqs = 4 * hw;
        }
    }
    __device__ int pos(int s) const { return DIR > 0 ? s : len - 1 - s; }
    // The L word is the same for every lane (lane offset 0).  imm: a constant
    // byte offset, which the compiler folds into the instruction.
    __device__ uint2 lw(int q, int imm = 0) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(rL, imm, lq0 + q * RQS(), 0);
        return make_uint2(v[0], v[1]);
    }
    __device__ uint2 rw(int q, int k, int imm = 0) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(rR, (int)roffb[k] + imm, rq0 + q * RQS(), 0);
        return make_uint2(v[0], v[1]);
    }
    __device__ int RQS() const { return VERT ? rqs : 8; }
    __device__ uint32_t pload(int p) const { return __builtin_amdgcn_raw_buffer_load_b32(rP, (int)laneb, pq0 + p * pqs, 0); }
    __device__ void pstore(int p, uint32_t v) const { __builtin_amdgcn_raw_buffer_store_b32(v, rP, (int)laneb, pq0 + p * pqs, 0); }
    __device__ uint32_t win(int q, int k) const { return win_sad(lw(q), rw(q, k)); }

    __device__ void recur(uint32_t c0, uint32_t c1) {
        // min over the d -/+ 1 neighbours; lanes 0 and 63 see their own value
        // in place of the missing one (d = -1, d = 128), which changes nothing
        const uint32_t n0 = min(Lp[1], (uint32_t)__builtin_amdgcn_update_dpp((int)Lp[1], (int)Lp[1], 0x138, 0xf, 0xf, false));
        const uint32_t n1 = min(Lp[0], (uint32_t)__builtin_amdgcn_update_dpp((int)Lp[0], (int)Lp[0], 0x130, 0xf, 0xf, false));
        const uint32_t mp = M + (uint32_t)a.P2, p1 = (uint32_t)a.P1;
        const uint32_t b0 = min(min(Lp[0], n0 + p1), mp) - M;
        const uint32_t b1 = min(min(Lp[1], n1 + p1), mp) - M;
        Lp[0] = c0 + b0;
        Lp[1] = c1 + b1;
        M = wave_min_all(min(Lp[0], Lp[1]));
    }

    template <int MODE>
    __device__ void emit(int p, uint32_t pv) {
        // Lp[0] < 2^16 wherever it is stored; an invalid Lp[1] (D odd) only
        // fills the unused upper half.  No carry between the halves: each
        // stays below 2^16
        const uint32_t packed = Lp[0] | (Lp[1] << 16);
        pstore(p, MODE == SGM_ADD ? pv + packed : packed);
    }

    // steps [sa, sb_end) of the walk (step 0 is the first pixel of the path).
    // Loads run one block ahead and may pass the end of the walk by up to
    // SGM_SLACK - 1 positions (into the padding and the slack of P).
    template <int MODE, bool NEEDP>
    __device__ void walk(int sa, int sb_end) {
        if (sa >= sb_end) return;
        if (sa == 0) {
            const int p = pos(0);
            const uint32_t pv = NEEDP ? pload(p) : 0u;
            c[0] = valid[0] ? 0u : SGM_BIAS;
            c[1] = valid[1] ? 0u : SGM_BIAS;
            for (int j = -KA; j <= KA; ++j) {
                c[0] += win(p + j, 0);
                c[1] += win(p + j, 1);
            }
            Lp[0] = c[0];
            Lp[1] = c[1];
            M = wave_min_all(min(Lp[0], Lp[1]));
            emit<MODE>(p, pv);
            sa = 1;
        }
        const int qa = DIR * KA, qs = -DIR * (KA + 1);  // entering / leaving window offsets
        uint2 la[SGM_U], ls[SGM_U], ra[SGM_U][2], rs[SGM_U][2];
        uint32_t pv[SGM_U];
        // L and R words and partial sums of steps s .. s + SGM_U - 1.  Rows:
        // one scalar offset per block and array, the step in the
        // instruction's immediate (from the lowest address of the block up);
        // columns: a scalar offset per load.
        auto issue = [&](int s) {
            const int p0 = pos(s), plo = DIR > 0 ? p0 : p0 - (SGM_U - 1);
#pragma unroll
            for (int u = 0; u < SGM_U; ++u) {
                if (VERT) {
                    const int p = p0 + DIR * u;
                    la[u] = lw(p + qa);
                    ls[u] = lw(p + qs);
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        ra[u][k] = rw(p + qa, k);
                        rs[u][k] = rw(p + qs, k);
                    }
                } else {
                    const int imm = 8 * (DIR > 0 ? u : SGM_U - 1 - u);
                    la[u] = lw(plo + qa, imm);
                    ls[u] = lw(plo + qs, imm);
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        ra[u][k] = rw(plo + qa, k, imm);
                        rs[u][k] = rw(plo + qs, k, imm);
                    }
                }
                if (NEEDP) pv[u] = pload(p0 + DIR * u);
            }
        };
        auto block = [&](int s, int n) {  // steps s .. s + n - 1 from the loaded block
            uint32_t dc[SGM_U][2], pc[SGM_U];
#pragma unroll
            for (int u = 0; u < SGM_U; ++u) {
#pragma unroll
                for (int k = 0; k < 2; ++k) dc[u][k] = win_sad(la[u], ra[u][k]) - win_sad(ls[u], rs[u][k]);
                pc[u] = NEEDP ? pv[u] : 0u;
            }
            if (n == SGM_U) {
                // The next block's loads reuse this block's registers: its
                // start position passes through empty asm statements that
                // consume the differences and partial sums, so no load is
                // placed before them.
                int sn = s + SGM_U;
#pragma unroll
                for (int u = 0; u < SGM_U; ++u) asm volatile("" : "+s"(sn) : "v"(dc[u][0]), "v"(dc[u][1]), "v"(pc[u]));
                issue(sn);  // in-bounds within the slack
            }
            const int p0 = pos(s);
#pragma unroll
            for (int u = 0; u < SGM_U; ++u) {
                if (u < n) {
                    c[0] += dc[u][0];
                    c[1] += dc[u][1];
                    recur(c[0], c[1]);
                    emit<MODE>(p0 + DIR * u, pc[u]);
                }
            }
        };
        issue(sa);
        int s = sa;
        for (; s + SGM_U <= sb_end; s += SGM_U) block(s, SGM_U);  // full blocks: no per-step branches
        if (s < sb_end) block(s, sb_end - s);
    }
};

template <bool VERT, int DIR>
__device__ void sgm_line(const SgmLines &a, bool active, int line, int lane) {
    LineWalk<VERT, DIR> lw(a, active ? line : 0, lane);
    const int m = lw.len / 2, n1 = DIR > 0 ? m : lw.len - m;
    if (active) lw.template walk<SGM_STORE, false>(0, n1);
    __syncthreads();
    if (active) lw.template walk<SGM_ADD, true>(n1, lw.len);
}

template <bool VERT>
__device__ void sgm_lines(const SgmLines &a, int blk, int w, int lane) {
    const int line = blk * 2 + (w >> 1);
    const bool active = line < (VERT ? a.g.W : a.g.H);
    if (w & 1)
        sgm_line<VERT, -1>(a, active, line, lane);
    else
        sgm_line<VERT, 1>(a, active, line, lane);
}

// All four paths in one launch.  Workgroup = 4 waves = two lines, each walked
// both ways by a pair of waves that meet in the middle: before the barrier
// each wave stores its L_r over its half of the line (forward: positions
// [0, len/2), backward: [len/2, len)), after it adds its L_r to what the
// other wave stored.  Workgroups [0, h_blocks) take rows (the longer walks,
// dispatched first), the rest columns.  Every wave reaches the barrier once.
__global__ __launch_bounds__(256) void sgm_paths_kernel(SgmLines a) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int b = blockIdx.x;
    if (b < a.h_blocks)
        sgm_lines<false>(a, b, w, lane);
    else
        sgm_lines<true>(a, b - a.h_blocks, w, lane);
}

// S = Pv + Ph per pixel; out = (first i with S minimal, that S).  A wave per
// pixel at a time (lane l holds i = 2l, 2l + 1), grid-stride over pixels.
__global__ __launch_bounds__(256) void sgm_argmin_kernel(const uint32_t *__restrict__ Pv,
                                                         const uint32_t *__restrict__ Ph, int npix, int D,
                                                         int32_t *__restrict__ out) {
    const int lane = threadIdx.x & 63, hw = (D + (D & 1)) >> 1;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    const bool v0 = 2 * lane < D, v1 = 2 * lane + 1 < D;
    for (int p = wave; p < npix; p += nwaves) {
        const size_t o = (size_t)p * hw + lane;
        const uint32_t a = v0 ? Pv[o] : 0u, b = v0 ? Ph[o] : 0u;
        // halves stay below 2^16 each, so one add sums both disparities
        const uint32_t s = a + b;
        uint32_t key = v0 ? (s & 0xffffu) << 8 | (uint32_t)(2 * lane) : 0xffffffffu;
        key = v1 ? min(key, (s >> 16) << 8 | (uint32_t)(2 * lane + 1)) : key;
        key = wave_min_all(key);
        if (lane == 0) {
            out[2 * (size_t)p] = (int32_t)(key & 0xff);
            out[2 * (size_t)p + 1] = (int32_t)(key >> 8);
        }
    }
}

umi_status sgm_check(int H, int W, int D, int s0, int step, int KY, int KX) {
    if (H < 1 || W < 1 || D < 1) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: H=%d W=%d D=%d", H, W, D);
    if (step != 1 && step != -1) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: step %d", step);
    if (KY < 0 || KY > SGM_MAXK || KX < 0 || KX > SGM_MAXK)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: window radius KY=%d KX=%d (0..3)", KY, KX);
    if (D > 128) return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: D=%d (at most 128)", D);
    if (s0 < -(1 << 28) || s0 > (1 << 28)) return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: s0 %d", s0);
    // the walks address each array with 32-bit byte offsets
    const SgmGeom g = sgm_geom(H, W, D, s0, step, KY, KX);
    if (4 * g.p_words >= (size_t)INT32_MAX || 8 * max(g.pack_h, g.pack_v) >= (size_t)INT32_MAX)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_sgm_u8: %dx%d with D=%d: partial sums of %zu B exceed 2 GiB", H, W,
                         D, 4 * g.p_words);
    return UMI_OK;
}

}  // namespace

extern "C" umi_status umi_sgm_workspace_size(int H, int W, int D, int s0, int step, int KY, int KX, size_t *bytes) {
    if (!bytes) return set_error(UMI_ERR_INVALID, "umi_sgm_workspace_size: null pointer");
    UMI_TRY(sgm_check(H, W, D, s0, step, KY, KX));
    *bytes = sgm_bytes(sgm_geom(H, W, D, s0, step, KY, KX));
    return UMI_OK;
}

extern "C" umi_status umi_sgm_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step,
                                 int KY, int KX, int P1, int P2, void *workspace, size_t workspace_bytes, int32_t *out,
                                 umi_stream s) {
    if (!L || !R || !out || !workspace) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: null pointer");
    UMI_TRY(sgm_check(H, W, D, s0, step, KY, KX));
    if (pitch < W) return set_error(UMI_ERR_INVALID, "umi_sgm_u8: pitch %d < W %d", pitch, W);
    const long cmax = 255L * (2 * KY + 1) * (2 * KX + 1);
    if (P1 < 0 || P2 < P1 || 4 * (cmax + P2) >= 65536)
        return set_error(UMI_ERR_INVALID, "umi_sgm_u8: P1=%d P2=%d (0 <= P1 <= P2, 4 (%ld + P2) < 65536)", P1, P2, cmax);
    const SgmGeom g = sgm_geom(H, W, D, s0, step, KY, KX);
    const size_t need = sgm_bytes(g);
    if (workspace_bytes < need)
        return set_error(UMI_ERR_INVALID, "umi_sgm_u8: workspace %zu B < %zu B", workspace_bytes, need);
    hipStream_t st = as_stream(s);
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    const size_t pb = sgm_p_bytes(g);
    uint32_t *Pv = reinterpret_cast<uint32_t *>(ws), *Ph = reinterpret_cast<uint32_t *>(ws + pb);
    uint2 *Lv = reinterpret_cast<uint2 *>(ws + 2 * pb);
    uint2 *Rv = Lv + g.pack_h, *Lh = Rv + g.pack_h, *Rh = Lh + g.pack_v;
    hipLaunchKernelGGL(sgm_pack_kernel, dim3(ceil_div(max(g.Wh, g.Wv), 256), H + 2 * SGM_PR), dim3(256), 0, st, L, R,
                       pitch, g, Lv, Rv, Lh, Rh);
    UMI_TRY(launch_check("sgm_pack_kernel"));
    SgmLines a{};
    a.Lv = Lv, a.Rv = Rv, a.Lh = Lh, a.Rh = Rh, a.Pv = Pv, a.Ph = Ph;
    a.g = g;
    a.s0 = s0, a.step = step, a.P1 = P1, a.P2 = P2;
    a.h_blocks = ceil_div(H, 2);
    hipLaunchKernelGGL(sgm_paths_kernel, dim3(a.h_blocks + ceil_div(W, 2)), dim3(256), 0, st, a);
    UMI_TRY(launch_check("sgm_paths_kernel"));
    const int npix = H * W;
    hipLaunchKernelGGL(sgm_argmin_kernel, dim3(min(ceil_div(npix, 4), 4096)), dim3(256), 0, st, Pv + g.p_slack,
                       Ph + g.p_slack, npix, D, out);
    return launch_check("sgm_argmin_kernel");
}
