/*
 * umi_oracle.c — CPU oracle for the UMI operator workloads.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the CPU check
 * loops that johnjohnlin/UMI uses to validate each example (the main/ programs), with
 * the generalisations the MI355X library implements (two frames, uint8
 * pixels, argmin, 2-D windows).  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as the checker.  The
 * library never calls it.
 *
 * Every function cites the reference loop it follows.  How each is pinned
 * (closed forms, brute force, invariants) is listed in the tests that load it.
 * Build: oracle/Makefile (gcc, -O2, -ffp-contract=off: float expressions are
 * evaluated exactly as written, fused multiply-adds only where fmaf() is
 * spelled out; -fopenmp).
 *
 * Threads: oracle_set_threads(n) spreads the outer loop (block rows, image
 * rows) of the ME, cost-volume, filter, bilateral, GEMM and convolution
 * loops over n OpenMP threads (default 1, the reference's single thread);
 * rows are independent, so the results do not depend on n.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int g_threads = 1;
void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/*
 * Full-search block motion estimation.
 * Follows main/full_me.cu:66-87: for each block (raster), for each search
 * offset (raster, dy outer), SAD over the block with the searched pixel
 * clamped to the frame (max(min(.,h-1),0), full_me.cu:77-78).  Generalised:
 * cur/ref are two uint8 frames, offsets run -R..R, and the first minimum in
 * raster order (strict <) is reported as (dy, dx, sad).
 * Blocks are processed for block rows [by_begin, by_end) only (a bounded
 * sample for timing); pass 0, H/B for the whole frame.
 */
void oracle_me_full_search_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int B, int R,
                              int by_begin, int by_end, int32_t *mv_sad, int32_t *volume) {
    const int nbx = W / B, S = 2 * R + 1;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int by = by_begin; by < by_end; ++by) {
        for (int bx = 0; bx < nbx; ++bx) {
            const int y = by * B, x = bx * B;
            const size_t blk = (size_t)by * nbx + bx;
            int32_t best = INT32_MAX, bdy = 0, bdx = 0;
            for (int sy = 0; sy < S; ++sy) {
                for (int sx = 0; sx < S; ++sx) {
                    int32_t sum = 0;
                    for (int py = 0; py < B; ++py) {
                        const int spy = clampi(y + py + sy - R, 0, H - 1);
                        const uint8_t *c = cur + (size_t)(y + py) * pitch + x;
                        const uint8_t *r = ref + (size_t)spy * pitch;
                        for (int px = 0; px < B; ++px) {
                            const int spx = clampi(x + px + sx - R, 0, W - 1);
                            sum += abs((int)c[px] - (int)r[spx]);
                        }
                    }
                    if (volume) volume[(blk * S + sy) * S + sx] = sum;
                    if (sum < best) {
                        best = sum;
                        bdy = sy - R;
                        bdx = sx - R;
                    }
                }
            }
            if (mv_sad) {
                mv_sad[blk * 3 + 0] = bdy;
                mv_sad[blk * 3 + 1] = bdx;
                mv_sad[blk * 3 + 2] = best;
            }
        }
    }
}

/*
 * Full-search SAD volume on int32 frames: the exact check loop of
 * main/full_me.cu:66-87 (one frame searched against itself there; here cur
 * and ref may differ), search offsets ofs..ofs+S-1 in each axis.
 */
void oracle_sad_volume_i32(const int32_t *cur, const int32_t *ref, int H, int W, int B, int S, int ofs,
                           int32_t *volume) {
    size_t c = 0;
    for (int y = 0; y + B <= H; y += B)
        for (int x = 0; x + B <= W; x += B)
            for (int sy = 0; sy < S; ++sy)
                for (int sx = 0; sx < S; ++sx) {
                    int32_t sum = 0;
                    for (int py = 0; py < B; ++py)
                        for (int px = 0; px < B; ++px) {
                            const int spy = clampi(y + py + sy + ofs, 0, H - 1);
                            const int spx = clampi(x + px + sx + ofs, 0, W - 1);
                            sum += abs(cur[(size_t)(y + py) * W + (x + px)] - ref[(size_t)spy * W + spx]);
                        }
                    volume[c++] = sum;
                }
}

/*
 * SAD cost volume.  Follows main/cost_volume.cu:67-80 (clamp applied to each
 * final coordinate, cost_volume.cu:73-74), generalised to a left/right pair
 * and a (2KY+1)x(2KX+1) window: shift s = s0 + step*i, rows [y_begin, y_end).
 */
void oracle_cost_volume_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step,
                           int KY, int KX, int y_begin, int y_end, int32_t *out) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int y = y_begin; y < y_end; ++y)
        for (int x = 0; x < W; ++x)
            for (int i = 0; i < D; ++i) {
                const int s = s0 + step * i;
                int32_t sad = 0;
                for (int dy = -KY; dy <= KY; ++dy) {
                    const int cy = clampi(y + dy, 0, H - 1);
                    for (int dx = -KX; dx <= KX; ++dx) {
                        const int lx = clampi(x + dx, 0, W - 1);
                        const int rx = clampi(x + dx + s, 0, W - 1);
                        sad += abs((int)L[(size_t)cy * pitch + lx] - (int)R[(size_t)cy * pitch + rx]);
                    }
                }
                out[((size_t)y * W + x) * D + i] = sad;
            }
}

/*
 * Winner-take-all over the cost volume above, per pixel: the first i (in
 * increasing order) whose cost is minimal, and that cost.  No loop of the
 * reference computes it; it is the consumer of the volume that
 * main/cost_volume.cu:35-65 writes.  Rows [y_begin, y_end); out[y][x] = (i,
 * cost).
 */
void oracle_stereo_wta_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step,
                          int KY, int KX, int y_begin, int y_end, int32_t *out) {
    for (int y = y_begin; y < y_end; ++y)
        for (int x = 0; x < W; ++x) {
            int32_t best = 0;
            int bi = -1;
            for (int i = 0; i < D; ++i) {
                const int s = s0 + step * i;
                int32_t sad = 0;
                for (int dy = -KY; dy <= KY; ++dy) {
                    const int cy = clampi(y + dy, 0, H - 1);
                    for (int dx = -KX; dx <= KX; ++dx) {
                        const int lx = clampi(x + dx, 0, W - 1);
                        const int rx = clampi(x + dx + s, 0, W - 1);
                        sad += abs((int)L[(size_t)cy * pitch + lx] - (int)R[(size_t)cy * pitch + rx]);
                    }
                }
                if (bi < 0 || sad < best) {
                    best = sad;
                    bi = i;
                }
            }
            out[((size_t)y * W + x) * 2] = bi;
            out[((size_t)y * W + x) * 2 + 1] = best;
        }
}

/*
 * Box sums of a uint8 image by brute force: the sum over rows
 * [max(y-ry,0), min(y+ry,H-1)] and columns [max(x-rx,0), min(x+rx,W-1)]
 * (the window clipped to the image), uint32.  No loop of the reference
 * computes it; it checks the box sums read from the integral image of
 * main/iimg.cu:37-82.
 */
void oracle_box_sum_u8(const uint8_t *in, int H, int W, int ip, int ry, int rx, uint32_t *out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint32_t s = 0;
            for (int yy = y - ry < 0 ? 0 : y - ry; yy <= (y + ry > H - 1 ? H - 1 : y + ry); ++yy)
                for (int xx = x - rx < 0 ? 0 : x - rx; xx <= (x + rx > W - 1 ? W - 1 : x + rx); ++xx)
                    s += in[(size_t)yy * ip + xx];
            out[(size_t)y * W + x] = s;
        }
}

/*
 * Rectangle (Haar-like) features by brute force over the uint8 image: at
 * every window position, sum over the feature's rectangles (x, y, w, h,
 * weight; weight 0 = unused) of weight x the pixel sum of the rectangle
 * moved to the position; unsigned arithmetic (int32 with wrap-around).  No
 * loop of the reference computes it; it checks features read from the
 * integral image of main/iimg.cu:37-82.
 */
void oracle_haar_u8(const uint8_t *in, int H, int W, int ip, const int32_t *feats, int nfeat, int win_h, int win_w,
                    int32_t *out) {
    const int OH = H - win_h + 1, OW = W - win_w + 1;
    for (int f = 0; f < nfeat; ++f)
        for (int y = 0; y < OH; ++y)
            for (int x = 0; x < OW; ++x) {
                uint32_t v = 0;
                for (int k = 0; k < 3; ++k) {
                    const int32_t *r = feats + (f * 3 + k) * 5;
                    if (r[4] == 0) continue;
                    uint32_t s = 0;
                    for (int yy = y + r[1]; yy < y + r[1] + r[3]; ++yy)
                        for (int xx = x + r[0]; xx < x + r[0] + r[2]; ++xx) s += in[(size_t)yy * ip + xx];
                    v += (uint32_t)r[4] * s;
                }
                out[((size_t)f * OH + y) * OW + x] = (int32_t)v;
            }
}

/*
 * Quarter-pel refinement of 8x8 block motion vectors (no loop of the
 * reference computes it; it refines the integer vectors of the full search,
 * main/full_me.cu:66-87 generalised).  Reference samples at quarter-pel
 * positions by bilinear interpolation with rounding, coordinates clamped to
 * the frame; from c = 4 * (integer MV): the 8 half-pel neighbours c + 2 o,
 * then the 8 quarter-pel neighbours of the best, o in {-1,0,1}^2 \ {0} in
 * raster order, each replacing the centre only when its SAD is smaller.
 * mv_in / mv_out: int32 [H/8][W/8][3] = (dy, dx, SAD); mv_out in quarter pels.
 */
static int floor4(int q) { return q >= 0 ? q / 4 : -((-q + 3) / 4); }

static int refq(const uint8_t *ref, int H, int W, int pitch, int qy, int qx) {
    const int Y = floor4(qy), X = floor4(qx), fy = qy - 4 * Y, fx = qx - 4 * X;
    const int y0 = clampi(Y, 0, H - 1), y1 = clampi(Y + 1, 0, H - 1);
    const int x0 = clampi(X, 0, W - 1), x1 = clampi(X + 1, 0, W - 1);
    const int A = ref[(size_t)y0 * pitch + x0], B = ref[(size_t)y0 * pitch + x1];
    const int C = ref[(size_t)y1 * pitch + x0], D = ref[(size_t)y1 * pitch + x1];
    return ((4 - fx) * (4 - fy) * A + fx * (4 - fy) * B + (4 - fx) * fy * C + fx * fy * D + 8) >> 4;
}

static int32_t subpel_sad(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int by, int bx, int vy,
                          int vx) {
    int32_t s = 0;
    for (int py = 0; py < 8; ++py)
        for (int px = 0; px < 8; ++px) {
            const int y = 8 * by + py, x = 8 * bx + px;
            s += abs((int)cur[(size_t)y * pitch + x] - refq(ref, H, W, pitch, 4 * y + vy, 4 * x + vx));
        }
    return s;
}

void oracle_me_subpel_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, const int32_t *mv_in,
                         int32_t *mv_out) {
    const int nby = H / 8, nbx = W / 8;
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            const int32_t *m = mv_in + ((size_t)by * nbx + bx) * 3;
            int cy = 4 * m[0], cx = 4 * m[1];
            int32_t best = subpel_sad(cur, ref, H, W, pitch, by, bx, cy, cx);
            for (int step = 2; step >= 1; step /= 2) {
                const int oy = cy, ox = cx;
                for (int k = 0; k < 9; ++k) {
                    if (k == 4) continue;
                    const int vy = oy + step * (k / 3 - 1), vx = ox + step * (k % 3 - 1);
                    const int32_t s = subpel_sad(cur, ref, H, W, pitch, by, bx, vy, vx);
                    if (s < best) {
                        best = s;
                        cy = vy;
                        cx = vx;
                    }
                }
            }
            int32_t *o = mv_out + ((size_t)by * nbx + bx) * 3;
            o[0] = cy;
            o[1] = cx;
            o[2] = best;
        }
}

/* Reads in[y][x] or applies the boundary: mode 0 -> pad value, 1 -> clamp
 * (the PAD / "WRAP"=clamp modes of src/UMI.h:205-218). */
#define BOUNDED(T, in, H, W, ip, y, x, mode, pad) \
    (((y) >= 0 && (y) < (H) && (x) >= 0 && (x) < (W)) ? (in)[(size_t)(y) * (ip) + (x)] \
     : ((mode) == 0 ? (pad) : (in)[(size_t)clampi((y), 0, (H)-1) * (ip) + clampi((x), 0, (W)-1)]))

/*
 * 3x3 box filter: main/filter3x3.cu:60-78 (sum = init; sum += pixel or pad;
 * sum *= mul).  Integer form wraps (unsigned arithmetic), float adds in order.
 */
void oracle_filter3x3_i32(const int32_t *in, int H, int W, int ip, int32_t init, int32_t mul, int32_t pad,
                          int mode, int32_t *out, int op) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint32_t sum = (uint32_t)init;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) sum += (uint32_t)BOUNDED(int32_t, in, H, W, ip, y + dy, x + dx, mode, pad);
            out[(size_t)y * op + x] = (int32_t)(sum * (uint32_t)mul);
        }
}

void oracle_filter3x3_f32(const float *in, int H, int W, int ip, float init, float mul, float pad, int mode,
                          float *out, int op) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum = init;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) sum += BOUNDED(float, in, H, W, ip, y + dy, x + dx, mode, pad);
            out[(size_t)y * op + x] = sum * mul;
        }
}

/*
 * Weighted 3x3 filter: main/wfilter3x3.cu:71-89 (sum = init; sum += w * x,
 * dy outer, dx inner).  The float form states the multiply-add as fmaf, the
 * contraction the strategy's `accum += x * w` (wfilter3x3.cu:13) gets on the
 * GPU.  _f64 evaluates the same sum in double (error reference).
 */
void oracle_wfilter3x3_i32(const int32_t *in, int H, int W, int ip, const int32_t *w, int32_t init, int32_t pad,
                           int mode, int32_t *out, int op) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint32_t sum = (uint32_t)init;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    sum += (uint32_t)w[3 * (dy + 1) + dx + 1] *
                           (uint32_t)BOUNDED(int32_t, in, H, W, ip, y + dy, x + dx, mode, pad);
            out[(size_t)y * op + x] = (int32_t)sum;
        }
}

void oracle_wfilter3x3_f32(const float *in, int H, int W, int ip, const float *w, float init, float pad, int mode,
                           float *out, int op) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum = init;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    sum = fmaf(BOUNDED(float, in, H, W, ip, y + dy, x + dx, mode, pad), w[3 * (dy + 1) + dx + 1], sum);
            out[(size_t)y * op + x] = sum;
        }
}

/* FP64 value and magnitude sum_{taps} |w x| (+|init|) of the weighted filter. */
void oracle_wfilter3x3_f64(const float *in, int H, int W, int ip, const float *w, float init, float pad, int mode,
                           double *out, double *mag) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double sum = init, m = fabs((double)init);
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const double t = (double)BOUNDED(float, in, H, W, ip, y + dy, x + dx, mode, pad) *
                                     (double)w[3 * (dy + 1) + dx + 1];
                    sum += t;
                    m += fabs(t);
                }
            out[(size_t)y * W + x] = sum;
            mag[(size_t)y * W + x] = m;
        }
}

/*
 * Bilateral filter: main/bilateral.cu:103-124 with the strategy's order
 * (bilateral.cu:12-23): for dy (outer), dx: n = pixel or pad,
 * w = expf(d*d*norm) * ws[dy] * ws[dx], wsum += w, wxsum += w*n (as fmaf),
 * out = wxsum / wsum.  _f64 is the same in double with exp().
 */
void oracle_bilateral_f32(const float *in, int H, int W, int ip, int r, const float *ws, float norm, float pad,
                          int mode, float *out, int op) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float c = in[(size_t)y * ip + x];
            float wsum = 0.f, wxsum = 0.f;
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx) {
                    const float n = BOUNDED(float, in, H, W, ip, y + dy, x + dx, mode, pad);
                    const float d = n - c;
                    const float w = expf(d * d * norm) * ws[dy + r] * ws[dx + r];
                    wsum += w;
                    wxsum = fmaf(w, n, wxsum);
                }
            out[(size_t)y * op + x] = wxsum / wsum;
        }
}

void oracle_bilateral_f64(const float *in, int H, int W, int ip, int r, const float *ws, float norm, float pad,
                          int mode, double *out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double c = in[(size_t)y * ip + x];
            double wsum = 0, wxsum = 0;
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx) {
                    const double n = BOUNDED(float, in, H, W, ip, y + dy, x + dx, mode, pad);
                    const double d = n - c;
                    const double w = exp(d * d * (double)norm) * (double)ws[dy + r] * (double)ws[dx + r];
                    wsum += w;
                    wxsum += w * n;
                }
            out[(size_t)y * W + x] = wxsum / wsum;
        }
}

/*
 * Integral image: main/iimg.cu:84-112 (first check: ans[y][x] = rowsum +
 * ans[y-1][x]).  u8 -> u32 and i32 -> i32 in wrapping unsigned arithmetic;
 * the f64 form accumulates float inputs exactly enough to serve as the error
 * reference of the float path.
 */
void oracle_iimg_u8_u32(const uint8_t *in, int H, int W, int ip, uint32_t *out) {
    for (int y = 0; y < H; ++y) {
        uint32_t rowsum = 0;
        for (int x = 0; x < W; ++x) {
            rowsum += in[(size_t)y * ip + x];
            out[(size_t)y * W + x] = rowsum + (y ? out[(size_t)(y - 1) * W + x] : 0u);
        }
    }
}

void oracle_iimg_i32(const int32_t *in, int H, int W, int ip, int32_t *out) {
    for (int y = 0; y < H; ++y) {
        uint32_t rowsum = 0;
        for (int x = 0; x < W; ++x) {
            rowsum += (uint32_t)in[(size_t)y * ip + x];
            out[(size_t)y * W + x] = (int32_t)(rowsum + (y ? (uint32_t)out[(size_t)(y - 1) * W + x] : 0u));
        }
    }
}

void oracle_iimg_f64(const float *in, int H, int W, int ip, double *out) {
    for (int y = 0; y < H; ++y) {
        double rowsum = 0;
        for (int x = 0; x < W; ++x) {
            rowsum += in[(size_t)y * ip + x];
            out[(size_t)y * W + x] = rowsum + (y ? out[(size_t)(y - 1) * W + x] : 0.0);
        }
    }
}

/*
 * Dense GEMM, the loop of main/gemm.cu:69-78 (for y, for x: sum over k of
 * A[y][k] * B[k][x]) in float, for rows [r0, r1) of C: the CPU baseline of
 * the precision config (the FP64 oracle of the tests is numpy).
 */
void oracle_gemm_f32_rows(const float *A, const float *B, float *C, int M, int N, int K, int r0, int r1) {
    (void)M;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int y = r0; y < r1; ++y)
        for (int x = 0; x < N; ++x) {
            float acc = 0.f;
            for (int k = 0; k < K; ++k) acc = fmaf(A[(size_t)y * K + k], B[(size_t)k * N + x], acc);
            C[(size_t)y * N + x] = acc;
        }
}

/*
 * Direct convolution, the loop of main/fprop.cu:84-108 (NCHW input, KCRS
 * filter, zero padding) in float, for images [n0, n1): the CPU baseline of
 * the precision config.
 */
void oracle_conv_f32(const float *x, const float *w, float *out, int N, int C, int H, int W, int K, int R, int S,
                     int stride, int pad, int n0, int n1) {
    (void)N;
    const int OH = (H + 2 * pad - R) / stride + 1, OW = (W + 2 * pad - S) / stride + 1;
#pragma omp parallel for num_threads(g_threads) schedule(static) collapse(2)
    for (int n = n0; n < n1; ++n)
        for (int k = 0; k < K; ++k)
            for (int oy = 0; oy < OH; ++oy)
                for (int ox = 0; ox < OW; ++ox) {
                    float acc = 0.f;
                    for (int c = 0; c < C; ++c)
                        for (int r = 0; r < R; ++r)
                            for (int s2 = 0; s2 < S; ++s2) {
                                const int iy = oy * stride + r - pad, ix = ox * stride + s2 - pad;
                                if (iy < 0 || iy >= H || ix < 0 || ix >= W) continue;
                                acc = fmaf(w[(((size_t)k * C + c) * R + r) * S + s2],
                                           x[(((size_t)n * C + c) * H + iy) * W + ix], acc);
                            }
                    out[(((size_t)n * K + k) * OH + oy) * OW + ox] = acc;
                }
}

/*
 * Full-search ME over NV12 frames: luma SAD of the 8x8 block plus the SAD of
 * its 4x4 U and V blocks.  No reference loop computes it (main/full_me.cu is
 * luma-only); it extends that search to the chroma planes of a 4:2:0 frame.
 * The chroma predictor of an integer luma vector (dy, dx) follows H.264
 * 8.4.2.2.2 (chroma sample interpolation) with the vector in quarter-luma
 * units (4dy, 4dx): chroma offset (4d) >> 3 = floor(d / 2), eighth-pel
 * fraction f = (4d) & 7 in {0, 4},
 *   pred = ((8-fx)(8-fy) A + fx (8-fy) B + (8-fx) fy C + fx fy D + 32) >> 6,
 * A..D = ref chroma at (Y,X), (Y,X+1), (Y+1,X), (Y+1,X+1), clamped to the
 * chroma plane ((H/2) x (W/2) samples, U at byte 2x and V at byte 2x + 1 of
 * a uv row).  cost = SAD_Y + SAD_U + SAD_V; out int32 [H/8][W/8][3] = (dy,
 * dx, cost) of the first minimum in raster order over (dy, dx) in [-R, R]^2.
 */
int oracle_nv12_chroma_pred(const uint8_t *uv, int HC, int WC, int pitch, int ch, int Y, int X, int fy, int fx) {
    const int y0 = clampi(Y, 0, HC - 1), y1 = clampi(Y + 1, 0, HC - 1);
    const int x0 = clampi(X, 0, WC - 1), x1 = clampi(X + 1, 0, WC - 1);
    const int A = uv[(size_t)y0 * pitch + 2 * x0 + ch], B = uv[(size_t)y0 * pitch + 2 * x1 + ch];
    const int C = uv[(size_t)y1 * pitch + 2 * x0 + ch], D = uv[(size_t)y1 * pitch + 2 * x1 + ch];
    return ((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6;
}

void oracle_me_nv12_u8(const uint8_t *cy, const uint8_t *cuv, const uint8_t *ry, const uint8_t *ruv, int H, int W,
                       int pitch_y, int pitch_uv, int R, int32_t *out) {
    const int nby = H / 8, nbx = W / 8, HC = H / 2, WC = W / 2, S = 2 * R + 1;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            int32_t best = INT32_MAX, bdy = 0, bdx = 0;
            for (int sy = 0; sy < S; ++sy)
                for (int sx = 0; sx < S; ++sx) {
                    const int dy = sy - R, dx = sx - R;
                    int32_t sum = 0;
                    for (int py = 0; py < 8; ++py)
                        for (int px = 0; px < 8; ++px) {
                            const int c = cy[(size_t)(8 * by + py) * pitch_y + 8 * bx + px];
                            const int r = ry[(size_t)clampi(8 * by + py + dy, 0, H - 1) * pitch_y +
                                             clampi(8 * bx + px + dx, 0, W - 1)];
                            sum += abs(c - r);
                        }
                    const int oy = (4 * dy) >> 3, fy = (4 * dy) & 7, ox = (4 * dx) >> 3, fx = (4 * dx) & 7;
                    for (int ch = 0; ch < 2; ++ch)
                        for (int j = 0; j < 4; ++j)
                            for (int i = 0; i < 4; ++i) {
                                const int c = cuv[(size_t)(4 * by + j) * pitch_uv + 2 * (4 * bx + i) + ch];
                                const int p = oracle_nv12_chroma_pred(ruv, HC, WC, pitch_uv, ch, 4 * by + j + oy,
                                                               4 * bx + i + ox, fy, fx);
                                sum += abs(c - p);
                            }
                    if (sum < best) {
                        best = sum;
                        bdy = dy;
                        bdx = dx;
                    }
                }
            int32_t *o = out + ((size_t)by * nbx + bx) * 3;
            o[0] = bdy;
            o[1] = bdx;
            o[2] = best;
        }
}

/*
 * Semi-global matching over the cost volume of oracle_cost_volume_u8 (the
 * volume main/cost_volume.cu:35-65 computes, generalised as above).  No
 * reference loop computes it: it is the aggregation step downstream of that
 * volume.  Four paths r (from the left, right, top, bottom):
 *   L_r(p, i) = C(p, i) + min(L_r(q, i), L_r(q, i-1) + P1, L_r(q, i+1) + P1,
 *                            M_r(q) + P2) - M_r(q),   q = p - r,
 *   M_r(q) = min_k L_r(q, k); the first pixel of a path has L_r = C;
 *   neighbours outside [0, D) are left out.  S = sum of the four L_r; out =
 *   (first i with S minimal, that S).  Integer arithmetic, exact.
 */
static void sgm_path(const int32_t *C, int H, int W, int D, int P1, int P2, int dy, int dx, int64_t *S) {
    int32_t *prev = (int32_t *)malloc(sizeof(int32_t) * D), *cur = (int32_t *)malloc(sizeof(int32_t) * D);
    /* walk each path line: rows for horizontal paths, columns for vertical */
    const int nlines = dy ? W : H, len = dy ? H : W;
    for (int l = 0; l < nlines; ++l) {
        for (int t = 0; t < len; ++t) {
            const int s = (dx > 0 || dy > 0) ? t : len - 1 - t;
            const int y = dy ? s : l, x = dy ? l : s;
            const int32_t *c = C + ((size_t)y * W + x) * D;
            if (t == 0) {
                for (int i = 0; i < D; ++i) cur[i] = c[i];
            } else {
                int32_t m = prev[0];
                for (int i = 1; i < D; ++i) m = prev[i] < m ? prev[i] : m;
                for (int i = 0; i < D; ++i) {
                    int32_t best = prev[i];
                    if (i > 0 && prev[i - 1] + P1 < best) best = prev[i - 1] + P1;
                    if (i + 1 < D && prev[i + 1] + P1 < best) best = prev[i + 1] + P1;
                    if (m + P2 < best) best = m + P2;
                    cur[i] = c[i] + best - m;
                }
            }
            for (int i = 0; i < D; ++i) S[((size_t)y * W + x) * D + i] += cur[i];
            int32_t *tmp = prev;
            prev = cur;
            cur = tmp;
        }
    }
    free(prev);
    free(cur);
}

void oracle_sgm_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step, int KY,
                   int KX, int P1, int P2, int32_t *out) {
    int32_t *C = (int32_t *)malloc(sizeof(int32_t) * (size_t)H * W * D);
    int64_t *S = (int64_t *)calloc((size_t)H * W * D, sizeof(int64_t));
    oracle_cost_volume_u8(L, R, H, W, pitch, D, s0, step, KY, KX, 0, H, C);
    sgm_path(C, H, W, D, P1, P2, 0, 1, S);
    sgm_path(C, H, W, D, P1, P2, 0, -1, S);
    sgm_path(C, H, W, D, P1, P2, 1, 0, S);
    sgm_path(C, H, W, D, P1, P2, -1, 0, S);
    for (size_t p = 0; p < (size_t)H * W; ++p) {
        int best = 0;
        for (int i = 1; i < D; ++i)
            if (S[p * D + i] < S[p * D + best]) best = i;
        out[2 * p] = best;
        out[2 * p + 1] = (int32_t)S[p * D + best];
    }
    free(C);
    free(S);
}
