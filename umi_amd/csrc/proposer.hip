// Learned motion proposals for full-search block ME (gfx950).
//
// No reference counterpart: johnjohnlin/UMI has only the exhaustive search
// (main/full_me.cu).  This file holds
//  * synthetic frame pairs: smoothed noise, warped by a smooth motion field
//    (sum of 4 Gaussian-weighted affine bumps, components clamped to +-32)
//    plus N(0, sigma^2) noise, from a counter-based hash (deterministic);
//  * the proposal network's training kernels in fp32 (forward, backward to
//    inputs and to weights/biases, Huber loss, Adam): NCHW, 3x3 dilated;
//  * the inference kernels around the bf16 MFMA layers (gemm.hip):
//    downsampling, the 2->32 first layer and 32->2 last layer, block
//    proposals, the exact +-2 SAD refinement with the neighbourhood test and
//    the fallback list.
#include "common.h"

namespace {
using namespace umi;

// ---------------------------------------------------------------- hashing
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
    return mix64(seed ^ mix64(a * 0x100000001b3ull + b));
}
__device__ inline float u01(uint64_t h) { return (float)((h >> 40) + 0.5) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- synthesis
// f0 = 5x5 box (edge-clamped, (sum + 12) / 25) of uniform uint8 noise.
__global__ void synth_frame0_kernel(uint8_t *f0, int H, int W, int pitch, uint64_t seed) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    int acc = 0;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            acc += (int)(hash3(seed, clampi(y + dy, 0, H - 1), clampi(x + dx, 0, W - 1)) & 0xff);
    f0[(size_t)y * pitch + x] = (uint8_t)((acc + 12) / 25);
}

struct Bump {  // mv(p) += exp(-|p-c|^2 / (2 s^2)) * (t + A (p - c))
    float cy, cx, inv2s2, ty, tx, a00, a01, a10, a11;
};

__device__ inline void field_at(const Bump *b, float y, float x, float &my, float &mx) {
    my = mx = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float dy = y - b[k].cy, dx = x - b[k].cx;
        const float g = __expf(-(dy * dy + dx * dx) * b[k].inv2s2);
        my += g * (b[k].ty + b[k].a00 * dy + b[k].a01 * dx);
        mx += g * (b[k].tx + b[k].a10 * dy + b[k].a11 * dx);
    }
    my = fminf(fmaxf(my, -32.f), 32.f);
    mx = fminf(fmaxf(mx, -32.f), 32.f);
}

// f1(y, x) = clip(round(f0[cl(y + rint(my))][cl(x + rint(mx))] + N(0, sigma^2))).
__global__ void synth_warp_kernel(const uint8_t *f0, uint8_t *f1, int H, int W, int pitch, const Bump *bumps,
                                  float sigma, uint64_t seed) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    __shared__ Bump b[4];
    if (threadIdx.x < 4) b[threadIdx.x] = bumps[threadIdx.x];
    __syncthreads();
    float my, mx;
    field_at(b, (float)y, (float)x, my, mx);
    const int sy = clampi(y + (int)rintf(my), 0, H - 1), sx = clampi(x + (int)rintf(mx), 0, W - 1);
    const uint64_t h = hash3(seed ^ 0x5bd1e995ull, y, x);
    const float u1 = u01(h), u2 = u01(mix64(h));
    const float n = sigma * sqrtf(-2.f * __logf(u1)) * __cosf(6.2831853f * u2);
    const float v = rintf((float)f0[(size_t)sy * pitch + sx] + n);
    f1[(size_t)y * pitch + x] = (uint8_t)fminf(fmaxf(v, 0.f), 255.f);
}

// ---------------------------------------------------------------- network
// Quarter-resolution input: out[n][c][y][x] = (mean of the 4x4 box - 128) / 8,
// c = 0 current frame, 1 reference frame.  (The box mean of smoothed noise
// has a standard deviation of about 8 grey levels, so inputs are O(1).)
__global__ void downsample_pair_kernel(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, float *out,
                                       int n) {
    const int H4 = H / 4, W4 = W / 4;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W4) return;
    int a = 0, b = 0;
    for (int dy = 0; dy < 4; ++dy)
        for (int dx = 0; dx < 4; ++dx) {
            a += cur[(size_t)(4 * y + dy) * pitch + 4 * x + dx];
            b += ref[(size_t)(4 * y + dy) * pitch + 4 * x + dx];
        }
    const size_t plane = (size_t)H4 * W4;
    out[(size_t)n * 2 * plane + (size_t)y * W4 + x] = (a * (1.f / 16.f) - 128.f) * 0.125f;
    out[(size_t)n * 2 * plane + plane + (size_t)y * W4 + x] = (b * (1.f / 16.f) - 128.f) * 0.125f;
}

// Forward 3x3 conv, dilation d, zero padding d (same size), + bias, optional
// ReLU.  One thread per pixel computes all COUT outputs; weights [co][ci][3][3]
// are staged in LDS as [ci][tap][co] for broadcast reads.
template <int COUT>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const float *__restrict__ in, int N, int Cin, int H, int W,
                                                       int d, const float *__restrict__ w,
                                                       const float *__restrict__ bias, int relu,
                                                       float *__restrict__ out) {
    extern __shared__ float wl[];  // Cin * 9 * COUT
    for (int i = threadIdx.x; i < Cin * 9 * COUT; i += blockDim.x) {
        const int co = i % COUT, t = (i / COUT) % 9, ci = i / (9 * COUT);
        wl[i] = w[(co * Cin + ci) * 9 + t];
    }
    __syncthreads();
    const long plane = (long)H * W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * plane) return;
    const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W, x = rem - y * W;
    float acc[COUT];
#pragma unroll
    for (int c = 0; c < COUT; ++c) acc[c] = bias[c];
    const float *src = in + (long)n * Cin * plane;
    for (int ci = 0; ci < Cin; ++ci) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
            const float v = src[(long)ci * plane + (long)yy * W + xx];
            const float *wr = wl + (ci * 9 + t) * COUT;
#pragma unroll
            for (int c = 0; c < COUT; ++c) acc[c] = __builtin_fmaf(v, wr[c], acc[c]);
        }
    }
    float *dst = out + (long)n * COUT * plane + rem;
#pragma unroll
    for (int c = 0; c < COUT; ++c) dst[(long)c * plane] = relu ? fmaxf(acc[c], 0.f) : acc[c];
}

// Backward to the input: din[n][ci][y][x] = sum_{co,t} w[co][ci][t] * g[n][co][y - dy_t][x - dx_t],
// where g = dout masked by the layer's ReLU (act > 0) when relu is set.
template <int CIN>
__global__ __launch_bounds__(256) void conv_bwd_data_kernel(const float *__restrict__ dout,
                                                            const float *__restrict__ act, int N, int Cout, int H,
                                                            int W, int d, const float *__restrict__ w, int relu,
                                                            float *__restrict__ din) {
    extern __shared__ float wl[];  // Cout * 9 * CIN as [co][tap][ci]
    for (int i = threadIdx.x; i < Cout * 9 * CIN; i += blockDim.x) {
        const int ci = i % CIN, t = (i / CIN) % 9, co = i / (9 * CIN);
        wl[i] = w[(co * CIN + ci) * 9 + t];
    }
    __syncthreads();
    const long plane = (long)H * W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * plane) return;
    const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W, x = rem - y * W;
    float acc[CIN];
#pragma unroll
    for (int c = 0; c < CIN; ++c) acc[c] = 0.f;
    const float *g = dout + (long)n * Cout * plane;
    const float *a = act + (long)n * Cout * plane;
    for (int co = 0; co < Cout; ++co) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y - (t / 3 - 1) * d, xx = x - (t % 3 - 1) * d;
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
            const long o = (long)co * plane + (long)yy * W + xx;
            const float v = (relu && a[o] <= 0.f) ? 0.f : g[o];
            const float *wr = wl + (co * 9 + t) * CIN;
#pragma unroll
            for (int c = 0; c < CIN; ++c) acc[c] = __builtin_fmaf(v, wr[c], acc[c]);
        }
    }
    float *dst = din + (long)n * CIN * plane + rem;
#pragma unroll
    for (int c = 0; c < CIN; ++c) dst[(long)c * plane] = acc[c];
}

// Backward to weights and bias, stage 1: workgroup b sums over its pixel range
// into part[b][Cout*Cin*9 + Cout].  Thread t owns (co, ci) pairs t, t+256, ...
// (at most 4) with their 9 taps; pairs with ci == 0 also sum the bias grad.
constexpr int WB_THREADS = 256;
__global__ __launch_bounds__(WB_THREADS) void conv_bwd_weight_kernel(const float *__restrict__ dout,
                                                                     const float *__restrict__ act,
                                                                     const float *__restrict__ in, int N, int Cin,
                                                                     int Cout, int H, int W, int d, int relu,
                                                                     long pix_per_wg, float *__restrict__ part) {
    const long plane = (long)H * W, total = N * plane;
    const int P = Cout * Cin;
    const int tid = threadIdx.x;
    float acc[4][9], bacc[4];
    int co_[4], ci_[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int pr = tid + WB_THREADS * j;
        co_[j] = pr < P ? pr / Cin : -1;
        ci_[j] = pr < P ? pr % Cin : 0;
        bacc[j] = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[j][t] = 0.f;
    }
    const long p0 = (long)blockIdx.x * pix_per_wg, p1 = min(total, p0 + pix_per_wg);
    for (long p = p0; p < p1; ++p) {
        const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W, x = rem - y * W;
        const float *src = in + (long)n * Cin * plane;
        const float *g = dout + (long)n * Cout * plane + rem;
        const float *a = act + (long)n * Cout * plane + rem;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (co_[j] < 0) continue;
            const float gv = (relu && a[(long)co_[j] * plane] <= 0.f) ? 0.f : g[(long)co_[j] * plane];
            if (gv == 0.f) continue;
            if (ci_[j] == 0) bacc[j] += gv;
            const float *s = src + (long)ci_[j] * plane;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) acc[j][t] = __builtin_fmaf(gv, s[(long)yy * W + xx], acc[j][t]);
            }
        }
    }
    float *o = part + (long)blockIdx.x * (P * 9 + Cout);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (co_[j] < 0) continue;
        const int pr = co_[j] * Cin + ci_[j];
#pragma unroll
        for (int t = 0; t < 9; ++t) o[pr * 9 + t] = acc[j][t];
        if (ci_[j] == 0) o[P * 9 + co_[j]] = bacc[j];
    }
}

// Stage 2: grad[i] = sum over workgroups of part[b][i] (fixed order).
__global__ void sum_parts_kernel(const float *__restrict__ part, int nparts, int len, float *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[(long)b * len + i];
    grad[i] = s;
}

// Huber loss (delta 1) of the quarter-resolution flow against block labels:
// target(y, x) = label[y/2][x/2] / 4 (dy, dx).  Writes dloss/dpred (mean over
// all components) and per-workgroup loss partial sums.
__global__ void huber_kernel(const float *__restrict__ pred, const int32_t *__restrict__ labels, int N, int H4,
                             int W4, float *__restrict__ grad, float *__restrict__ loss_part) {
    __shared__ float red[256];
    const long plane = (long)H4 * W4;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < N * plane) {
        const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W4, x = rem - y * W4;
        const int nbx = W4 / 2;
        const int32_t *lab = labels + (((long)n * (H4 / 2) + y / 2) * nbx + x / 2) * 3;
        const float inv = 1.0f / (float)(2 * N * plane);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const long o = ((long)n * 2 + c) * plane + rem;
            const float e = pred[o] - lab[c] * 0.25f;
            const float ae = fabsf(e);
            l += ae <= 1.f ? 0.5f * e * e : ae - 0.5f;
            grad[o] = fminf(fmaxf(e, -1.f), 1.f) * inv;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- inference
__device__ inline uint16_t bf16_rne(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// First layer (2 -> 32, dilation d) from the fp32 NCHW input to NHWC bf16,
// bias + ReLU (weights are the bf16-rounded training weights).
__global__ __launch_bounds__(256) void first_layer_kernel(const float *__restrict__ in, int N, int H, int W, int d,
                                                          const float *__restrict__ w, const float *__restrict__ bias,
                                                          uint16_t *__restrict__ out) {
    __shared__ float wl[2 * 9 * 32];
    for (int i = threadIdx.x; i < 2 * 9 * 32; i += blockDim.x) {
        const int co = i % 32, t = (i / 32) % 9, ci = i / (9 * 32);
        wl[i] = bf16f(bf16_rne(w[(co * 2 + ci) * 9 + t]));
    }
    __syncthreads();
    const long plane = (long)H * W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * plane) return;
    const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W, x = rem - y * W;
    float acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = bias[c];
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
            const float v = bf16f(bf16_rne(in[((long)n * 2 + ci) * plane + (long)yy * W + xx]));
#pragma unroll
            for (int c = 0; c < 32; ++c) acc[c] = __builtin_fmaf(v, wl[(ci * 9 + t) * 32 + c], acc[c]);
        }
    uint16_t *o = out + p * 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) o[c] = bf16_rne(fmaxf(acc[c], 0.f));
}

// Last layer (32 -> 2, dilation d) from NHWC bf16 to the fp32 flow [n][2][H][W], no ReLU.
__global__ __launch_bounds__(256) void last_layer_kernel(const uint16_t *__restrict__ in, int N, int H, int W, int d,
                                                         const float *__restrict__ w, const float *__restrict__ bias,
                                                         float *__restrict__ flow) {
    __shared__ float wl[32 * 9 * 2];
    for (int i = threadIdx.x; i < 32 * 9 * 2; i += blockDim.x) {
        const int co = i % 2, t = (i / 2) % 9, ci = i / 18;
        wl[i] = bf16f(bf16_rne(w[(co * 32 + ci) * 9 + t]));
    }
    __syncthreads();
    const long plane = (long)H * W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * plane) return;
    const int n = (int)(p / plane), rem = (int)(p - n * plane), y = rem / W, x = rem - y * W;
    float a0 = bias[0], a1 = bias[1];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
        const uint4 *src = (const uint4 *)(in + ((long)n * plane + (long)yy * W + xx) * 32);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 u = src[q];
            const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const int ci = q * 8 + h;
                const float v = bf16f((uint16_t)(wd[h >> 1] >> (16 * (h & 1))));
                a0 = __builtin_fmaf(v, wl[(ci * 9 + t) * 2 + 0], a0);
                a1 = __builtin_fmaf(v, wl[(ci * 9 + t) * 2 + 1], a1);
            }
        }
    }
    flow[((long)n * 2 + 0) * plane + rem] = a0;
    flow[((long)n * 2 + 1) * plane + rem] = a1;
}

// Block proposals: mean of the block's 2x2 quarter-resolution flow, x4,
// rounded, clamped to [-R, R].  prop[blk] = (dy, dx).
__global__ void proposals_kernel(const float *__restrict__ flow, int H4, int W4, int R, int32_t *__restrict__ prop) {
    const int nbx = W4 / 2, nby = H4 / 2;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbx * nby) return;
    const int by = b / nbx, bx = b - by * nbx;
    const long plane = (long)H4 * W4;
    for (int c = 0; c < 2; ++c) {
        const float *f = flow + c * plane + (long)(2 * by) * W4 + 2 * bx;
        const float m = (f[0] + f[1] + f[W4] + f[W4 + 1]) * 0.25f * 4.f;
        prop[b * 2 + c] = clampi((int)rintf(m), -R, R);
    }
}

// SAD of the 8x8 block (by, bx) of cur against ref at offset (dy, dx), ref clamped.
__device__ inline int sad8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int by, int bx, int dy,
                           int dx) {
    int s = 0;
    const int y0 = by * 8, x0 = bx * 8;
    for (int py = 0; py < 8; ++py) {
        const uint8_t *c = cur + (size_t)(y0 + py) * pitch + x0;
        const uint8_t *r = ref + (size_t)clampi(y0 + py + dy, 0, H - 1) * pitch;
#pragma unroll
        for (int px = 0; px < 8; ++px) s += abs((int)c[px] - (int)r[clampi(x0 + px + dx, 0, W - 1)]);
    }
    return s;
}

// Refinement: 64 threads per block.  Lanes 0..24 evaluate the +-2 window
// around the block's proposal (raster order), lanes 25..32 the 8 neighbours'
// proposals.  The refined result is the first minimum of the window; the
// block falls back to full search when that minimum lies on the window's
// border or exceeds 1.1 x the smallest SAD found at the neighbours' proposals.
__global__ __launch_bounds__(256) void refine_kernel(const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref,
                                                     int H, int W, int pitch, int R, const int32_t *__restrict__ prop,
                                                     int32_t *__restrict__ mv_sad, int32_t *__restrict__ fb_list,
                                                     int32_t *__restrict__ fb_count, float slack) {
    const int nbx = W / 8, nby = H / 8;
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const bool active = b < nbx * nby;
    const int by = active ? b / nbx : 0, bx = active ? b - by * nbx : 0;
    const int py = prop[(active ? b : 0) * 2], px = prop[(active ? b : 0) * 2 + 1];
    uint32_t key = 0xffffffffu;  // (sad << 8) | window index, window lanes only
    int nb_sad = 0x7fffffff;
    if (active && lane < 25) {
        const int dy = clampi(py + lane / 5 - 2, -R, R), dx = clampi(px + lane % 5 - 2, -R, R);
        key = ((uint32_t)sad8(cur, ref, H, W, pitch, by, bx, dy, dx) << 8) | (uint32_t)lane;
    } else if (active && lane < 33) {
        const int k = lane - 25;                  // neighbours in raster order, skipping the centre
        const int o = k < 4 ? k : k + 1;
        const int ny = by + o / 3 - 1, nx = bx + o % 3 - 1;
        if (ny >= 0 && ny < nby && nx >= 0 && nx < nbx) {
            const int nb = ny * nbx + nx;
            nb_sad = sad8(cur, ref, H, W, pitch, by, bx, prop[nb * 2], prop[nb * 2 + 1]);
        }
    }
    for (int s = 32; s > 0; s >>= 1) {
        key = min(key, (uint32_t)__shfl_xor((int)key, s, 64));
        nb_sad = min(nb_sad, __shfl_xor(nb_sad, s, 64));
    }
    if (!active || lane != 0) return;
    const int idx = key & 0xff, sad = (int)(key >> 8);
    const int dy = clampi(py + idx / 5 - 2, -R, R), dx = clampi(px + idx % 5 - 2, -R, R);
    const bool border = (idx / 5 == 0 || idx / 5 == 4 || idx % 5 == 0 || idx % 5 == 4);
    const bool worse = (float)sad > slack * (float)min(nb_sad, sad);
    int32_t *o = mv_sad + (size_t)b * 3;
    o[0] = dy;
    o[1] = dx;
    o[2] = sad;
    if (border || worse) fb_list[atomicAdd(fb_count, 1)] = b;
}

__global__ void zero_i32_kernel(int32_t *p) { *p = 0; }

}  // namespace

extern "C" {

umi_status umi_synth_frame0(uint8_t *f0, int H, int W, int pitch, uint64_t seed, umi_stream s) {
    if (!f0 || H < 1 || W < 1 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_synth_frame0: bad arguments");
    hipLaunchKernelGGL(synth_frame0_kernel, dim3(ceil_div(W, 256), H), dim3(256), 0, as_stream(s), f0, H, W, pitch, seed);
    return launch_check("synth_frame0_kernel");
}

umi_status umi_synth_warp(const uint8_t *f0, uint8_t *f1, int H, int W, int pitch, const float *bumps, float sigma,
                          uint64_t seed, umi_stream s) {
    if (!f0 || !f1 || !bumps || H < 1 || W < 1 || pitch < W)
        return set_error(UMI_ERR_INVALID, "umi_synth_warp: bad arguments");
    hipLaunchKernelGGL(synth_warp_kernel, dim3(ceil_div(W, 256), H), dim3(256), 0, as_stream(s), f0, f1, H, W, pitch,
                       (const Bump *)bumps, sigma, seed);
    return launch_check("synth_warp_kernel");
}

umi_status umi_cnn_downsample_pair(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, float *out, int n,
                                   umi_stream s) {
    if (!cur || !ref || !out || H < 4 || W < 4 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_cnn_downsample_pair");
    hipLaunchKernelGGL(downsample_pair_kernel, dim3(ceil_div(W / 4, 256), H / 4), dim3(256), 0, as_stream(s), cur, ref,
                       H, W, pitch, out, n);
    return launch_check("downsample_pair_kernel");
}

umi_status umi_cnn_conv_fwd(const float *in, int N, int Cin, int Cout, int H, int W, int dil, const float *w,
                            const float *bias, int relu, float *out, umi_stream s) {
    if (!in || !w || !bias || !out || N < 1 || Cin < 1 || H < 1 || W < 1 || dil < 1)
        return set_error(UMI_ERR_INVALID, "umi_cnn_conv_fwd: bad arguments");
    const long total = (long)N * H * W;
    const size_t lds = sizeof(float) * Cin * 9 * Cout;
    if (lds > 64 * 1024) return set_error(UMI_ERR_UNSUPPORTED, "umi_cnn_conv_fwd: Cin*Cout too large");
    dim3 grid((unsigned)((total + 255) / 256));
    if (Cout == 32)
        hipLaunchKernelGGL(conv_fwd_kernel<32>, grid, dim3(256), lds, as_stream(s), in, N, Cin, H, W, dil, w, bias, relu, out);
    else if (Cout == 2)
        hipLaunchKernelGGL(conv_fwd_kernel<2>, grid, dim3(256), lds, as_stream(s), in, N, Cin, H, W, dil, w, bias, relu, out);
    else
        return set_error(UMI_ERR_UNSUPPORTED, "umi_cnn_conv_fwd: Cout %d (32 or 2)", Cout);
    return launch_check("conv_fwd_kernel");
}

umi_status umi_cnn_conv_bwd_data(const float *dout, const float *act, int N, int Cin, int Cout, int H, int W, int dil,
                                 const float *w, int relu, float *din, umi_stream s) {
    if (!dout || !act || !w || !din || N < 1 || H < 1 || W < 1 || dil < 1)
        return set_error(UMI_ERR_INVALID, "umi_cnn_conv_bwd_data: bad arguments");
    if (Cin != 32) return set_error(UMI_ERR_UNSUPPORTED, "umi_cnn_conv_bwd_data: Cin %d (32)", Cin);
    const long total = (long)N * H * W;
    const size_t lds = sizeof(float) * Cout * 9 * 32;
    hipLaunchKernelGGL(conv_bwd_data_kernel<32>, dim3((unsigned)((total + 255) / 256)), dim3(256), lds, as_stream(s),
                       dout, act, N, Cout, H, W, dil, w, relu, din);
    return launch_check("conv_bwd_data_kernel");
}

umi_status umi_cnn_conv_bwd_weight_workspace(int N, int Cin, int Cout, int H, int W, size_t *bytes) {
    if (!bytes) return set_error(UMI_ERR_INVALID, "umi_cnn_conv_bwd_weight_workspace: null pointer");
    const long total = (long)N * H * W;
    const int nwg = (int)std::min<long>(1024, (total + 255) / 256);
    *bytes = sizeof(float) * (size_t)nwg * (Cout * Cin * 9 + Cout);
    return UMI_OK;
}

umi_status umi_cnn_conv_bwd_weight(const float *dout, const float *act, const float *in, int N, int Cin, int Cout,
                                   int H, int W, int dil, int relu, float *dw_db, void *workspace,
                                   size_t workspace_bytes, umi_stream s) {
    if (!dout || !act || !in || !dw_db || !workspace) return set_error(UMI_ERR_INVALID, "umi_cnn_conv_bwd_weight: null");
    if (Cout * Cin > 4 * WB_THREADS) return set_error(UMI_ERR_UNSUPPORTED, "umi_cnn_conv_bwd_weight: Cout*Cin > 1024");
    const long total = (long)N * H * W;
    const int nwg = (int)std::min<long>(1024, (total + 255) / 256);
    const int len = Cout * Cin * 9 + Cout;
    if (workspace_bytes < sizeof(float) * (size_t)nwg * len)
        return set_error(UMI_ERR_INVALID, "umi_cnn_conv_bwd_weight: workspace too small");
    const long per = (total + nwg - 1) / nwg;
    hipStream_t st = as_stream(s);
    hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(nwg), dim3(WB_THREADS), 0, st, dout, act, in, N, Cin, Cout, H, W, dil,
                       relu, per, (float *)workspace);
    UMI_TRY(launch_check("conv_bwd_weight_kernel"));
    hipLaunchKernelGGL(sum_parts_kernel, dim3(ceil_div(len, 256)), dim3(256), 0, st, (const float *)workspace, nwg, len,
                       dw_db);
    return launch_check("sum_parts_kernel");
}

umi_status umi_cnn_huber(const float *pred, const int32_t *labels, int N, int H4, int W4, float *grad,
                         float *loss_parts, int *nparts, umi_stream s) {
    if (!pred || !labels || !grad || !loss_parts || !nparts) return set_error(UMI_ERR_INVALID, "umi_cnn_huber: null");
    const long total = (long)N * H4 * W4;
    *nparts = (int)((total + 255) / 256);
    hipLaunchKernelGGL(huber_kernel, dim3(*nparts), dim3(256), 0, as_stream(s), pred, labels, N, H4, W4, grad, loss_parts);
    return launch_check("huber_kernel");
}

umi_status umi_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                    int step, umi_stream s) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return set_error(UMI_ERR_INVALID, "umi_adam: bad arguments");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(s), w, g, m, v, n, lr, b1, b2, eps,
                       c1, c2);
    return launch_check("adam_kernel");
}

umi_status umi_cnn_first_layer(const float *in, int N, int H, int W, int dil, const float *w, const float *bias,
                               uint16_t *out_nhwc, umi_stream s) {
    if (!in || !w || !bias || !out_nhwc) return set_error(UMI_ERR_INVALID, "umi_cnn_first_layer: null");
    const long total = (long)N * H * W;
    hipLaunchKernelGGL(first_layer_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, as_stream(s), in, N, H,
                       W, dil, w, bias, out_nhwc);
    return launch_check("first_layer_kernel");
}

umi_status umi_cnn_last_layer(const uint16_t *in_nhwc, int N, int H, int W, int dil, const float *w, const float *bias,
                              float *flow, umi_stream s) {
    if (!in_nhwc || !w || !bias || !flow) return set_error(UMI_ERR_INVALID, "umi_cnn_last_layer: null");
    const long total = (long)N * H * W;
    hipLaunchKernelGGL(last_layer_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, as_stream(s), in_nhwc, N,
                       H, W, dil, w, bias, flow);
    return launch_check("last_layer_kernel");
}

umi_status umi_me_proposals(const float *flow, int H4, int W4, int R, int32_t *prop, umi_stream s) {
    if (!flow || !prop) return set_error(UMI_ERR_INVALID, "umi_me_proposals: null");
    const int nb = (H4 / 2) * (W4 / 2);
    hipLaunchKernelGGL(proposals_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, as_stream(s), flow, H4, W4, R, prop);
    return launch_check("proposals_kernel");
}

umi_status umi_me_refine(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R, const int32_t *prop,
                         float slack, int32_t *mv_sad, int32_t *fb_list, int32_t *fb_count, umi_stream s) {
    if (!cur || !ref || !prop || !mv_sad || !fb_list || !fb_count) return set_error(UMI_ERR_INVALID, "umi_me_refine: null");
    if (H < 8 || W < 8 || pitch < W || R < 2) return set_error(UMI_ERR_INVALID, "umi_me_refine: bad frame");
    const int nb = (H / 8) * (W / 8);
    hipStream_t st = as_stream(s);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, st, fb_count);
    UMI_TRY(launch_check("zero_i32_kernel"));
    hipLaunchKernelGGL(refine_kernel, dim3(ceil_div(nb, 4)), dim3(256), 0, st, cur, ref, H, W, pitch, R, prop, mv_sad,
                       fb_list, fb_count, slack);
    return launch_check("refine_kernel");
}

}  // extern "C"
