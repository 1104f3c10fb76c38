// Learned motion proposals for full-search block ME (gfx950).
//
// No reference counterpart: johnjohnlin/UMI has only the exhaustive search
// (main/full_me.cu).  This file holds
//  * synthetic frame pairs: smoothed noise, warped by a smooth motion field
//    (sum of 4 Gaussian-weighted affine bumps, components clamped to +-32)
//    plus N(0, sigma^2) noise, from a counter-based hash (deterministic);
//  * the proposal network's training kernels in fp32 (forward, backward to
//    inputs and to weights/biases, Huber loss, Adam): NCHW, 3x3 dilated;
//  * inference: downsampling, the four layers on bf16 MFMA, block
//    proposals;
//  * the exact SAD refinement of the proposals (seed, +-2 descent,
//    propagation rounds, fallback check) and the fallback list.
#include "common.h"
#include "stage.h"
#include "wave.h"

#include <algorithm>

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
// Two values rounded to bf16 (RNE) by one v_cvt_pk_bf16_f32, packed low = a;
// for finite inputs the same bits as bf16_rne.
__device__ inline uint32_t bf16x2_rne(float a, float b) {
    typedef __attribute__((ext_vector_type(2))) float f32x2;
    typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

// Inference layers of the proposal network on v_mfma_f32_32x32x16_bf16,
// out[co][pixel] = sum_k W[co][k] x in[k][pixel]:
//   CNN_FIRST  2 -> 32: fp32 NCHW input, k = 2 t + ci over the 9 taps t (18
//              of 32 k, two chunks of 16), bf16 NHWC output, ReLU;
//   CNN_MID   32 -> 32: bf16 NHWC input, chunk 2 t + h = tap t, channels
//              16 h .. 16 h + 15 (18 chunks), bf16 NHWC output, ReLU;
//   CNN_LAST  32 -> 2:  as CNN_MID with output rows 2..31 zero, fp32 NCHW out.
// Operand layout (per 16-k chunk): lane l holds A row l & 31 (output
// channel) and B column l & 31 (pixel), each for k = 8 (l >> 5) .. + 7;
// accumulator r of lane l is (row 8 (r >> 2) + 4 (l >> 5) + (r & 3), column
// l & 31).  Weights and inputs are rounded to bf16 (RNE); zero padding d.
// Each wave walks 32-pixel groups (pixels in N, y, x order).
enum { CNN_FIRST = 0, CNN_MID = 1, CNN_LAST = 2 };

// The A operands, packed once per layer: wp[c * 64 + lane] (16 B each).
template <int MODE>
__global__ void cnn_pack_kernel(const float *__restrict__ w, bf16x8 *__restrict__ wp) {
    constexpr int NCH = MODE == CNN_FIRST ? 2 : 18;
    constexpr int COUT = MODE == CNN_LAST ? 2 : 32;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NCH * 64) return;
    const int c = i / 64, lane = i % 64, lr = lane & 31, lh = lane >> 5;
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float x = 0.f;
        if (lr < COUT) {
            if constexpr (MODE == CNN_FIRST) {
                const int k = 16 * c + 8 * lh + j;
                if (k < 18) x = w[(lr * 2 + (k & 1)) * 9 + (k >> 1)];
            } else {
                x = w[(lr * 32 + 16 * (c & 1) + 8 * lh + j) * 9 + (c >> 1)];
            }
        }
        v[j] = bf16_rne(x);
    }
    wp[i] = __builtin_bit_cast(bf16x8, v);
}

template <int MODE>
__global__ __launch_bounds__(256) void cnn_layer_kernel(const void *__restrict__ in_, int N, int H, int W, int d,
                                                        const bf16x8 *__restrict__ wp,
                                                        const float *__restrict__ bias, void *__restrict__ out_) {
    constexpr int NCH = MODE == CNN_FIRST ? 2 : 18;
    constexpr int COUT = MODE == CNN_LAST ? 2 : 32;
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5;
    bf16x8 A[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) A[c] = wp[c * 64 + lane];
    float bv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = 8 * (r >> 2) + 4 * lh + (r & 3);
        bv[r] = m < COUT ? bias[m] : 0.f;
    }
    const long plane = (long)H * W, npx = (long)N * plane;
    const long ngroups = (npx + 31) / 32;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long g = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < ngroups; g += nwaves) {
        const long p = g * 32 + lr;
        const bool valid = p < npx;
        const int n = valid ? (int)(p / plane) : 0;
        const int rem = valid ? (int)(p - n * plane) : 0, y = rem / W, x = rem - y * W;
        bf16x8 B[NCH];
        if constexpr (MODE == CNN_FIRST) {
            const float *in = (const float *)in_;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                uint16_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = 16 * c + 8 * lh + j, t = k >> 1;
                    const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
                    float val = 0.f;
                    if (k < 18 && valid && yy >= 0 && yy < H && xx >= 0 && xx < W)
                        val = in[((long)n * 2 + (k & 1)) * plane + (long)yy * W + xx];
                    v[j] = bf16_rne(val);
                }
                B[c] = __builtin_bit_cast(bf16x8, v);
            }
        } else {
            const uint16_t *in = (const uint16_t *)in_;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
                u32x4 v0 = {0, 0, 0, 0}, v1 = {0, 0, 0, 0};
                if (valid && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const u32x4 *src = (const u32x4 *)(in + (((long)n * plane + (long)yy * W + xx) * 32 + 8 * lh));
                    v0 = src[0];
                    v1 = src[2];  // channels 16 + 8 lh ..
                }
                B[2 * t] = __builtin_bit_cast(bf16x8, v0);
                B[2 * t + 1] = __builtin_bit_cast(bf16x8, v1);
            }
        }
        f32x16 acc = {};
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[c], B[c], acc, 0, 0, 0);
        if (!valid) continue;
        if constexpr (MODE == CNN_LAST) {
            if (lh == 0) {
                float *flow = (float *)out_;
                flow[((long)n * 2 + 0) * plane + rem] = acc[0] + bv[0];
                flow[((long)n * 2 + 1) * plane + rem] = acc[1] + bv[1];
            }
        } else {
            uint16_t *o = (uint16_t *)out_ + p * 32;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t h[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const float a0 = fmaxf(acc[4 * q + 2 * e] + bv[4 * q + 2 * e], 0.f);
                    const float a1 = fmaxf(acc[4 * q + 2 * e + 1] + bv[4 * q + 2 * e + 1], 0.f);
                    h[e] = bf16x2_rne(a0, a1);
                }
                *(uint2 *)(o + 8 * q + 4 * lh) = make_uint2(h[0], h[1]);
            }
        }
    }
}

// CNN_MID with the input staged in LDS: a workgroup computes an output tile
// of CT_H x CT_W pixels of one image; its (CT_H + 2d) x (CT_W + 2d) input
// pixels (zero outside the image) sit in LDS at 80 B per pixel (64 B of
// channels + 16 B that spread the 16-byte operand reads of 16 consecutive
// pixels over all banks).  Wave w computes rows 2w and 2w + 1 (32 pixels
// each).  Same operands, order and results as cnn_layer_kernel<CNN_MID>.
constexpr int CT_H = 8, CT_W = 32, CT_PX = 80;

// The stored form of a layer output: bias, ReLU, bf16 (RNE); lane (l & 31,
// l >> 5) writes channels 8 q + 4 (l >> 5) .. + 3, q = 0..3, of its pixel.
__device__ inline void put_relu_bf16(const f32x16 &acc, const float (&bv)[16], int lh, uint16_t *px) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t h[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float a0 = fmaxf(acc[4 * q + 2 * e] + bv[4 * q + 2 * e], 0.f);
            const float a1 = fmaxf(acc[4 * q + 2 * e + 1] + bv[4 * q + 2 * e + 1], 0.f);
            h[e] = bf16x2_rne(a0, a1);
        }
        *(uint2 *)(px + 8 * q + 4 * lh) = make_uint2(h[0], h[1]);
    }
}

// A operands (18 chunks) and bias of a 32 -> 32 layer, per lane.
__device__ inline void mid_operands(const bf16x8 *__restrict__ wp, const float *__restrict__ bias, bf16x8 (&A)[18],
                                    float (&bv)[16]) {
    const int lane = threadIdx.x & 63, lh = lane >> 5;
#pragma unroll
    for (int c = 0; c < 18; ++c) A[c] = wp[c * 64 + lane];
#pragma unroll
    for (int r = 0; r < 16; ++r) bv[r] = bias[8 * (r >> 2) + 4 * lh + (r & 3)];
}

// A 32 -> 32 layer of dilation d on the CT_H x CT_W output tile at (y0, x0)
// of an H x W image, from its (CT_H + 2d) x (CT_W + 2d) input pixels in LDS
// (hw = CT_W + 2d per row, CT_PX bytes per pixel, zero outside the image).
// Wave w computes rows 2w and 2w + 1; bf16 NHWC output.
__device__ inline void mid_from_tile(const uint8_t *tile, int hw, int d, int y0, int x0, int H, int W,
                                     const bf16x8 (&A)[18], const float (&bv)[16], uint16_t *__restrict__ out) {
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5, w = threadIdx.x >> 6;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int r = 2 * w + rr, y = y0 + r, x = x0 + lr;
        f32x16 acc = {};
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) {
            const int tr = r + d + (tp / 3 - 1) * d, tc = lr + d + (tp % 3 - 1) * d;
            const uint8_t *src = tile + (tr * hw + tc) * CT_PX + 16 * lh;
            const bf16x8 b0 = __builtin_bit_cast(bf16x8, *(const u32x4 *)src);
            const bf16x8 b1 = __builtin_bit_cast(bf16x8, *(const u32x4 *)(src + 32));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * tp], b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * tp + 1], b1, acc, 0, 0, 0);
        }
        if (y < H && x < W) put_relu_bf16(acc, bv, lh, out + ((long)y * W + x) * 32);
    }
}

// D > 0: the dilation as a compile-time constant (d is ignored); 0: d.
template <int D>
__global__ __launch_bounds__(256) void cnn_mid_tile_kernel(const uint16_t *__restrict__ in, int H, int W, int d_,
                                                           const bf16x8 *__restrict__ wp,
                                                           const float *__restrict__ bias,
                                                           uint16_t *__restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) uint8_t tile[];
    const int d = D > 0 ? D : d_;
    const int tiles_x = (W + CT_W - 1) / CT_W, tiles_y = (H + CT_H - 1) / CT_H;
    const int n = blockIdx.x / (tiles_x * tiles_y), t = blockIdx.x % (tiles_x * tiles_y);
    const int y0 = (t / tiles_x) * CT_H, x0 = (t % tiles_x) * CT_W;
    const int hw = CT_W + 2 * d, hh = CT_H + 2 * d, nchunk = hh * hw * 4;
    const long plane = (long)H * W;
    const uint16_t *img = in + (long)n * plane * 32;
    // stage: 16-byte chunks, 8 in flight per thread, zero outside the image
    for (int base = 0; base < nchunk; base += 8 * 256) {
        u32x4 v[8];
        int idx[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            idx[u] = min(base + u * 256 + (int)threadIdx.x, nchunk - 1);
            const int px = idx[u] >> 2, q = idx[u] & 3, r = px / hw, c = px - r * hw;
            const int yy = y0 - d + r, xx = x0 - d + c;
            const bool in_img = yy >= 0 && yy < H && xx >= 0 && xx < W;
            const u32x4 *src = (const u32x4 *)(img + ((long)clampi(yy, 0, H - 1) * W + clampi(xx, 0, W - 1)) * 32) + q;
            v[u] = *src;
            if (!in_img) v[u] = u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int px = idx[u] >> 2, q = idx[u] & 3;
            *(u32x4 *)(tile + px * CT_PX + 16 * q) = v[u];
        }
    }
    bf16x8 A[18];
    float bv[16];
    mid_operands(wp, bias, A, bv);
    __syncthreads();
    mid_from_tile(tile, hw, d, y0, x0, H, W, A, bv, out + (long)n * plane * 32);
}

// The first layer read straight from the frames: the workgroup computes its
// (CT_H + 2) x (CT_W + 2) input pixels (dilation 1) of the quarter-
// resolution pair, (mean of the 4x4 pixels - 128) / 8 of cur and of ref as
// in umi_cnn_downsample_pair, rounded to bf16 (zero outside the quarter-
// resolution image), into LDS, then runs cnn_layer_kernel<CNN_FIRST>'s
// operands and order on them: bf16 NHWC output, ReLU.  One image (N = 1).
__global__ __launch_bounds__(256) void cnn_first_frames_kernel(const uint8_t *__restrict__ cur,
                                                               const uint8_t *__restrict__ ref, int H, int W,
                                                               int pitch, const bf16x8 *__restrict__ wp,
                                                               const float *__restrict__ bias,
                                                               uint16_t *__restrict__ out) {
    constexpr int HW = CT_W + 2, HH = CT_H + 2;
    __shared__ uint16_t xin[HH * HW][2];
    const int H4 = H / 4, W4 = W / 4;
    const int tiles_x = (W4 + CT_W - 1) / CT_W;
    const int y0 = (blockIdx.x / tiles_x) * CT_H, x0 = (blockIdx.x % tiles_x) * CT_W;
    for (int i = threadIdx.x; i < HH * HW * 2; i += 256) {
        const int px = i >> 1, ch = i & 1, r = px / HW, c = px - r * HW;
        const int Y = y0 - 1 + r, X = x0 - 1 + c;
        float v = 0.f;
        if (Y >= 0 && Y < H4 && X >= 0 && X < W4) {
            const uint8_t *f = (ch ? ref : cur) + (size_t)(4 * Y) * pitch + 4 * X;
            int sum = 0;
#pragma unroll
            for (int dy = 0; dy < 4; ++dy)
#pragma unroll
                for (int dx = 0; dx < 4; ++dx) sum += f[(size_t)dy * pitch + dx];
            v = (sum * (1.f / 16.f) - 128.f) * 0.125f;
        }
        xin[px][ch] = bf16_rne(v);
    }
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5, w = threadIdx.x >> 6;
    const bf16x8 A0 = wp[lane], A1 = wp[64 + lane];
    float bv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bv[r] = bias[8 * (r >> 2) + 4 * lh + (r & 3)];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int r = 2 * w + rr, y = y0 + r, x = x0 + lr;
        bf16x8 B[2];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 16 * ch + 8 * lh + j, t = k >> 1;
                v[j] = k < 18 ? xin[(r + 1 + t / 3 - 1) * HW + lr + 1 + t % 3 - 1][k & 1] : (uint16_t)0;
            }
            B[ch] = __builtin_bit_cast(bf16x8, v);
        }
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B[1], acc, 0, 0, 0);
        if (y < H4 && x < W4) {
            uint16_t *o = out + ((long)y * W4 + x) * 32;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t h[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const float a0 = fmaxf(acc[4 * q + 2 * e] + bv[4 * q + 2 * e], 0.f);
                    const float a1 = fmaxf(acc[4 * q + 2 * e + 1] + bv[4 * q + 2 * e + 1], 0.f);
                    h[e] = bf16x2_rne(a0, a1);
                }
                *(uint2 *)(o + 8 * q + 4 * lh) = make_uint2(h[0], h[1]);
            }
        }
    }
}

// The first two layers in one launch, for the second layer's CT_H x CT_W
// output tile at (y0, x0): the workgroup computes the quarter-resolution
// inputs of the (CT_H + 2d + 2) x (CT_W + 2d + 2) pixels around it from the
// frames (as cnn_first_frames_kernel: dword row loads when A4, the frames
// and pitch being 4-byte aligned), the first layer's outputs on the
// (CT_H + 2d) x (CT_W + 2d) pixels into LDS (zero outside the image: the
// second layer's padding), and the second layer (dilation d) from there as
// cnn_mid_tile_kernel.  Same operands, order and bf16 roundings as those two
// launches; one image (N = 1).
template <bool A4, int D>
__global__ __launch_bounds__(256) void cnn_first_mid_frames_kernel(
    const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref, int H, int W, int pitch,
    const bf16x8 *__restrict__ wp0, const float *__restrict__ bias0, int d_, const bf16x8 *__restrict__ wp1,
    const float *__restrict__ bias1, uint16_t *__restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) uint8_t tile[];
    const int d = D > 0 ? D : d_;  // D > 0: compile-time dilation
    const int H4 = H / 4, W4 = W / 4;
    const int tiles_x = (W4 + CT_W - 1) / CT_W;
    const int y0 = (blockIdx.x / tiles_x) * CT_H, x0 = (blockIdx.x % tiles_x) * CT_W;
    const int hw = CT_W + 2 * d, hh = CT_H + 2 * d;  // first-layer outputs
    const int iw = hw + 2, ih = hh + 2;              // their inputs
    uint16_t(*xin)[2] = reinterpret_cast<uint16_t(*)[2]>(tile + hh * hw * CT_PX);
    // item i: quarter-resolution pixel i / 2 of cur (i even) or ref (i odd);
    // all row loads of a pass are issued before the first sum
    constexpr int U = 6;
    const int nitem = ih * iw * 2;
    for (int base = 0; base < nitem; base += U * 256) {
        uint32_t v[U][4];
        int idx[U];
        bool inside[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            idx[u] = min(base + u * 256 + (int)threadIdx.x, nitem - 1);
            const int px = idx[u] >> 1, r = px / iw, c = px - r * iw;
            const int Y = y0 - d - 1 + r, X = x0 - d - 1 + c;
            inside[u] = Y >= 0 && Y < H4 && X >= 0 && X < W4;
            const uint8_t *f =
                ((idx[u] & 1) ? ref : cur) + (size_t)(4 * clampi(Y, 0, H4 - 1)) * pitch + 4 * clampi(X, 0, W4 - 1);
#pragma unroll
            for (int dy = 0; dy < 4; ++dy) {
                const uint8_t *q = f + (size_t)dy * pitch;
                if constexpr (A4)
                    v[u][dy] = *reinterpret_cast<const uint32_t *>(q);
                else
                    v[u][dy] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint32_t sum = 0;
#pragma unroll
            for (int dy = 0; dy < 4; ++dy) sum = __builtin_amdgcn_sad_u8(v[u][dy], 0u, sum);  // byte sum
            const float val = inside[u] ? ((int)sum * (1.f / 16.f) - 128.f) * 0.125f : 0.f;
            xin[idx[u] >> 1][idx[u] & 1] = bf16_rne(val);
        }
    }
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5, w = threadIdx.x >> 6;
    const bf16x8 A0 = wp0[lane], A1 = wp0[64 + lane];
    float b0v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) b0v[r] = bias0[8 * (r >> 2) + 4 * lh + (r & 3)];
    __syncthreads();
    // first layer: 32-pixel groups of the hh x hw region in raster order;
    // lanes past its end repeat its last pixel (same value, same slot)
    const int npx = hh * hw;
    for (int g = w; g < (npx + 31) / 32; g += 4) {
        const int p = min(32 * g + lr, npx - 1), r = p / hw, c = p - r * hw;
        bf16x8 B[2];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 16 * ch + 8 * lh + j, t = k >> 1;
                v[j] = k < 18 ? xin[(r + t / 3) * iw + c + t % 3][k & 1] : (uint16_t)0;
            }
            B[ch] = __builtin_bit_cast(bf16x8, v);
        }
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B[1], acc, 0, 0, 0);
        const int Y = y0 - d + r, X = x0 - d + c;
        uint16_t *dst = reinterpret_cast<uint16_t *>(tile + p * CT_PX);
        if (Y >= 0 && Y < H4 && X >= 0 && X < W4) {
            put_relu_bf16(acc, b0v, lh, dst);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) *(uint2 *)(dst + 8 * q + 4 * lh) = make_uint2(0, 0);
        }
    }
    bf16x8 A[18];
    float bv[16];
    mid_operands(wp1, bias1, A, bv);
    __syncthreads();
    mid_from_tile(tile, hw, d, y0, x0, H4, W4, A, bv, out);
}

// The last layer with the block proposals: a 32-pixel group is 2 rows x 16
// columns (lane lr: row lr / 16, column lr % 16), so each 2x2 quad of the
// flow lies in one group.  Per pixel the operands and order of
// cnn_layer_kernel<CNN_LAST>; the flow is written as that kernel writes it,
// and, when prop is given, the proposal of block (Y / 2, X / 2) is
// round(4 x ((f(Y, X) + f(Y, X + 1)) + f(Y + 1, X)) + f(Y + 1, X + 1)) / 4),
// clamped to [-R, R], as umi_me_proposals.  One image (N = 1).
__global__ __launch_bounds__(256) void cnn_last_prop_kernel(const uint16_t *__restrict__ in, int H, int W, int d,
                                                            const bf16x8 *__restrict__ wp,
                                                            const float *__restrict__ bias, int R,
                                                            float *__restrict__ flow, int32_t *__restrict__ prop) {
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5;
    bf16x8 A[18];
#pragma unroll
    for (int c = 0; c < 18; ++c) A[c] = wp[c * 64 + lane];
    const float b0 = bias[0], b1 = bias[1];
    const int gx = (W + 15) / 16, gy = (H + 1) / 2, ngroups = gx * gy;
    const long plane = (long)H * W;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < ngroups; g += nwaves) {
        const int y = (g / gx) * 2 + (lr >> 4), x = (g % gx) * 16 + (lr & 15);
        const bool valid = y < H && x < W;
        bf16x8 B[18];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y + (t / 3 - 1) * d, xx = x + (t % 3 - 1) * d;
            u32x4 v0 = {0, 0, 0, 0}, v1 = {0, 0, 0, 0};
            if (valid && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const u32x4 *src = (const u32x4 *)(in + (((long)yy * W + xx) * 32 + 8 * lh));
                v0 = src[0];
                v1 = src[2];
            }
            B[2 * t] = __builtin_bit_cast(bf16x8, v0);
            B[2 * t + 1] = __builtin_bit_cast(bf16x8, v1);
        }
        f32x16 acc = {};
#pragma unroll
        for (int c = 0; c < 18; ++c) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[c], B[c], acc, 0, 0, 0);
        const float f0 = acc[0] + b0, f1 = acc[1] + b1;  // rows 0, 1 of lanes 0..31
        if (lh == 0 && valid) {
            flow[(long)y * W + x] = f0;
            flow[plane + (long)y * W + x] = f1;
        }
        if (prop) {
            // quad of lane 2q (row 0): +1 (column + 1), +16 (row + 1), +17
            const float r0 = __shfl(f0, lane + 1), r1 = __shfl(f1, lane + 1);
            const float d0 = __shfl(f0, lane + 16), d1 = __shfl(f1, lane + 16);
            const float e0 = __shfl(f0, lane + 17), e1 = __shfl(f1, lane + 17);
            const int X = x / 2, Y = y / 2;
            if (lh == 0 && lr < 16 && (lr & 1) == 0 && Y < H / 2 && X < W / 2) {
                const float m0 = (f0 + r0 + d0 + e0) * 0.25f * 4.f, m1 = (f1 + r1 + d1 + e1) * 0.25f * 4.f;
                int32_t *o = prop + ((long)Y * (W / 2) + X) * 2;
                o[0] = clampi((int)rintf(m0), -R, R);
                o[1] = clampi((int)rintf(m1), -R, R);
            }
        }
    }
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

// ---------------------------------------------------------------- refinement
// Work layout: a workgroup owns a tile of RT_Y x RT_X blocks and stages in LDS
// the cur tile and the ref region every vector in [-R, R]^2 can reach (ref
// coordinates clamped at fill time); 8 lanes per block then evaluate the
// block's candidate vectors from LDS.
constexpr int RT_X = 8, RT_Y = 4, RT_THREADS = 256;  // 32 blocks x 8 lanes

// Block state between refinement passes: motion vector, SAD, and whether the
// descent stopped at its iteration limit.
struct RefState {
    int dy, dx, sad, unconv;
};

struct RefineArgs {
    const uint8_t *cur, *ref;
    int H, W, pitch, R, iters;
    int tiles_x, ntiles;  // tile grid
    int stride4, rows;    // LDS ref region: rows x stride4 dwords
    int ref16, cur16;     // frame base and pitch are 16-byte aligned and W >= 16
};

// LDS geometry of a tile's ref region: rows 8 RT_Y + 2R; a row starts at the
// 16-byte boundary xs at or below x0 - R (x0 - R - xs <= 15) and reaches far
// enough that a four-dword read starting at or below any block's column for
// a vector in [-R, R] stays inside the row: byte 8 (RT_X - 1) + 2R + 15 + 15
// < 4 stride4; stride4 is a multiple of 4 (whole 16-byte chunks).
inline void region_shape(int R, int &stride4, int &rows) {
    stride4 = (8 * (RT_X - 1) + 2 * R + 31 + 16) / 16 * 4;
    rows = 8 * RT_Y + 2 * R;
}
// LDS of a tile: ref region and cur tile.
inline size_t region_lds_bytes(int R) {
    int s4, rows;
    region_shape(R, s4, rows);
    return (size_t)rows * s4 * 4 + 8 * RT_Y * 8 * RT_X;
}

// Workgroups of consecutive ids run on different XCDs; give each XCD a
// contiguous band of tiles (the grid is padded to a multiple of 8).
__device__ inline int tile_of_workgroup() {
    const int per = gridDim.x / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

__device__ inline int floor16(int x) { return x >= 0 ? x & ~15 : -((-x + 15) & ~15); }

// Stages the tile's ref region (columns from xs = floor16(x0 - R), rows from
// y0 - R, coordinates clamped to the frame) and cur tile.  Returns x0 - R - xs.
// Frames with 16-byte-aligned base and pitch (and W >= 16) load 16-byte
// chunks, all issued before the first LDS store (stage_rows16); others
// assemble clamped dwords.
__device__ inline int fill_tile(uint32_t *reg, uint32_t *ctile, const RefineArgs &a, int bx0, int by0) {
    const int xr = bx0 * 8 - a.R, y0 = by0 * 8 - a.R, xs = floor16(xr);
    const auto ref_row = [&](int r) { return a.ref + (size_t)clampi(y0 + r, 0, a.H - 1) * a.pitch; };
    // cur tile: 8 RT_Y rows of 8 RT_X bytes; rows past the frame belong to no block
    const auto cur_row = [&](int r) { return a.cur + (size_t)min(by0 * 8 + r, a.H - 1) * a.pitch; };
    if (a.ref16) {
        stage_rows16<4>(reg, a.rows, a.stride4 / 4, a.stride4 / 4, xs, a.W, threadIdx.x, RT_THREADS, ref_row);
    } else {
        stage_clamped<8>(reg, a.rows * a.stride4, threadIdx.x, RT_THREADS, a.W,
                         [&](int i, const uint8_t *&row, int &gx) {
                             const int r = i / a.stride4;
                             row = ref_row(r);
                             gx = xs + 4 * (i - r * a.stride4);
                         });
    }
    static_assert(8 * RT_Y * RT_X / 16 <= RT_THREADS, "one cur chunk per thread");
    if (a.cur16) {
        stage_rows16<1>(ctile, 8 * RT_Y, RT_X / 2, RT_X / 2, bx0 * 8, a.W, threadIdx.x, RT_THREADS, cur_row);
    } else {
        stage_clamped<2>(ctile, 8 * RT_Y * 2 * RT_X, threadIdx.x, RT_THREADS, a.W,
                         [&](int i, const uint8_t *&row, int &gx) {
                             const int r = i / (2 * RT_X);
                             row = cur_row(r);
                             gx = bx0 * 8 + 4 * (i - r * (2 * RT_X));
                         });
    }
    return xr - xs;
}

// SAD sources for one block's candidate vectors.  LdsSad reads the tile's
// staged ref region; GlobalSad reads ref directly (coordinates clamped),
// three aligned dwords per row for columns in [3, W - 12] (never outside the
// row), clamped bytes elsewhere.
struct LdsSad {
    const uint32_t *reg;
    int stride4, o0;  // o0: byte offset of the block's vector (-R, -R) in the region
    // All 24 reads are issued before the first SAD; even and odd rows sum in
    // two interleaved chains.
    __device__ uint32_t operator()(const uint32_t (&cw)[16], int R, int dy, int dx) const {
        const int o = o0 + (dy + R) * stride4 * 4 + dx + R;
        const uint32_t *p = reg + (o >> 2);
        const uint32_t sh = (uint32_t)(o & 3);
        uint32_t d[8][3];
#pragma unroll
        for (int py = 0; py < 8; ++py)
#pragma unroll
            for (int k = 0; k < 3; ++k) d[py][k] = p[py * stride4 + k];
        uint32_t s[2] = {0, 0};
#pragma unroll
        for (int py = 0; py < 8; ++py) {
            uint32_t &a = s[py & 1];
            a = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d[py][1], d[py][0], sh), cw[2 * py], a);
            a = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d[py][2], d[py][1], sh), cw[2 * py + 1], a);
        }
        return s[0] + s[1];
    }
    // SADs at (dy, dx0 + c), c = 0..4 (all within [-R, R]) from one pass over
    // 8 rows of 16 bytes: v_qsad_pk_u16_u8 gives the 4 horizontal offsets of
    // c = 0..3, two v_sad_u8 the fifth.  All 32 reads are issued first; even
    // and odd rows accumulate in two interleaved chains (each packed u16 SAD
    // stays below 2^15, so one 64-bit add merges them without carries between
    // fields).
    __device__ void row5(const uint32_t (&cw)[16], int R, int dy, int dx0, uint32_t (&s)[5]) const {
        const int o = o0 + (dy + R) * stride4 * 4 + dx0 + R;
        const uint32_t *p = reg + (o >> 2);
        const uint32_t sh = (uint32_t)(o & 3);
        uint32_t d[8][4];
#pragma unroll
        for (int py = 0; py < 8; ++py)
#pragma unroll
            for (int k = 0; k < 4; ++k) d[py][k] = p[py * stride4 + k];
        uint64_t acc[2] = {0, 0};
        uint32_t s4[2] = {0, 0};
#pragma unroll
        for (int py = 0; py < 8; ++py) {
            const int h = py & 1;
            const uint32_t a0 = __builtin_amdgcn_alignbyte(d[py][1], d[py][0], sh),
                           a1 = __builtin_amdgcn_alignbyte(d[py][2], d[py][1], sh),
                           a2 = __builtin_amdgcn_alignbyte(d[py][3], d[py][2], sh);
            acc[h] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)a0 | ((uint64_t)a1 << 32), cw[2 * py], acc[h]);
            acc[h] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)a1 | ((uint64_t)a2 << 32), cw[2 * py + 1], acc[h]);
            s4[h] = __builtin_amdgcn_sad_u8(a1, cw[2 * py], s4[h]);
            s4[h] = __builtin_amdgcn_sad_u8(a2, cw[2 * py + 1], s4[h]);
        }
        const uint64_t t = acc[0] + acc[1];
#pragma unroll
        for (int c = 0; c < 4; ++c) s[c] = (uint32_t)(t >> (16 * c)) & 0xffffu;
        s[4] = s4[0] + s4[1];
    }
};
__device__ inline LdsSad lds_sad(const uint32_t *reg, const RefineArgs &a, int xoff, int ty, int tx) {
    return {reg, a.stride4, 8 * ty * a.stride4 * 4 + 8 * tx + xoff};
}

struct GlobalSad {
    const uint8_t *ref;
    int H, W, pitch, y0, x0;
    __device__ uint32_t operator()(const uint32_t (&cw)[16], int /*R*/, int dy, int dx) const {
        const int x = x0 + dx;
        uint32_t s = 0;
        if (x >= 3 && x + 12 <= W) {
            uint32_t d[8][3], sh[8];
#pragma unroll
            for (int py = 0; py < 8; ++py) {  // all 24 loads before the first use
                const uint8_t *p = ref + (size_t)clampi(y0 + dy + py, 0, H - 1) * pitch + x;
                sh[py] = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 3);
                const uint32_t *q = reinterpret_cast<const uint32_t *>(p - sh[py]);
                d[py][0] = q[0];
                d[py][1] = q[1];
                d[py][2] = q[2];
            }
#pragma unroll
            for (int py = 0; py < 8; ++py) {
                s = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d[py][1], d[py][0], sh[py]), cw[2 * py], s);
                s = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d[py][2], d[py][1], sh[py]), cw[2 * py + 1],
                                            s);
            }
        } else {
            for (int py = 0; py < 8; ++py) {
                const uint8_t *r = ref + (size_t)clampi(y0 + dy + py, 0, H - 1) * pitch;
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    lo |= (uint32_t)r[clampi(x + i, 0, W - 1)] << (8 * i);
                    hi |= (uint32_t)r[clampi(x + 4 + i, 0, W - 1)] << (8 * i);
                }
                s = __builtin_amdgcn_sad_u8(lo, cw[2 * py], s);
                s = __builtin_amdgcn_sad_u8(hi, cw[2 * py + 1], s);
            }
        }
        return s;
    }
    // As LdsSad::row5; columns x in [3, W - 16] read four aligned dwords per
    // row (inside the row), other columns fall back to single positions.
    __device__ void row5(const uint32_t (&cw)[16], int R, int dy, int dx0, uint32_t (&s)[5]) const {
        const int x = x0 + dx0;
        if (x >= 3 && x + 16 <= W) {
            uint64_t acc = 0;
            uint32_t s4 = 0;
            uint32_t d[8][4], sh[8];
#pragma unroll
            for (int py = 0; py < 8; ++py) {  // all 32 loads before the first use
                const uint8_t *p = ref + (size_t)clampi(y0 + dy + py, 0, H - 1) * pitch + x;
                sh[py] = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 3);
                const uint32_t *q = reinterpret_cast<const uint32_t *>(p - sh[py]);
#pragma unroll
                for (int k = 0; k < 4; ++k) d[py][k] = q[k];
            }
#pragma unroll
            for (int py = 0; py < 8; ++py) {
                const uint32_t a0 = __builtin_amdgcn_alignbyte(d[py][1], d[py][0], sh[py]),
                               a1 = __builtin_amdgcn_alignbyte(d[py][2], d[py][1], sh[py]),
                               a2 = __builtin_amdgcn_alignbyte(d[py][3], d[py][2], sh[py]);
                acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)a0 | ((uint64_t)a1 << 32), cw[2 * py], acc);
                acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)a1 | ((uint64_t)a2 << 32), cw[2 * py + 1], acc);
                s4 = __builtin_amdgcn_sad_u8(a1, cw[2 * py], s4);
                s4 = __builtin_amdgcn_sad_u8(a2, cw[2 * py + 1], s4);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) s[c] = (uint32_t)(acc >> (16 * c)) & 0xffffu;
            s[4] = s4;
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) s[c] = (*this)(cw, R, dy, dx0 + c);
        }
    }
};

__device__ inline void tile_cur(uint32_t (&cw)[16], const uint32_t *ctile, int ty, int tx) {
#pragma unroll
    for (int py = 0; py < 8; ++py) {
        cw[2 * py] = ctile[(8 * ty + py) * 2 * RT_X + 2 * tx];
        cw[2 * py + 1] = ctile[(8 * ty + py) * 2 * RT_X + 2 * tx + 1];
    }
}

// Cur rows of the block at (y0, x0) from global memory.
__device__ inline void global_cur(uint32_t (&cw)[16], const uint8_t *cur, int pitch, int y0, int x0) {
#pragma unroll
    for (int py = 0; py < 8; ++py) {
        const uint8_t *r = cur + (size_t)(y0 + py) * pitch + x0;
        if ((reinterpret_cast<uintptr_t>(r) & 3) == 0) {
            cw[2 * py] = reinterpret_cast<const uint32_t *>(r)[0];
            cw[2 * py + 1] = reinterpret_cast<const uint32_t *>(r)[1];
        } else {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                lo |= (uint32_t)r[i] << (8 * i);
                hi |= (uint32_t)r[4 + i] << (8 * i);
            }
            cw[2 * py] = lo;
            cw[2 * py + 1] = hi;
        }
    }
}

// Minimum over the 8 lanes of a block (DPP, no LDS round trip).
__device__ inline uint32_t group_min(uint32_t k) { return min8_u32(k); }

// Iterated +-2 descent from (cy, cx).  Each step takes the 25 positions
// clamp(c + (i / 5 - 2, i % 5 - 2), -R, R), i = 0..24, and moves to the
// first minimum in i order; it stops when that minimum is within 1 of the
// centre in both axes, else repeats up to `iters` steps (unconv = 1 if it
// never stops).  Lane r < 5 evaluates window row r (its 5 columns in one
// pass when none of them is clamped).  The 8 lanes call it together.
template <class Sad>
__device__ inline RefState descend(const RefineArgs &a, const uint32_t (&cw)[16], const Sad &sad, int l, int cy,
                                   int cx) {
    RefState st{cy, cx, 0, 1};
    for (int t = 0; t < a.iters; ++t) {
        uint32_t key = 0xffffffffu;
        if (l < 5) {
            const int vy = clampi(cy + l - 2, -a.R, a.R);
            uint32_t sv[5];
            if (cx - 2 >= -a.R && cx + 2 <= a.R) {
                sad.row5(cw, a.R, vy, cx - 2, sv);
            } else {
#pragma unroll
                for (int c = 0; c < 5; ++c) sv[c] = sad(cw, a.R, vy, clampi(cx + c - 2, -a.R, a.R));
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) key = min(key, (sv[c] << 5) | (uint32_t)(5 * l + c));
        }
        key = group_min(key);
        const int i = (int)(key & 31);
        const int ny = clampi(cy + i / 5 - 2, -a.R, a.R), nx = clampi(cx + i % 5 - 2, -a.R, a.R);
        const bool moved = abs(ny - cy) == 2 || abs(nx - cx) == 2;
        cy = ny;
        cx = nx;
        st = {cy, cx, (int)(key >> 5), 1};
        if (!moved) {
            st.unconv = 0;
            break;
        }
    }
    return st;
}

// Neighbour o (0..8, raster order of the 3x3 neighbourhood, 4 = the block)
// at distance d.  nb is always a valid index (the block itself when the
// neighbour is outside the frame), so that loads through it can be issued
// unconditionally.
__device__ inline bool neighbour(int by, int bx, int o, int d, int nby, int nbx, int &nb) {
    const int ny = by + (o / 3 - 1) * d, nx = bx + (o % 3 - 1) * d;
    const bool in = ny >= 0 && ny < nby && nx >= 0 && nx < nbx;
    nb = in ? ny * nbx + nx : by * nbx + bx;
    return in;
}

// The 8 neighbours in raster order, skipping the block itself.
__device__ inline int ring(int k) { return k < 4 ? k : k + 1; }

// Candidate c is evaluated by lane c % 8 in slot c / 8; after the group
// minimum every lane fetches the winner's vector from that lane.
template <int NS>
__device__ inline void winner(uint32_t key, const int (&my)[NS], const int (&mx)[NS], int &cy, int &cx) {
    const int c = (int)(key & 31), pick = c >> 3;
    int vy = my[0], vx = mx[0];
#pragma unroll
    for (int j = 1; j < NS; ++j)
        if (pick == j) {
            vy = my[j];
            vx = mx[j];
        }
    cy = __shfl(vy, c & 7, 8);
    cx = __shfl(vx, c & 7, 8);
}

// Distances of in-tile round p: 1 and local_jump(p) (4, 2, 1: none);
// of global round p: 1, global_jump(p, 0) and global_jump(p, 1).
__device__ inline int local_jump(int p) { return 4 >> (p % 3); }
__host__ __device__ inline int global_jump(int p, int i) {
    constexpr int kJ[4][2] = {{8, 32}, {4, 16}, {2, 8}, {4, 32}};
    return kJ[p & 3][i];
}

// The block a thread works on: 8 lanes per block, RT_Y x RT_X blocks per tile.
struct BlockCtx {
    int by, bx, ty, tx, l, b;  // b: block index (0 for blocks past the frame)
    bool valid;
};
__device__ inline BlockCtx block_ctx(const RefineArgs &a, int tile) {
    BlockCtx c;
    const int g = threadIdx.x >> 3;
    c.l = threadIdx.x & 7;
    c.ty = g / RT_X;
    c.tx = g % RT_X;
    c.by = (tile / a.tiles_x) * RT_Y + c.ty;
    c.bx = (tile % a.tiles_x) * RT_X + c.tx;
    c.valid = tile < a.ntiles && c.by < a.H / 8 && c.bx < a.W / 8;
    c.b = c.valid ? c.by * (a.W / 8) + c.bx : 0;
    return c;
}

// First minimum over a block's candidates (lane l: candidates l + 8 j,
// used when ok), skipping a candidate equal to `skip_y, skip_x` when
// skip is set (its SAD equals the block's own and never wins).
template <int NS, class Sad>
__device__ inline uint32_t candidates_min(const RefineArgs &a, const uint32_t (&cw)[16], const Sad &sad, int l,
                                          const int (&my)[NS], const int (&mx)[NS], const bool (&ok)[NS], bool skip,
                                          int skip_y, int skip_x) {
    uint32_t key = 0xffffffffu;
#pragma unroll
    for (int j = 0; j < NS; ++j)
        if (ok[j] && !(skip && my[j] == skip_y && mx[j] == skip_x))
            key = min(key, (sad(cw, a.R, my[j], mx[j]) << 5) | (uint32_t)(l + 8 * j));
    return group_min(key);
}

// Seed of one tile: stage the ref region; per block the first minimum over
// its own and its 8 neighbours' proposals (candidates 0..8), the descent,
// then up to local_rounds in-tile rounds (Jacobi over the tile's states in
// LDS; candidates are the neighbours at distance 1 and local_jump(p) inside
// the tile; the tile stops after a round without a change).  All threads of
// the workgroup call it (it contains barriers); it returns the block's state
// (every lane of the block holds it) and the region offset in xoff.
__device__ inline RefState seed_tile(const RefineArgs &a, int local_rounds, const int32_t *__restrict__ prop,
                                     uint32_t *reg, uint32_t *ctile, RefState *tstate, const BlockCtx &c, int tile,
                                     int &xoff) {
    const int nby = a.H / 8, nbx = a.W / 8;
    int my[2], mx[2];
    bool ok[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {  // proposal loads first: in flight with the fill
        const int k = c.l + 8 * j;
        int nb;
        const bool in = neighbour(c.by, c.bx, k == 0 ? 4 : ring(min(k, 8) - 1), 1, nby, nbx, nb);
        ok[j] = c.valid && k < 9 && in;
        const int i = ok[j] ? nb : c.b;
        my[j] = clampi(prop[2 * i], -a.R, a.R);
        mx[j] = clampi(prop[2 * i + 1], -a.R, a.R);
    }
    xoff = fill_tile(reg, ctile, a, (tile % a.tiles_x) * RT_X, (tile / a.tiles_x) * RT_Y);
    __syncthreads();
    const LdsSad sad = lds_sad(reg, a, xoff, c.ty, c.tx);
    uint32_t cw[16];
    tile_cur(cw, ctile, c.ty, c.tx);
    RefState st{0, 0, 0, 0};
    if (c.valid) {
        int cy, cx;
        winner<2>(candidates_min<2>(a, cw, sad, c.l, my, mx, ok, false, 0, 0), my, mx, cy, cx);
        st = descend(a, cw, sad, c.l, cy, cx);
    }
    const int g = c.ty * RT_X + c.tx;
    if (c.l == 0) tstate[g] = st;
    __syncthreads();
    for (int p = 0; p < local_rounds; ++p) {
        const int k = local_jump(p);
#pragma unroll
        for (int j = 0; j < 2; ++j) {  // candidates from the tile's states before the round
            const int o = ring(c.l), d = j ? k : 1;
            const int ty = c.ty + (o / 3 - 1) * d, tx = c.tx + (o % 3 - 1) * d;
            ok[j] = c.valid && (j == 0 || k > 1) && ty >= 0 && ty < RT_Y && tx >= 0 && tx < RT_X &&
                    c.by - c.ty + ty < nby && c.bx - c.tx + tx < nbx;
            const RefState n = tstate[ok[j] ? ty * RT_X + tx : g];
            my[j] = n.dy;
            mx[j] = n.dx;
        }
        __syncthreads();  // all reads of the round precede its writes
        bool improved = false;
        if (c.valid) {
            const uint32_t key = candidates_min<2>(a, cw, sad, c.l, my, mx, ok, true, st.dy, st.dx);
            improved = (int)(key >> 5) < st.sad;  // no candidate: key >> 5 = 2^27 - 1
            if (improved) {
                int cy, cx;
                winner<2>(key, my, mx, cy, cx);
                st = descend(a, cw, sad, c.l, cy, cx);
            }
        }
        if (c.l == 0) tstate[g] = st;
        if (!__syncthreads_or(improved)) break;  // uniform
    }
    return st;
}

// Seed kernel: seed_tile per tile, the state to `out`.  Also clears the
// round flags and the fallback count.
__global__ __launch_bounds__(RT_THREADS) void refine_seed_kernel(RefineArgs a, int local_rounds,
                                                                 const int32_t *__restrict__ prop,
                                                                 RefState *__restrict__ out, int32_t *changed,
                                                                 int rounds, int32_t *fb_count) {
    extern __shared__ uint32_t lds[];
    __shared__ RefState tstate[RT_Y * RT_X];
    const int tile = tile_of_workgroup();
    if (blockIdx.x == 0) {
        if ((int)threadIdx.x < rounds) changed[threadIdx.x] = 0;
        if (threadIdx.x == 0) *fb_count = 0;
    }
    if (tile >= a.ntiles) return;  // whole workgroup
    const BlockCtx c = block_ctx(a, tile);
    int xoff;
    const RefState st = seed_tile(a, local_rounds, prop, lds, lds + a.rows * a.stride4, tstate, c, tile, xoff);
    if (c.valid && c.l == 0) out[c.b] = st;
}

// One global round p: each block tries the vectors of its 8 neighbours at
// distance 1, global_jump(p, 0) and global_jump(p, 1) (candidates 0..23,
// state `in`) and descends from the first minimum when it beats its own SAD;
// SADs from the tile's staged ref region.  A round whose predecessor changed
// nothing does nothing.
__global__ __launch_bounds__(RT_THREADS) void refine_round_kernel(RefineArgs a, int p, const RefState *__restrict__ in,
                                                                  RefState *__restrict__ out,
                                                                  const int32_t *prev_changed, int32_t *changed) {
    extern __shared__ uint32_t lds[];
    const int tile = tile_of_workgroup();
    if (tile >= a.ntiles) return;
    // read now, tested after the loads below are in flight (most rounds run)
    const int32_t run = prev_changed ? *prev_changed : 1;
    uint32_t *reg = lds, *ctile = lds + a.rows * a.stride4;
    const BlockCtx c = block_ctx(a, tile);
    const int nby = a.H / 8, nbx = a.W / 8;
    RefState st = in[c.b];
    int my[3], mx[3];
    bool ok[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {  // state loads first: in flight with the fill
        int nb;
        const bool inside = neighbour(c.by, c.bx, ring(c.l), j ? global_jump(p, j - 1) : 1, nby, nbx, nb);
        ok[j] = c.valid && inside;
        const RefState n = in[ok[j] ? nb : c.b];
        my[j] = n.dy;
        mx[j] = n.dx;
    }
    const int xoff = fill_tile(reg, ctile, a, (tile % a.tiles_x) * RT_X, (tile / a.tiles_x) * RT_Y);
    if (run == 0) return;  // whole workgroup
    __syncthreads();
    bool improved = false;
    if (c.valid) {
        const LdsSad sad = lds_sad(reg, a, xoff, c.ty, c.tx);
        uint32_t cw[16];
        tile_cur(cw, ctile, c.ty, c.tx);
        const uint32_t key = candidates_min<3>(a, cw, sad, c.l, my, mx, ok, true, st.dy, st.dx);
        improved = (int)(key >> 5) < st.sad;  // no candidate: key >> 5 = 2^27 - 1
        if (improved) {
            int cy, cx;
            winner<3>(key, my, mx, cy, cx);
            st = descend(a, cw, sad, c.l, cy, cx);
        }
        if (c.l == 0) out[c.b] = st;
    }
    if (__syncthreads_or(improved) && threadIdx.x == 0) *changed = 1;
}

// Writes lane 0's result and appends fallback blocks, one atomic per wave
// (the wave's fallback blocks take consecutive slots).  All lanes call it.
__device__ inline void emit(const BlockCtx &c, const RefState &own, bool fallback, int32_t *__restrict__ mv_sad,
                            int32_t *__restrict__ fb_list, int32_t *__restrict__ fb_count) {
    const bool lead = c.valid && c.l == 0;
    if (lead) {
        int32_t *o = mv_sad + (size_t)c.b * 3;
        o[0] = own.dy;
        o[1] = own.dx;
        o[2] = own.sad;
    }
    const int lane = threadIdx.x & 63;
    const uint64_t fm = __ballot(lead && fallback);
    if (fm) {
        const int leader = __ffsll((unsigned long long)fm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(fb_count, __popcll(fm));
        base = __shfl(base, leader, 64);
        if (lead && fallback) fb_list[base + __popcll(fm & ((1ull << lane) - 1))] = c.b;
    }
}

// Final pass: the state after the last global round that ran (round p reads
// buffer p & 1 and writes (p + 1) & 1; rounds run while the previous one
// changed something).  A block falls back to the full search when its
// descent never stopped or when its SAD exceeds slack x the smallest SAD it
// has at its 8 neighbours' vectors (a neighbour with the block's own vector
// has the block's SAD and is not evaluated).
__global__ __launch_bounds__(RT_THREADS) void refine_final_kernel(RefineArgs a, const RefState *__restrict__ buf0,
                                                                  const RefState *__restrict__ buf1,
                                                                  const int32_t *__restrict__ changed, int rounds,
                                                                  float slack, int32_t *__restrict__ mv_sad,
                                                                  int32_t *__restrict__ fb_list,
                                                                  int32_t *__restrict__ fb_count) {
    const int tile = tile_of_workgroup();
    if (tile >= a.ntiles) return;
    const int lane = threadIdx.x & 63;
    const int32_t flag = changed[min(lane, max(rounds - 1, 0))];
    const BlockCtx c = block_ctx(a, tile);
    int nb;
    const bool inside = neighbour(c.by, c.bx, ring(c.l), 1, a.H / 8, a.W / 8, nb);
    const bool ok = c.valid && inside;
    const int nbi = ok ? nb : c.b;
    // both buffers are read; which one is final is known after the ballot
    const RefState own0 = buf0[c.b], own1 = buf1[c.b], nb0 = buf0[nbi], nb1 = buf1[nbi];
    uint32_t cw[16];
    global_cur(cw, a.cur, a.pitch, c.valid ? 8 * c.by : 0, c.valid ? 8 * c.bx : 0);
    const uint64_t zero = __ballot(lane < rounds && flag == 0);
    const int last = zero ? __ffsll((unsigned long long)zero) - 1 : rounds - 1;
    const bool one = (last + 1) & 1;
    const RefState own = one ? own1 : own0, n = one ? nb1 : nb0;
    const GlobalSad gsad{a.ref, a.H, a.W, a.pitch, 8 * c.by, 8 * c.bx};
    bool fb = false;
    if (c.valid) {
        uint32_t m = 0xffffffffu;
        if (ok && (n.dy != own.dy || n.dx != own.dx)) m = gsad(cw, a.R, n.dy, n.dx);
        m = group_min(m);
        const bool worse = (float)own.sad > slack * (float)min((uint32_t)own.sad, m);
        fb = own.unconv || worse;
    }
    emit(c, own, fb, mv_sad, fb_list, fb_count);
}

// ---------------------------------------------------------------- fused refinement
// The refinement in one launch: seed, every global round, the check and the
// fallback blocks' full search.  A workgroup per tile keeps the tile's ref
// region in LDS for the whole call.  Block states pass between tiles as
// tagged 64-bit words (tag << 32 | packed state), written and read as relaxed
// agent-scope atomics, which are coherent across the XCDs' L2 caches; the
// tag is the call's epoch + 1.  A lane that needs a neighbour's state of
// round p polls that word until it carries the tag, so a tile waits only for
// the tiles whose states it reads (no barrier across the grid).  Every tile
// runs every global round (no stop after a round without a change).  The
// check publishes one tagged word per tile with its fallback blocks; once a
// workgroup (the grid's padding included) has seen every tile's word, it
// takes the fallback blocks k = its id, + grid, ... of the grid's list (tile
// order, then block order in the tile) and runs their full search over
// [-R, R]^2.  Each workgroup keeps its own epoch word, read at the start and
// advanced at the end, so all agree on the tag without a reset between
// calls.  Every wait is bounded by the wall clock; a wait that runs out sets
// the workspace's timeout word and the kernel still ends.  Tiles wait on
// each other, so the whole grid must be resident at once: the launcher
// checks the occupancy first.
constexpr long long RF_WAIT_TICKS = 20000000;  // 0.2 s of the 100 MHz wall clock per wait
constexpr int RF_MAX_TILES = 4 * RT_THREADS;   // the listed phase scans 4 tiles per thread
constexpr int RF_MAX_ROUNDS = 32;              // improved-round bits per tile word
// s_sleep units (64 clocks) between polls: of a neighbour's state, and of
// every tile's check word (a longer sleep keeps the memory-side traffic of
// 1024 waiting workgroups away from the tiles still working; measured
// against 2/8: 0.5-1 us off the refinement)
constexpr int RF_SLEEP_NEAR = 8, RF_SLEEP_GRID = 32;

struct FusedWs {
    uint32_t *ctl;  // [0] timeout, [1 + p] round p changed some block (written by the last workgroup)
    uint32_t *ep;   // epoch per workgroup (tiles, then the grid's padding)
    uint64_t *st;   // (rounds + 1) x nb state words: seed, then after each round
    uint64_t *sum;  // per tile: fallback-block word, then (ntiles on) improved-rounds word
};

// Packed state: dx + 64 (8 bits) | dy + 64 (8) | SAD (15) | unconv (1).
__device__ inline uint64_t state_word(uint32_t tag, const RefState &s) {
    return ((uint64_t)tag << 32) | (uint32_t)(s.dx + 64) | ((uint32_t)(s.dy + 64) << 8) | ((uint32_t)s.sad << 16) |
           ((uint32_t)s.unconv << 31);
}
__device__ inline RefState word_state(uint64_t w) {
    const uint32_t v = (uint32_t)w;
    return {(int)((v >> 8) & 0xffu) - 64, (int)(v & 0xffu) - 64, (int)((v >> 16) & 0x7fffu), (int)(v >> 31)};
}
__device__ inline void put_word(uint64_t *p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t get_word(const uint64_t *p) {
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Polls the words p[i] with need[i] (already loaded once into w[i]) until
// each carries `tag`, reloading all those still stale in one round trip per
// try, for at most RF_WAIT_TICKS, and not at all once some wait of the grid
// has run out (so that a launch whose grid is not resident still ends within
// one timeout).
template <int N, int SLEEP = 2>
__device__ inline void await_words(const uint64_t *const (&p)[N], uint64_t (&w)[N], const bool (&need)[N],
                                   uint32_t tag, uint32_t *timeout) {
    const auto stale = [&]() {
        bool any = false;
#pragma unroll
        for (int i = 0; i < N; ++i) any = any || (need[i] && (uint32_t)(w[i] >> 32) != tag);
        return any;
    };
    if (!stale()) return;
    const uint64_t t0 = wall_clock64();
    bool late;
    do {
        __builtin_amdgcn_s_sleep(SLEEP);
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (need[i] && (uint32_t)(w[i] >> 32) != tag) w[i] = get_word(p[i]);
        late = (long long)(wall_clock64() - t0) >= RF_WAIT_TICKS ||
               __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    } while (stale() && !late);
    if (stale()) __hip_atomic_store(timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Full search of block (by, bx) over [-R, R]^2 by the whole workgroup (all
// threads call it): the (8 + 2R)-row clamped window in LDS; thread t takes
// the 4 horizontal offsets 4 (t % G) .. + 3 (v_qsad_pk_u16_u8) of the
// vertical offsets of chunk t / G, two at a time (one pass over 9 window
// rows feeds both); the first minimum in raster order as the minimum of
// (SAD << idx_bits | dyi S + dxi), per wave by shuffles, then an LDS atomic.
__device__ inline void full_search_one(const RefineArgs &a, int by, int bx, uint32_t *win, uint32_t *key,
                                       int32_t *__restrict__ mv_sad) {
    const int R = a.R, S = 2 * R + 1, G = ceil_div(S, 4), nch = RT_THREADS / G, DCH = ceil_div(S, nch);
    const int ws4 = ceil_div(8 + 2 * R + 4, 4), wrows = 8 + 2 * R;
    const int idx_bits = 32 - __builtin_clz((unsigned)(S * S - 1));
    const int x0 = 8 * bx, y0 = 8 * by, tid = threadIdx.x;
    stage_clamped<8>(win, wrows * ws4, tid, RT_THREADS, a.W, [&](int k, const uint8_t *&row, int &gx) {
        const int r = k / ws4;
        row = a.ref + (size_t)clampi(y0 - R + r, 0, a.H - 1) * a.pitch;
        gx = x0 - R + 4 * (k - r * ws4);
    });
    uint32_t cw[16];
    global_cur(cw, a.cur, a.pitch, y0, x0);
    if (tid == 0) *key = 0xffffffffu;
    __syncthreads();
    const int g = tid % G, ch = tid / G, d0 = ch * DCH, d1 = min(S, d0 + DCH);
    uint32_t bestk = 0xffffffffu;
    const auto close = [&](uint64_t acc, int dyi) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int dxi = 4 * g + q;
            if (dxi < S)
                bestk = min(bestk, (((uint32_t)(acc >> (16 * q)) & 0xffffu) << idx_bits) | (uint32_t)(dyi * S + dxi));
        }
    };
    if (ch < nch) {
        // vertical offsets in pairs: two independent qsad chains over 9 window rows
        for (int dyi = d0; dyi < d1; dyi += 2) {
            const bool two = dyi + 1 < d1;
            uint64_t acc0 = 0, acc1 = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const uint32_t *wr = win + min(dyi + r, wrows - 1) * ws4 + g;
                const uint32_t w0 = wr[0], w1 = wr[1], w2 = wr[2];
                const uint64_t s0 = (uint64_t)w0 | ((uint64_t)w1 << 32), s1 = (uint64_t)w1 | ((uint64_t)w2 << 32);
                if (r < 8) {
                    acc0 = __builtin_amdgcn_qsad_pk_u16_u8(s0, cw[2 * r], acc0);
                    acc0 = __builtin_amdgcn_qsad_pk_u16_u8(s1, cw[2 * r + 1], acc0);
                }
                if (r > 0) {
                    acc1 = __builtin_amdgcn_qsad_pk_u16_u8(s0, cw[2 * (r - 1)], acc1);
                    acc1 = __builtin_amdgcn_qsad_pk_u16_u8(s1, cw[2 * (r - 1) + 1], acc1);
                }
            }
            close(acc0, dyi);
            if (two) close(acc1, dyi + 1);
        }
    }
    // first the wave's minimum, then one LDS atomic per wave
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) bestk = min(bestk, (uint32_t)__shfl_xor((int)bestk, o, 64));
    if ((tid & 63) == 0) atomicMin(key, bestk);
    __syncthreads();
    if (tid == 0) {
        const uint32_t k = *key, idx = k & ((1u << idx_bits) - 1u);
        int32_t *o = mv_sad + ((size_t)by * (a.W / 8) + bx) * 3;
        o[0] = (int32_t)(idx / S) - R;
        o[1] = (int32_t)(idx % S) - R;
        o[2] = (int32_t)(k >> idx_bits);
    }
    __syncthreads();  // the window and key are reused by the next block
}

__global__ __launch_bounds__(RT_THREADS) void refine_fused_kernel(RefineArgs a, int local_rounds, int rounds,
                                                                  float slack, const int32_t *__restrict__ prop,
                                                                  FusedWs w, int32_t *__restrict__ mv_sad,
                                                                  int32_t *__restrict__ fb_list,
                                                                  int32_t *__restrict__ fb_count) {
    extern __shared__ uint32_t lds[];
    __shared__ RefState tstate[RT_Y * RT_X];
    __shared__ uint32_t sh_fb, sh_chg, sh_key, sh_owner;
    __shared__ uint32_t sh_part[RT_THREADS / 64];
    const int tile = tile_of_workgroup();  // past a.ntiles: grid padding, which joins the fallback search only
    const int nby = a.H / 8, nbx = a.W / 8, nb = nby * nbx, tid = threadIdx.x, lane = tid & 63;
    const uint32_t tag = w.ep[tile] + 1u;  // every workgroup holds the same epoch
    if (tile < a.ntiles) {
        if (tid == 0) {
            sh_fb = 0;
            sh_chg = 0;
        }
        uint32_t *reg = lds, *ctile = lds + a.rows * a.stride4;
        const BlockCtx c = block_ctx(a, tile);
        int xoff;
        RefState st = seed_tile(a, local_rounds, prop, reg, ctile, tstate, c, tile, xoff);
        if (c.valid && c.l == 0) put_word(w.st + c.b, state_word(tag, st));
        const LdsSad sad = lds_sad(reg, a, xoff, c.ty, c.tx);
        uint32_t cw[16];
        tile_cur(cw, ctile, c.ty, c.tx);

        // global rounds: candidates from the neighbours' states after round p - 1
        for (int p = 0; p < rounds; ++p) {
            const uint64_t *in = w.st + (size_t)p * nb;
            const uint64_t *src[3];
            uint64_t wv[3];
            bool ok[3];
            int my[3], mx[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {  // all three loads in flight before the first wait
                int n;
                const bool inside = neighbour(c.by, c.bx, ring(c.l), j ? global_jump(p, j - 1) : 1, nby, nbx, n);
                ok[j] = c.valid && inside;
                src[j] = in + n;
                wv[j] = ok[j] ? get_word(src[j]) : 0;
            }
            await_words<3, RF_SLEEP_NEAR>(src, wv, ok, tag, w.ctl);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const RefState n = word_state(wv[j]);
                my[j] = ok[j] ? n.dy : st.dy;
                mx[j] = ok[j] ? n.dx : st.dx;
            }
            bool improved = false;
            if (c.valid) {
                const uint32_t key = candidates_min<3>(a, cw, sad, c.l, my, mx, ok, true, st.dy, st.dx);
                improved = (int)(key >> 5) < st.sad;  // no candidate: key >> 5 = 2^27 - 1
                if (improved) {
                    int cy, cx;
                    winner<3>(key, my, mx, cy, cx);
                    st = descend(a, cw, sad, c.l, cy, cx);
                }
                if (c.l == 0) put_word(w.st + (size_t)(p + 1) * nb + c.b, state_word(tag, st));
            }
            if (__ballot(improved) && lane == 0) atomicOr(&sh_chg, 1u << p);
        }

        // check against the 8 neighbours' final vectors
        {
            const uint64_t *fin = w.st + (size_t)rounds * nb;
            int n1;
            const bool ok1 = c.valid && neighbour(c.by, c.bx, ring(c.l), 1, nby, nbx, n1);
            const uint64_t *src[1] = {fin + n1};
            uint64_t wv[1] = {ok1 ? get_word(src[0]) : 0};
            const bool need[1] = {ok1};
            await_words<1, RF_SLEEP_NEAR>(src, wv, need, tag, w.ctl);
            const RefState n = ok1 ? word_state(wv[0]) : st;
            if (c.valid) {
                uint32_t m = 0xffffffffu;
                if (ok1 && (n.dy != st.dy || n.dx != st.dx)) m = sad(cw, a.R, n.dy, n.dx);
                m = group_min(m);
                const bool worse = (float)st.sad > slack * (float)min((uint32_t)st.sad, m);
                const bool fb = st.unconv || worse;
                if (c.l == 0) {
                    if (fb) {
                        atomicOr(&sh_fb, 1u << (c.ty * RT_X + c.tx));
                    } else {  // fallback blocks are written by the full search only
                        int32_t *o = mv_sad + (size_t)c.b * 3;
                        o[0] = st.dy;
                        o[1] = st.dx;
                        o[2] = st.sad;
                    }
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            put_word(w.sum + tile, ((uint64_t)tag << 32) | sh_fb);
            put_word(w.sum + a.ntiles + tile, ((uint64_t)tag << 32) | sh_chg);
        }
    }

    // every tile's fallback word, polled by the first wave only (16 words a
    // lane, in groups of 4; fewer loads in flight while the last tiles run);
    // prefix counts over the tiles in LDS
    uint32_t *masks = lds, *pre = lds + RF_MAX_TILES, *win = pre + RF_MAX_TILES + 1;
    if (tid < 64) {
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const uint64_t *src[4];
            uint64_t v[4];
            bool need[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = 256 * q + 4 * tid + k;
                need[k] = i < a.ntiles;
                src[k] = w.sum + (need[k] ? i : 0);
                v[k] = need[k] ? get_word(src[k]) : ((uint64_t)tag << 32);
            }
            await_words<4, RF_SLEEP_GRID>(src, v, need, tag, w.ctl);
#pragma unroll
            for (int k = 0; k < 4; ++k) masks[256 * q + 4 * tid + k] = (uint32_t)v[k];
        }
    }
    __syncthreads();
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) cnt += __popc(masks[4 * tid + k]);
    // exclusive scan of the threads' counts (4 tiles each)
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) sh_part[tid >> 6] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < RT_THREADS / 64; ++k) {
        base += k < (tid >> 6) ? sh_part[k] : 0u;
        total += sh_part[k];
    }
    {
        uint32_t run = base + incl - cnt;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pre[4 * tid + k] = run;
            run += __popc(masks[4 * tid + k]);
        }
    }
    __syncthreads();
    const auto block_of = [&](int i, uint32_t m, int rank) {  // rank-th set bit of tile i's mask
        for (int r = 0; r < rank; ++r) m &= m - 1;
        const int gb = __ffs(m) - 1;
        const int by = (i / a.tiles_x) * RT_Y + gb / RT_X, bx = (i % a.tiles_x) * RT_X + gb % RT_X;
        return by * nbx + bx;
    };
    if (blockIdx.x == gridDim.x - 1) {  // the fallback list, its length and the rounds' change flags
        const uint64_t *src[4];
        uint64_t v[4];
        bool need[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = 4 * tid + k;
            need[k] = i < a.ntiles;
            src[k] = w.sum + a.ntiles + (need[k] ? i : 0);
            v[k] = need[k] ? get_word(src[k]) : ((uint64_t)tag << 32);
        }
        await_words<4>(src, v, need, tag, w.ctl);
        uint32_t chg = (uint32_t)v[0] | (uint32_t)v[1] | (uint32_t)v[2] | (uint32_t)v[3];
        for (int i = tid; i < a.ntiles; i += RT_THREADS) {
            const uint32_t m = masks[i];
            for (int r = 0; r < __popc(m); ++r) fb_list[pre[i] + r] = block_of(i, m, r);
        }
        if (tid == 0) {
            *fb_count = (int32_t)total;
            sh_chg = 0;
        }
        for (int o = 32; o >= 1; o >>= 1) chg |= (uint32_t)__shfl_xor((int)chg, o, 64);
        __syncthreads();
        if (lane == 0) atomicOr(&sh_chg, chg);
        __syncthreads();
        if (tid < rounds) w.ctl[1 + tid] = (sh_chg >> tid) & 1u;
    }
    // the fallback blocks k, k + grid, ... (consecutive k on consecutive
    // workgroups, so on different XCDs); the tile holding item k is the one
    // whose prefix range contains it (found by its owner thread)
    for (uint32_t k = blockIdx.x; k < total; k += gridDim.x) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 4 * tid + j;
            if (i < a.ntiles && pre[i] <= k && k < pre[i] + __popc(masks[i])) sh_owner = (uint32_t)i;
        }
        __syncthreads();
        const int i = (int)sh_owner;
        const int b = block_of(i, masks[i], (int)(k - pre[i]));
        full_search_one(a, b / nbx, b % nbx, win, &sh_key, mv_sad);  // ends with a barrier
    }
    if (tid == 0) w.ep[tile] = tag;
}

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
    if (!cur || !ref || !out || H < 4 || W < 4 || pitch < W || n < 0)
        return set_error(UMI_ERR_INVALID, "umi_cnn_downsample_pair: bad arguments");
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

static int cnn_chunks(int kind) { return kind == CNN_FIRST ? 2 : 18; }

umi_status umi_cnn_packed_bytes(int kind, size_t *bytes) {
    if (!bytes || kind < CNN_FIRST || kind > CNN_LAST) return set_error(UMI_ERR_INVALID, "umi_cnn_packed_bytes: kind %d", kind);
    *bytes = (size_t)cnn_chunks(kind) * 64 * 16;
    return UMI_OK;
}

umi_status umi_cnn_pack_layer(int kind, const float *w, void *packed, umi_stream s) {
    if (!w || !packed) return set_error(UMI_ERR_INVALID, "umi_cnn_pack_layer: null pointer");
    hipStream_t st = as_stream(s);
    const dim3 grid(ceil_div(cnn_chunks(kind) * 64, 256)), block(256);
    switch (kind) {
        case CNN_FIRST: hipLaunchKernelGGL(cnn_pack_kernel<CNN_FIRST>, grid, block, 0, st, w, (bf16x8 *)packed); break;
        case CNN_MID: hipLaunchKernelGGL(cnn_pack_kernel<CNN_MID>, grid, block, 0, st, w, (bf16x8 *)packed); break;
        case CNN_LAST: hipLaunchKernelGGL(cnn_pack_kernel<CNN_LAST>, grid, block, 0, st, w, (bf16x8 *)packed); break;
        default: return set_error(UMI_ERR_INVALID, "umi_cnn_pack_layer: kind %d (0, 1, 2)", kind);
    }
    return launch_check("cnn_pack_kernel");
}

umi_status umi_cnn_infer_layer(int kind, const void *in, int N, int H, int W, int dil, const void *packed,
                               const float *bias, void *out, umi_stream s) {
    if (!in || !packed || !bias || !out) return set_error(UMI_ERR_INVALID, "umi_cnn_infer_layer: null pointer");
    if (N < 0 || H < 1 || W < 1 || dil < 1) return set_error(UMI_ERR_INVALID, "umi_cnn_infer_layer: bad shape");
    const long groups = ((long)N * H * W + 31) / 32;
    if (groups == 0) return UMI_OK;
    // persistent waves: two workgroups of 4 waves per CU
    const unsigned grid = (unsigned)std::min<long>((groups + 3) / 4, 256L * 2);
    hipStream_t st = as_stream(s);
    const bf16x8 *wp = (const bf16x8 *)packed;
    switch (kind) {
        case CNN_FIRST:
            hipLaunchKernelGGL(cnn_layer_kernel<CNN_FIRST>, dim3(grid), dim3(256), 0, st, in, N, H, W, dil, wp, bias, out);
            break;
        case CNN_MID: {
            const size_t lds = (size_t)(CT_H + 2 * dil) * (CT_W + 2 * dil) * CT_PX;
            if (lds <= 64 * 1024) {  // staged tile; very large dilations read global memory per tap
                const long tiles = (long)N * ceil_div(H, CT_H) * ceil_div(W, CT_W);
                const auto k = dil == 2 ? cnn_mid_tile_kernel<2> : dil == 4 ? cnn_mid_tile_kernel<4>
                                                                             : cnn_mid_tile_kernel<0>;
                hipLaunchKernelGGL(k, dim3((unsigned)tiles), dim3(256), lds, st, (const uint16_t *)in, H, W, dil, wp,
                                   bias, (uint16_t *)out);
                return launch_check("cnn_mid_tile_kernel");
            }
            hipLaunchKernelGGL(cnn_layer_kernel<CNN_MID>, dim3(grid), dim3(256), 0, st, in, N, H, W, dil, wp, bias, out);
            break;
        }
        case CNN_LAST:
            hipLaunchKernelGGL(cnn_layer_kernel<CNN_LAST>, dim3(grid), dim3(256), 0, st, in, N, H, W, dil, wp, bias, out);
            break;
        default: return set_error(UMI_ERR_INVALID, "umi_cnn_infer_layer: kind %d (0, 1, 2)", kind);
    }
    return launch_check("cnn_layer_kernel");
}

umi_status umi_cnn_first_layer_frames(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                      const void *packed, const float *bias, void *out_nhwc, umi_stream s) {
    if (!cur || !ref || !packed || !bias || !out_nhwc)
        return set_error(UMI_ERR_INVALID, "umi_cnn_first_layer_frames: null pointer");
    if (H < 4 || W < 4 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_cnn_first_layer_frames: bad frame");
    const int tiles = ceil_div(H / 4, CT_H) * ceil_div(W / 4, CT_W);
    hipLaunchKernelGGL(cnn_first_frames_kernel, dim3(tiles), dim3(256), 0, as_stream(s), cur, ref, H, W, pitch,
                       (const bf16x8 *)packed, bias, (uint16_t *)out_nhwc);
    return launch_check("cnn_first_frames_kernel");
}

umi_status umi_cnn_first_two_layers_frames(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                          const void *packed0, const float *bias0, int dil, const void *packed1,
                                          const float *bias1, void *out_nhwc, umi_stream s) {
    if (!cur || !ref || !packed0 || !bias0 || !packed1 || !bias1 || !out_nhwc)
        return set_error(UMI_ERR_INVALID, "umi_cnn_first_two_layers_frames: null pointer");
    if (H < 4 || W < 4 || pitch < W || dil < 1)
        return set_error(UMI_ERR_INVALID, "umi_cnn_first_two_layers_frames: bad frame or dilation");
    const size_t lds = (size_t)(CT_H + 2 * dil) * (CT_W + 2 * dil) * CT_PX +
                       (size_t)(CT_H + 2 * dil + 2) * (CT_W + 2 * dil + 2) * 4;
    if (lds > 64 * 1024) return set_error(UMI_ERR_UNSUPPORTED, "umi_cnn_first_two_layers_frames: dilation %d", dil);
    const int tiles = ceil_div(H / 4, CT_H) * ceil_div(W / 4, CT_W);
    const bool a4 = ((reinterpret_cast<uintptr_t>(cur) | reinterpret_cast<uintptr_t>(ref) | (uintptr_t)pitch) & 3) == 0;
    const auto k = a4 ? (dil == 2 ? cnn_first_mid_frames_kernel<true, 2> : cnn_first_mid_frames_kernel<true, 0>)
                      : cnn_first_mid_frames_kernel<false, 0>;
    hipLaunchKernelGGL(k, dim3(tiles), dim3(256), lds, as_stream(s), cur, ref, H, W, pitch, (const bf16x8 *)packed0,
                       bias0, dil, (const bf16x8 *)packed1, bias1, (uint16_t *)out_nhwc);
    return launch_check("cnn_first_mid_frames_kernel");
}

umi_status umi_cnn_last_layer_proposals(const void *in_nhwc, int H4, int W4, int dil, const void *packed,
                                        const float *bias, int R, float *flow, int32_t *prop, umi_stream s) {
    if (!in_nhwc || !packed || !bias || !flow) return set_error(UMI_ERR_INVALID, "umi_cnn_last_layer_proposals: null");
    if (H4 < 1 || W4 < 1 || dil < 1 || R < 0) return set_error(UMI_ERR_INVALID, "umi_cnn_last_layer_proposals: shape");
    const long groups = (long)((H4 + 1) / 2) * ((W4 + 15) / 16);
    const unsigned grid = (unsigned)std::min<long>((groups + 3) / 4, 256L * 2);
    hipLaunchKernelGGL(cnn_last_prop_kernel, dim3(grid), dim3(256), 0, as_stream(s), (const uint16_t *)in_nhwc, H4,
                       W4, dil, (const bf16x8 *)packed, bias, R, flow, prop);
    return launch_check("cnn_last_prop_kernel");
}

umi_status umi_me_proposals(const float *flow, int H4, int W4, int R, int32_t *prop, umi_stream s) {
    if (!flow || !prop) return set_error(UMI_ERR_INVALID, "umi_me_proposals: null");
    const int nb = (H4 / 2) * (W4 / 2);
    hipLaunchKernelGGL(proposals_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, as_stream(s), flow, H4, W4, R, prop);
    return launch_check("proposals_kernel");
}

static size_t refine_ws_bytes(int H, int W, int rounds) {
    return 2 * round_up((size_t)(H / 8) * (W / 8) * sizeof(RefState), 256) + round_up((size_t)(rounds + 1) * 4, 256);
}

umi_status umi_me_refine_workspace(int H, int W, int rounds, size_t *bytes) {
    if (!bytes || H < 8 || W < 8 || rounds < 0) return set_error(UMI_ERR_INVALID, "umi_me_refine_workspace: bad args");
    *bytes = refine_ws_bytes(H, W, rounds);
    return UMI_OK;
}

// The fused refinement's workspace: control words, per-tile epochs, the
// (rounds + 1) state buffers, the per-tile summary words.
struct FusedLayout {
    size_t ctl, ep, st, sum, bytes;
};
static FusedLayout fused_layout(int H, int W, int rounds) {
    const int nbx = W / 8, nby = H / 8, ntiles = ceil_div(nbx, RT_X) * ceil_div(nby, RT_Y);
    FusedLayout L;
    L.ctl = 0;
    L.ep = 256;
    L.st = L.ep + round_up((size_t)round_up(ntiles, 8) * 4, 256);
    L.sum = L.st + round_up((size_t)nbx * nby * 8 * (rounds + 1), 256);
    L.bytes = L.sum + round_up((size_t)ntiles * 16, 256);
    return L;
}

umi_status umi_me_refine_fused_workspace(int H, int W, int rounds, size_t *bytes) {
    if (!bytes || H < 8 || W < 8 || rounds < 0 || rounds > RF_MAX_ROUNDS)
        return set_error(UMI_ERR_INVALID, "umi_me_refine_fused_workspace: bad args");
    *bytes = fused_layout(H, W, rounds).bytes;
    return UMI_OK;
}

umi_status umi_me_refine_fused(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R,
                               const int32_t *prop, int iters, int local_rounds, int rounds, float slack,
                               int32_t *mv_sad, int32_t *fb_list, int32_t *fb_count, void *ws, size_t ws_bytes,
                               umi_stream s) {
    if (!cur || !ref || !prop || !mv_sad || !fb_list || !fb_count || !ws)
        return set_error(UMI_ERR_INVALID, "umi_me_refine_fused: null pointer");
    if (H < 8 || W < 8 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_me_refine_fused: bad frame");
    if (R < 1 || R > 64) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_refine_fused: radius %d (1..64)", R);
    if (iters < 1 || iters > 64 || local_rounds < 0 || local_rounds > 64 || rounds < 0 || rounds > RF_MAX_ROUNDS)
        return set_error(UMI_ERR_INVALID, "umi_me_refine_fused: iters %d (1..64), local rounds %d, rounds %d (0..%d)",
                         iters, local_rounds, rounds, RF_MAX_ROUNDS);
    const FusedLayout L = fused_layout(H, W, rounds);
    if (ws_bytes < L.bytes)
        return set_error(UMI_ERR_INVALID, "umi_me_refine_fused: workspace %zu B < %zu B", ws_bytes, L.bytes);
    const int nbx = W / 8, nby = H / 8;
    RefineArgs a{cur, ref, H, W, pitch, R, iters, ceil_div(nbx, RT_X), 0, 0, 0, 0, 0};
    a.ref16 = ((reinterpret_cast<uintptr_t>(ref) | (uintptr_t)pitch) & 15) == 0 && W >= 16;
    a.cur16 = ((reinterpret_cast<uintptr_t>(cur) | (uintptr_t)pitch) & 15) == 0 && W >= 16;
    a.ntiles = a.tiles_x * ceil_div(nby, RT_Y);
    if (a.ntiles > RF_MAX_TILES)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_me_refine_fused: %d tiles of 4 x 8 blocks (at most %d)", a.ntiles,
                         RF_MAX_TILES);
    region_shape(R, a.stride4, a.rows);
    const size_t listed = 4 * (2 * (size_t)RF_MAX_TILES + 1) + 4 * (size_t)(8 + 2 * R) * ceil_div(8 + 2 * R + 4, 4);
    const size_t lds = std::max(region_lds_bytes(R), listed);
    // every tile waits on others: the grid must be resident at once (with a
    // fifth of the occupancy the runtime reports kept in reserve)
    const dim3 grid(round_up(a.ntiles, 8)), block(RT_THREADS);
    int dev = 0, cus = 0, per_cu = 0;
    UMI_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    UMI_TRY(hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"));
    UMI_TRY(hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)refine_fused_kernel,
                                                                   RT_THREADS, lds),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor"));
    if ((long)grid.x * 5 > (long)per_cu * cus * 4)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_me_refine_fused: %u workgroups, %d resident per CU x %d CUs",
                         grid.x, per_cu, cus);
    const FusedWs w{(uint32_t *)((char *)ws + L.ctl), (uint32_t *)((char *)ws + L.ep),
                    (uint64_t *)((char *)ws + L.st), (uint64_t *)((char *)ws + L.sum)};
    hipLaunchKernelGGL(refine_fused_kernel, grid, block, lds, as_stream(s), a, local_rounds, rounds, slack, prop, w,
                       mv_sad, fb_list, fb_count);
    return launch_check("refine_fused_kernel");
}

umi_status umi_me_refine_fused_status(void *ws, int rounds, int32_t *changed, int *timed_out) {
    if (!ws || !timed_out || rounds < 0 || rounds > RF_MAX_ROUNDS || (rounds > 0 && !changed))
        return set_error(UMI_ERR_INVALID, "umi_me_refine_fused_status: bad args");
    uint32_t ctl[1 + RF_MAX_ROUNDS];
    UMI_TRY(hip_check(hipMemcpy(ctl, ws, sizeof(ctl), hipMemcpyDeviceToHost), "hipMemcpy (refinement status)"));
    *timed_out = ctl[0] != 0;
    for (int p = 0; p < rounds; ++p) changed[p] = (int32_t)ctl[1 + p];
    if (ctl[0]) {
        const uint32_t z = 0;
        UMI_TRY(hip_check(hipMemcpy(ws, &z, 4, hipMemcpyHostToDevice), "hipMemcpy (refinement status)"));
    }
    return UMI_OK;
}

umi_status umi_me_refine(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R, const int32_t *prop,
                         int iters, int local_rounds, int rounds, float slack, int32_t *mv_sad, int32_t *fb_list,
                         int32_t *fb_count, void *ws, size_t ws_bytes, umi_stream s) {
    if (!cur || !ref || !prop || !mv_sad || !fb_list || !fb_count || !ws)
        return set_error(UMI_ERR_INVALID, "umi_me_refine: null pointer");
    if (H < 8 || W < 8 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_me_refine: bad frame");
    if (R < 1 || R > 64) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_refine: radius %d (1..64)", R);
    if (iters < 1 || iters > 64 || local_rounds < 0 || local_rounds > 64 || rounds < 0 || rounds > 64)
        return set_error(UMI_ERR_INVALID, "umi_me_refine: iters %d (1..64), local rounds %d, rounds %d (0..64)", iters,
                         local_rounds, rounds);
    if (ws_bytes < refine_ws_bytes(H, W, rounds))
        return set_error(UMI_ERR_INVALID, "umi_me_refine: workspace %zu B < %zu B", ws_bytes,
                         refine_ws_bytes(H, W, rounds));
    const int nbx = W / 8, nby = H / 8, nb = nbx * nby;
    const size_t sb = round_up((size_t)nb * sizeof(RefState), 256);
    RefState *buf[2] = {(RefState *)ws, (RefState *)((char *)ws + sb)};
    int32_t *changed = (int32_t *)((char *)ws + 2 * sb);
    RefineArgs a{cur, ref, H, W, pitch, R, iters, ceil_div(nbx, RT_X), 0, 0, 0, 0, 0};
    a.ref16 = ((reinterpret_cast<uintptr_t>(ref) | (uintptr_t)pitch) & 15) == 0 && W >= 16;
    a.cur16 = ((reinterpret_cast<uintptr_t>(cur) | (uintptr_t)pitch) & 15) == 0 && W >= 16;
    a.ntiles = a.tiles_x * ceil_div(nby, RT_Y);
    region_shape(R, a.stride4, a.rows);
    const size_t lds = region_lds_bytes(R);
    hipStream_t st = as_stream(s);
    const dim3 grid(round_up(a.ntiles, 8)), block(RT_THREADS);
    hipLaunchKernelGGL(refine_seed_kernel, grid, block, lds, st, a, local_rounds, prop, buf[0], changed, rounds,
                       fb_count);
    UMI_TRY(launch_check("refine_seed_kernel"));
    for (int p = 0; p < rounds; ++p) {
        hipLaunchKernelGGL(refine_round_kernel, grid, block, lds, st, a, p, buf[p & 1], buf[(p + 1) & 1],
                           p ? changed + p - 1 : (const int32_t *)nullptr, changed + p);
        UMI_TRY(launch_check("refine_round_kernel"));
    }
    hipLaunchKernelGGL(refine_final_kernel, grid, block, 0, st, a, buf[0], buf[1], changed, rounds, slack, mv_sad,
                       fb_list, fb_count);
    UMI_TRY(launch_check("refine_final_kernel"));
    // the fallback blocks: one workgroup each, over the list the final pass wrote
    return umi_me_full_search_listed_u8(cur, ref, H, W, pitch, 8, R, fb_list, fb_count, nb, mv_sad, s);
}

}  // extern "C"
