// Device-side tile bodies of the 3x3 filters and the bilateral filter,
// shared by the one-image kernels (stencil.hip) and the persistent batch
// kernels (batch.hip).  Semantics: main/filter3x3.cu:10-17,
// main/wfilter3x3.cu:8-15, main/bilateral.cu:9-24 (johnjohnlin/UMI).
#pragma once
#include "common.h"

namespace umi {

template <typename T>
struct Arith;
template <>
struct Arith<float> {
    static __device__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static __device__ float add(float a, float b) { return a + b; }
    static __device__ float mul(float a, float b) { return a * b; }
};
template <>
struct Arith<int32_t> {  // two's-complement wrap, as the reference's int32 examples rely on
    static __device__ int32_t fma(int32_t a, int32_t b, int32_t c) { return (int32_t)((uint32_t)a * (uint32_t)b + (uint32_t)c); }
    static __device__ int32_t add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
    static __device__ int32_t mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
};

// Halo tile staging: tile[e] = the image at (gy0 + e / LW, gx0 + e % LW) for
// e < n, out-of-image positions by `mode` (0: pad, 1: clamp to the edge).
// Every load is unconditional (clamped indices) and a thread's U loads are
// all issued before its first LDS store, so a tile costs one load latency
// rather than one per element; indices past n repeat element n - 1 (same
// value, same slot).  All 256 threads call it; the caller synchronises.
template <int U, typename T>
__device__ inline void stage_halo(T *tile, int n, int LW, const T *__restrict__ in, int H, int W, int ip, int gy0,
                                  int gx0, int mode, T pad) {
    const int tid = threadIdx.x;
    for (int base = 0; base < n; base += U * 256) {
        T v[U];
        int idx[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            idx[u] = min(base + u * 256 + tid, n - 1);
            const int r = idx[u] / LW, c = idx[u] - r * LW;
            const int gy = gy0 + r, gx = gx0 + c;
            ok[u] = mode != 0 || (gy >= 0 && gy < H && gx >= 0 && gx < W);
            v[u] = in[(size_t)clampi(gy, 0, H - 1) * ip + clampi(gx, 0, W - 1)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) tile[idx[u]] = ok[u] ? v[u] : pad;
    }
}

// One TX x TY output tile of a 3x3 filter by 256 threads: stage the
// (TY+2) x (TX+2) halo tile in LDS (`tile`), then each thread computes
// TX*TY/256 rows of one column.  WEIGHTED = false: mul * (init + sum x);
// true: init + sum x * w.  Ends with a barrier so `tile` can be reused.
template <typename T, bool WEIGHTED, int TX, int TY>
__device__ inline void filter3x3_tile(T *tile, const T *__restrict__ in, int H, int W, int ip, const T (&w)[9], T init,
                                      T mul, T pad, int mode, T *__restrict__ out, int op, int x0, int y0) {
    constexpr int LW = TX + 2, LH = TY + 2, ROWS = TX * TY / 256;
    static_assert(TX * TY % 256 == 0 && 256 % TX == 0, "tile shape");
    const int tid = threadIdx.x;
    stage_halo<(LH * LW + 255) / 256>(tile, LH * LW, LW, in, H, W, ip, y0 - 1, x0 - 1, mode, pad);
    __syncthreads();
    const int tx = tid % TX, ty = tid / TX;
    const int x = x0 + tx;
    if (x < W) {
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
            const int ly = ty * ROWS + i;
            const int y = y0 + ly;
            if (y < H) {
                T acc = init;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        const T v = tile[(ly + dy) * LW + tx + dx];
                        acc = WEIGHTED ? Arith<T>::fma(v, w[dy * 3 + dx], acc) : Arith<T>::add(acc, v);
                    }
                out[(size_t)y * op + x] = WEIGHTED ? acc : Arith<T>::mul(acc, mul);
            }
        }
    }
    __syncthreads();
}

// One TX x TY tile of the bilateral filter (one output per thread, 256
// threads): halo tile of (TY+2r) x (TX+2r) in `tile`, spatial weights in
// `wsh` (2r+1 values, already staged).  RT > 0 is the radius known at compile
// time (r == RT): the window loops unroll and the spatial weights stay in
// registers.  Ends with a barrier.
template <int TX, int TY, int RT = 0>
__device__ inline void bilateral_tile(float *tile, const float *wsh, const float *__restrict__ in, int H, int W, int ip,
                                      int r_, float norm, float pad, int mode, float *__restrict__ out, int op, int x0,
                                      int y0) {
    static_assert(TX * TY == 256, "one output per thread");
    const int r = RT > 0 ? RT : r_;
    const int LW = TX + 2 * r, LH = TY + 2 * r;
    const int tid = threadIdx.x;
    stage_halo<4>(tile, LH * LW, LW, in, H, W, ip, y0 - r, x0 - r, mode, pad);
    __syncthreads();
    const int tx = tid % TX, ty = tid / TX;
    const int x = x0 + tx, y = y0 + ty;
    // exp(v) = 2^(v log2 e) on v_exp_f32 (about 1 ulp): the range weight
    // without the library exp's range reduction and special cases
    const float norm2 = norm * 1.44269504088896340736f;
    if (x < W && y < H) {
        const float center = tile[(ty + r) * LW + tx + r];
        float wsum = 0.f, wxsum = 0.f;
        auto tap = [&](float n, float wy, float wx) {
            const float d = n - center;
            const float wgt = __builtin_amdgcn_exp2f(d * d * norm2) * wy * wx;
            wsum += wgt;
            wxsum = __builtin_fmaf(wgt, n, wxsum);
        };
        if constexpr (RT > 0) {
            float wv[2 * RT + 1];
#pragma unroll
            for (int k = 0; k <= 2 * RT; ++k) wv[k] = wsh[k];
#pragma unroll
            for (int dy = 0; dy <= 2 * RT; ++dy) {
                const float *row = tile + (ty + dy) * LW + tx;
#pragma unroll
                for (int dx = 0; dx <= 2 * RT; ++dx) tap(row[dx], wv[dy], wv[dx]);
            }
        } else {
            for (int dy = 0; dy <= 2 * r; ++dy) {
                const float wy = wsh[dy];
                const float *row = tile + (ty + dy) * LW + tx;
                for (int dx = 0; dx <= 2 * r; ++dx) tap(row[dx], wy, wsh[dx]);
            }
        }
        out[(size_t)y * op + x] = wxsum / wsum;
    }
    __syncthreads();
}

constexpr int BL_MAXR = 12;  // the reference example uses sigma 4 -> radius 12

}  // namespace umi
