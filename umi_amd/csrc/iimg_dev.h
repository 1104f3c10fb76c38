// Device-side bodies of the integral image (band scan, carry scan, carry
// add), shared by the one-image kernels (iimg.hip) and the batch kernels
// (batch.hip).  Semantics: main/iimg.cu:11-18,37-82,84-112 (johnjohnlin/UMI).
#pragma once
#include "common.h"
#include "wave.h"

namespace umi {

template <typename Tin, typename Trow, typename Tcol, typename Tout>
struct IimgTypes {
    using In = Tin;
    using Row = Trow;
    using Col = Tcol;
    using Out = Tout;
};
using IU8 = IimgTypes<uint8_t, uint32_t, uint32_t, uint32_t>;
using II32 = IimgTypes<int32_t, uint32_t, uint32_t, int32_t>;
using IF32 = IimgTypes<float, float, double, float>;

constexpr int IM_THREADS = 256;
constexpr int IM_LDS = 64 * 1024;

struct IimgPlan {
    int C, LS, TB, nb;
};
__host__ __device__ inline IimgPlan iimg_plan(int H, int W) {
    IimgPlan p;
    p.C = ceil_div(W, 64);
    p.LS = W + ceil_div(W, p.C) + 1;
    p.TB = IM_LDS / (4 * p.LS);
    if (p.TB > 32) p.TB = 32;
    p.nb = p.TB > 0 ? ceil_div(H, p.TB) : 0;
    return p;
}

// Band b (rows b*TB ..) of one image by 256 threads: band-local summed-area
// table into `out`, the band's column totals into totals[b][x].  `rows` is
// LDS of TB x LS row elements.  Ends with a barrier.
template <class T>
__device__ inline void iimg_band(typename T::Row *rows, const typename T::In *__restrict__ in, int H, int W, int ip,
                                 const IimgPlan &p, int b, typename T::Out *__restrict__ out, int op,
                                 typename T::Col *__restrict__ totals) {
    using Row = typename T::Row;
    using Col = typename T::Col;
    const int C = p.C, LS = p.LS;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int y0 = b * p.TB;
    const int nr = min(p.TB, H - y0);
    for (int r = 0; r < nr; ++r) {
        const typename T::In *src = in + (size_t)(y0 + r) * ip;
        for (int x = tid; x < W; x += IM_THREADS) rows[r * LS + x + x / C] = (Row)src[x];
    }
    __syncthreads();
    for (int r = wv; r < nr; r += IM_THREADS / 64) {
        Row *row = rows + r * LS + lane * (C + 1);
        const int n = min(C, W - lane * C);
        Row run = 0;
        for (int k = 0; k < n; ++k) {
            run = run + row[k];
            row[k] = run;
        }
        const Row incl = wave_inclusive_scan(run, [](Row a, Row c) { return a + c; });
        const Row excl = incl - run;
        for (int k = 0; k < n; ++k) row[k] = row[k] + excl;
    }
    __syncthreads();
    for (int x = tid; x < W; x += IM_THREADS) {
        Col acc = 0;
        for (int r = 0; r < nr; ++r) {
            acc = acc + (Col)rows[r * LS + x + x / C];
            out[(size_t)(y0 + r) * op + x] = (typename T::Out)acc;
        }
        totals[(size_t)b * W + x] = acc;
    }
    __syncthreads();
}

// Exclusive scan over the nb band totals of column x (in place).
template <class T>
__device__ inline void iimg_carry_scan(int W, int nb, int x, typename T::Col *__restrict__ totals) {
    typename T::Col run = 0;
    for (int b = 0; b < nb; ++b) {
        const typename T::Col t = totals[(size_t)b * W + x];
        totals[(size_t)b * W + x] = run;
        run = run + t;
    }
}

// out[y][x] += carry of y's band.
template <class T>
__device__ inline void iimg_carry_add(int W, int TB, int y, int x, const typename T::Col *__restrict__ carry,
                                      typename T::Out *__restrict__ out, int op) {
    typename T::Out *o = out + (size_t)y * op + x;
    *o = (typename T::Out)((typename T::Col)*o + carry[(size_t)(y / TB) * W + x]);
}

}  // namespace umi
