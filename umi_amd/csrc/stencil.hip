// 3x3 filters and the bilateral filter (gfx950).
//
// filter3x3  : out = mul * (init + sum_{3x3} x)           main/filter3x3.cu:10-17,36-58
// wfilter3x3 : out = init + sum_{dy,dx} x * w[dy][dx]      main/wfilter3x3.cu:8-15,39-67
// bilateral  : out = sum w*n / sum w, w = exp((n-c)^2 norm) ws[dy] ws[dx]
//                                                          main/bilateral.cu:9-24,54-99
// (johnjohnlin/UMI).  Out-of-image taps take the PAD value (the reference's
// PAD boundary, src/UMI.h:205-218) or, with mode = clamp, the edge pixel.
// Taps are accumulated in the strategy's order (dy outer, dx inner, starting
// from init) with fused multiply-adds, so float results equal the oracle's
// fmaf loop bit for bit (filters) or to a few ulp (bilateral: exp).
//
// Design: HBM-bound stencils.  A 256-thread workgroup stages a
// (TY+2r) x (TX+2r) halo tile in LDS (coalesced loads, boundary applied once
// at load time), then each thread computes a vertical strip of outputs,
// reusing the tile rows it has in flight; stores are coalesced along x.
#include "common.h"
#include "stencil_dev.h"

namespace {
using namespace umi;

// Tiles of 64 x 32 outputs; images with fewer than 2048 such tiles (below
// about 4 Mpx) take 64 x 16, whose twice as many workgroups overlap their
// load -> compute -> store chains (1024^2: 4.3 -> 3.8 us; larger images keep
// the taller tile's smaller halo share).
constexpr int F_TX = 64, F_TY = 32, F_TY_SMALL = 16;
constexpr int BL_TX = 32, BL_TY = 8;  // one output per thread

template <typename T, bool WEIGHTED, int TY>
__global__ __launch_bounds__(256) void filter3x3_kernel(const T *__restrict__ in, int H, int W, int ip,
                                                        const T *__restrict__ wts, T init, T mul, T pad, int mode,
                                                        T *__restrict__ out, int op) {
    __shared__ T tile[(TY + 2) * (F_TX + 2)];
    T w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) w[k] = WEIGHTED ? wts[k] : T(0);
    filter3x3_tile<T, WEIGHTED, F_TX, TY>(tile, in, H, W, ip, w, init, mul, pad, mode, out, op, blockIdx.x * F_TX,
                                          blockIdx.y * TY);
}

template <int RT>
__global__ __launch_bounds__(256) void bilateral_kernel(const float *__restrict__ in, int H, int W, int ip, int r,
                                                        const float *__restrict__ ws, float norm, float pad,
                                                        int mode, float *__restrict__ out, int op) {
    extern __shared__ float bl_lds[];
    float *wsh = bl_lds;
    if ((int)threadIdx.x < 2 * r + 1) wsh[threadIdx.x] = ws[threadIdx.x];
    bilateral_tile<BL_TX, BL_TY, RT>(bl_lds + 32, wsh, in, H, W, ip, r, norm, pad, mode, out, op, blockIdx.x * BL_TX,
                                     blockIdx.y * BL_TY);
}

template <typename T, bool WEIGHTED>
umi_status filter_launch(const T *in, int H, int W, int ip, const T *w, T init, T mul, T pad, int mode, T *out,
                         int op, umi_stream s, const char *name) {
    if (!in || !out || (WEIGHTED && !w)) return set_error(UMI_ERR_INVALID, "%s: null pointer", name);
    if (H < 0 || W < 0 || ip < W || op < W)
        return set_error(UMI_ERR_INVALID, "%s: H=%d W=%d in_pitch=%d out_pitch=%d", name, H, W, ip, op);
    if (mode != 0 && mode != 1) return set_error(UMI_ERR_INVALID, "%s: mode %d (0 pad, 1 clamp)", name, mode);
    if (H == 0 || W == 0) return UMI_OK;
    if ((long)ceil_div(W, F_TX) * ceil_div(H, F_TY) < 2048) {
        dim3 grid(ceil_div(W, F_TX), ceil_div(H, F_TY_SMALL));
        hipLaunchKernelGGL((filter3x3_kernel<T, WEIGHTED, F_TY_SMALL>), grid, dim3(256), 0, as_stream(s), in, H, W,
                           ip, w, init, mul, pad, mode, out, op);
    } else {
        dim3 grid(ceil_div(W, F_TX), ceil_div(H, F_TY));
        hipLaunchKernelGGL((filter3x3_kernel<T, WEIGHTED, F_TY>), grid, dim3(256), 0, as_stream(s), in, H, W, ip, w,
                           init, mul, pad, mode, out, op);
    }
    return launch_check(name);
}

}  // namespace

extern "C" {

umi_status umi_filter3x3_f32(const float *in, int H, int W, int in_pitch, float init, float mul, float pad,
                             int mode, float *out, int out_pitch, umi_stream s) {
    return filter_launch<float, false>(in, H, W, in_pitch, nullptr, init, mul, pad, mode, out, out_pitch, s,
                                       "umi_filter3x3_f32");
}

umi_status umi_filter3x3_i32(const int32_t *in, int H, int W, int in_pitch, int32_t init, int32_t mul,
                             int32_t pad, int mode, int32_t *out, int out_pitch, umi_stream s) {
    return filter_launch<int32_t, false>(in, H, W, in_pitch, nullptr, init, mul, pad, mode, out, out_pitch, s,
                                         "umi_filter3x3_i32");
}

umi_status umi_wfilter3x3_f32(const float *in, int H, int W, int in_pitch, const float *weights, float init,
                              float pad, int mode, float *out, int out_pitch, umi_stream s) {
    return filter_launch<float, true>(in, H, W, in_pitch, weights, init, 0.f, pad, mode, out, out_pitch, s,
                                      "umi_wfilter3x3_f32");
}

umi_status umi_wfilter3x3_i32(const int32_t *in, int H, int W, int in_pitch, const int32_t *weights,
                              int32_t init, int32_t pad, int mode, int32_t *out, int out_pitch, umi_stream s) {
    return filter_launch<int32_t, true>(in, H, W, in_pitch, weights, init, 0, pad, mode, out, out_pitch, s,
                                        "umi_wfilter3x3_i32");
}

umi_status umi_bilateral_f32(const float *in, int H, int W, int in_pitch, int radius, const float *spatial,
                             float norm, float pad, int mode, float *out, int out_pitch, umi_stream s) {
    if (!in || !out || !spatial) return set_error(UMI_ERR_INVALID, "umi_bilateral_f32: null pointer");
    if (H < 0 || W < 0 || in_pitch < W || out_pitch < W)
        return set_error(UMI_ERR_INVALID, "umi_bilateral_f32: H=%d W=%d pitches %d %d", H, W, in_pitch, out_pitch);
    if (radius < 0 || radius > BL_MAXR)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_bilateral_f32: radius %d (0..%d)", radius, BL_MAXR);
    if (mode != 0 && mode != 1) return set_error(UMI_ERR_INVALID, "umi_bilateral_f32: mode %d", mode);
    if (H == 0 || W == 0) return UMI_OK;
    const size_t lds = sizeof(float) * ((size_t)(BL_TX + 2 * radius) * (BL_TY + 2 * radius) + 32);
    dim3 grid(ceil_div(W, BL_TX), ceil_div(H, BL_TY));
    if (radius == 2)  // the 5 x 5 window: unrolled
        hipLaunchKernelGGL(bilateral_kernel<2>, grid, dim3(256), lds, as_stream(s), in, H, W, in_pitch, radius,
                           spatial, norm, pad, mode, out, out_pitch);
    else
        hipLaunchKernelGGL(bilateral_kernel<0>, grid, dim3(256), lds, as_stream(s), in, H, W, in_pitch, radius,
                           spatial, norm, pad, mode, out, out_pitch);
    return launch_check("umi_bilateral_f32");
}

}  // extern "C"
