// Integral image (summed-area table), gfx950.
//
//   out[y][x] = sum_{y' <= y, x' <= x} in[y'][x']
//
// The reference computes it with two UMI launches of a running-sum strategy
// (row scan into a transposed temporary, then a scan of that temporary;
// main/iimg.cu:11-18,37-82, check loop main/iimg.cu:84-112, johnjohnlin/UMI).
// Element forms here: uint8 -> uint32 and int32 -> int32 (exact, wrapping),
// float32 -> float32 with row sums inside a tile in fp32 and every vertical
// sum and carry in fp64.
//
// Design: three launches of the tiled reduce-then-scan in iimg_dev.h (tile
// totals, carries, tile output).  The input is read twice and the output
// written once; the carries are small (row, column and tile prefix sums).
#include "common.h"
#include "iimg_dev.h"

namespace {
using namespace umi;

template <class T>
__global__ __launch_bounds__(IM_THREADS) void iimg_totals_kernel(const typename T::In *__restrict__ in, int H, int W,
                                                                 int ip, typename T::Col *__restrict__ ws) {
    __shared__ typename T::Row lds[IM_TH * IM_TW];
    __shared__ typename T::Col tmp[4];
    iimg_tile_totals<T>(lds, tmp, in, H, W, ip, iimg_rows16(in, ip), blockIdx.x, blockIdx.y, iimg_ws<T>(ws, H, W));
}

template <class T>
__global__ __launch_bounds__(256) void iimg_carry_kernel(int H, int W, typename T::Col *__restrict__ ws) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;  // whole waves: the item count is a multiple of 64
    if (k < iimg_carry_items(H, W)) iimg_carry_item<T>(H, W, k, iimg_ws<T>(ws, H, W));
}

template <class T>
__global__ __launch_bounds__(IM_THREADS) void iimg_out_kernel(const typename T::In *__restrict__ in, int H, int W,
                                                              int ip, typename T::Col *__restrict__ ws,
                                                              typename T::Out *__restrict__ out, int op) {
    __shared__ typename T::Row lds[IM_TH * IM_TW];
    __shared__ typename T::Col rl[IM_TH];
    iimg_tile_out<T>(lds, rl, in, H, W, ip, iimg_rows16(in, ip), blockIdx.x, blockIdx.y, iimg_ws<T>(ws, H, W), out,
                     op);
}

template <class T>
size_t iimg_ws_bytes(int H, int W) {
    return (size_t)iimg_ws_elems(H, W) * sizeof(typename T::Col);
}

template <class T>
umi_status iimg_launch(const typename T::In *in, int H, int W, int ip, typename T::Out *out, int op, void *ws,
                       size_t ws_bytes, umi_stream s, const char *name) {
    if (!in || !out) return set_error(UMI_ERR_INVALID, "%s: null pointer", name);
    if (H < 0 || W < 0 || ip < W || op < W)
        return set_error(UMI_ERR_INVALID, "%s: H=%d W=%d in_pitch=%d out_pitch=%d", name, H, W, ip, op);
    if (H == 0 || W == 0) return UMI_OK;
    const IimgGeom g = iimg_geom(H, W);
    const size_t need = iimg_ws_bytes<T>(H, W);
    hipStream_t st = as_stream(s);
    void *tmp = ws;
    bool own = false;
    if (!tmp) {
        UMI_TRY(hip_check(hipMalloc(&tmp, need), "hipMalloc (integral image workspace)"));
        own = true;
    } else if (ws_bytes < need) {
        return set_error(UMI_ERR_INVALID, "%s: workspace %zu B < %zu B", name, ws_bytes, need);
    }
    auto *w = static_cast<typename T::Col *>(tmp);
    hipLaunchKernelGGL(iimg_totals_kernel<T>, dim3(g.ntx, g.nty), dim3(IM_THREADS), 0, st, in, H, W, ip, w);
    umi_status rc = launch_check(name);
    if (rc == UMI_OK) {
        hipLaunchKernelGGL(iimg_carry_kernel<T>, dim3(ceil_div(iimg_carry_items(H, W), 256)), dim3(256), 0, st, H, W,
                           w);
        rc = launch_check(name);
    }
    if (rc == UMI_OK) {
        hipLaunchKernelGGL(iimg_out_kernel<T>, dim3(g.ntx, g.nty), dim3(IM_THREADS), 0, st, in, H, W, ip, w, out, op);
        rc = launch_check(name);
    }
    if (own) {
        hipError_t e = hipStreamSynchronize(st);
        hipFree(tmp);
        if (rc == UMI_OK) rc = hip_check(e, "hipStreamSynchronize (integral image)");
    }
    return rc;
}

__global__ __launch_bounds__(256) void box_sum_kernel(const uint32_t *__restrict__ S, int H, int W, int ip, int ry,
                                                      int rx, uint32_t *__restrict__ out, int op) {
    box_sum_tile(S, H, W, ip, ry, rx, out, op, blockIdx.x * BX_TX, blockIdx.y * BX_TY);
}

}  // namespace

extern "C" {

umi_status umi_iimg_workspace_size(int H, int W, int kind, size_t *bytes) {
    if (!bytes) return set_error(UMI_ERR_INVALID, "umi_iimg_workspace_size: null pointer");
    if (H < 0 || W < 0) return set_error(UMI_ERR_INVALID, "umi_iimg_workspace_size: H=%d W=%d", H, W);
    switch (kind) {
        case 0: *bytes = iimg_ws_bytes<IU8>(H, W); break;
        case 1: *bytes = iimg_ws_bytes<II32>(H, W); break;
        case 2: *bytes = iimg_ws_bytes<IF32>(H, W); break;
        default: return set_error(UMI_ERR_INVALID, "umi_iimg_workspace_size: kind %d", kind);
    }
    return UMI_OK;
}

umi_status umi_iimg_u8_u32(const uint8_t *in, int H, int W, int in_pitch, uint32_t *out, int out_pitch,
                           void *workspace, size_t workspace_bytes, umi_stream s) {
    return iimg_launch<IU8>(in, H, W, in_pitch, out, out_pitch, workspace, workspace_bytes, s, "umi_iimg_u8_u32");
}

umi_status umi_iimg_i32(const int32_t *in, int H, int W, int in_pitch, int32_t *out, int out_pitch,
                        void *workspace, size_t workspace_bytes, umi_stream s) {
    return iimg_launch<II32>(in, H, W, in_pitch, out, out_pitch, workspace, workspace_bytes, s, "umi_iimg_i32");
}

umi_status umi_iimg_f32(const float *in, int H, int W, int in_pitch, float *out, int out_pitch, void *workspace,
                        size_t workspace_bytes, umi_stream s) {
    return iimg_launch<IF32>(in, H, W, in_pitch, out, out_pitch, workspace, workspace_bytes, s, "umi_iimg_f32");
}

umi_status umi_box_sum_iimg_u32(const uint32_t *iimg, int H, int W, int in_pitch, int ry, int rx, uint32_t *out,
                                int out_pitch, umi_stream s) {
    if (!iimg || !out) return set_error(UMI_ERR_INVALID, "umi_box_sum_iimg_u32: null pointer");
    if (H < 1 || W < 1 || in_pitch < W || out_pitch < W || ry < 0 || rx < 0)
        return set_error(UMI_ERR_INVALID, "umi_box_sum_iimg_u32: H=%d W=%d pitches %d %d radii %d %d", H, W, in_pitch,
                         out_pitch, ry, rx);
    hipLaunchKernelGGL(box_sum_kernel, dim3(ceil_div(W, BX_TX), ceil_div(H, BX_TY)), dim3(256), 0, as_stream(s), iimg,
                       H, W, in_pitch, ry, rx, out, out_pitch);
    return launch_check("box_sum_kernel");
}

}  // extern "C"
