// Integral image (summed-area table), gfx950.
//
//   out[y][x] = sum_{y' <= y, x' <= x} in[y'][x']
//
// The reference computes it with two UMI launches of a running-sum strategy
// (row scan into a transposed temporary, then a scan of that temporary;
// main/iimg.cu:11-18,37-82, check loop main/iimg.cu:84-112, johnjohnlin/UMI).
// Element forms here: uint8 -> uint32 and int32 -> int32 (exact, wrapping),
// float32 -> float32 with the row pass in fp32 and the column pass in fp64.
//
// Design (three launches, HBM-streaming):
//  K1 band kernel: a workgroup loads a band of TB rows into LDS (coalesced),
//     each wave scans whole rows (per-lane contiguous chunks + a DPP wave
//     scan of the chunk totals), then each thread runs down one column of the
//     band (column accumulator in Tcol), writing the band-local table and the
//     band's bottom row (its column totals) to the workspace;
//  K2 per column, an exclusive scan of the band totals (nb values);
//  K3 adds each band's carry to its rows (bands >= 1).
// LDS rows use the padded index e + e / C (C = per-lane chunk) so the
// chunk-strided accesses of a wave fall in distinct banks.
#include "common.h"
#include "iimg_dev.h"

namespace {
using namespace umi;

template <class T>
__global__ __launch_bounds__(IM_THREADS) void iimg_band_kernel(const typename T::In *__restrict__ in, int H, int W,
                                                               int ip, IimgPlan p, typename T::Out *__restrict__ out,
                                                               int op, typename T::Col *__restrict__ totals) {
    extern __shared__ uint32_t im_lds[];
    iimg_band<T>(reinterpret_cast<typename T::Row *>(im_lds), in, H, W, ip, p, blockIdx.x, out, op, totals);
}

template <class T>
__global__ void iimg_carry_scan_kernel(int W, int nb, typename T::Col *__restrict__ totals) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < W) iimg_carry_scan<T>(W, nb, x, totals);
}

template <class T>
__global__ void iimg_carry_add_kernel(int H, int W, int TB, const typename T::Col *__restrict__ carry,
                                      typename T::Out *__restrict__ out, int op) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = TB + blockIdx.y;  // band 0 has no carry
    if (x < W && y < H) iimg_carry_add<T>(W, TB, y, x, carry, out, op);
}

template <class T>
size_t iimg_ws_bytes(int H, int W) {
    const IimgPlan p = iimg_plan(H, W);
    return (size_t)p.nb * W * sizeof(typename T::Col);
}

template <class T>
umi_status iimg_launch(const typename T::In *in, int H, int W, int ip, typename T::Out *out, int op, void *ws,
                       size_t ws_bytes, umi_stream s, const char *name) {
    if (!in || !out) return set_error(UMI_ERR_INVALID, "%s: null pointer", name);
    if (H < 0 || W < 0 || ip < W || op < W)
        return set_error(UMI_ERR_INVALID, "%s: H=%d W=%d in_pitch=%d out_pitch=%d", name, H, W, ip, op);
    if (H == 0 || W == 0) return UMI_OK;
    const IimgPlan p = iimg_plan(H, W);
    if (p.TB < 1) return set_error(UMI_ERR_UNSUPPORTED, "%s: width %d exceeds one LDS row band", name, W);
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
    auto *tot = static_cast<typename T::Col *>(tmp);
    hipLaunchKernelGGL(iimg_band_kernel<T>, dim3(p.nb), dim3(IM_THREADS), (size_t)4 * p.LS * p.TB, st, in, H, W,
                       ip, p, out, op, tot);
    umi_status rc = launch_check(name);
    if (rc == UMI_OK && p.nb > 1) {
        hipLaunchKernelGGL(iimg_carry_scan_kernel<T>, dim3(ceil_div(W, 256)), dim3(256), 0, st, W, p.nb, tot);
        rc = launch_check(name);
        if (rc == UMI_OK) {
            hipLaunchKernelGGL(iimg_carry_add_kernel<T>, dim3(ceil_div(W, 256), H - p.TB), dim3(256), 0, st, H, W,
                               p.TB, tot, out, op);
            rc = launch_check(name);
        }
    }
    if (own) {
        hipError_t e = hipStreamSynchronize(st);
        hipFree(tmp);
        if (rc == UMI_OK) rc = hip_check(e, "hipStreamSynchronize (integral image)");
    }
    return rc;
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

}  // extern "C"
