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
// Design: one launch of the single-pass tile kernel in iimg_dev.h (tiles
// ordered by an atomic ticket; each publishes its column, row and tile sums
// and reads those of the tiles above and to its left).  The input is read
// once and the output written once.
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <mutex>
#include "common.h"
#include "iimg_dev.h"

namespace {
using namespace umi;

}  // namespace
int umi::g_iimg_per_cu = 0;
namespace {
// Tile height selection: 0 = by size, else RPL (rows per lane: 1, 2 or 4).
std::atomic<int> g_iimg_rpl{0};

template <class T, int RPL>
__global__ __launch_bounds__(IM_THREADS) void iimg_kernel(const typename T::In *__restrict__ in, int H, int W, int ip,
                                                          typename T::Out *__restrict__ out, int op,
                                                          uint8_t *__restrict__ ws) {
    extern __shared__ __attribute__((aligned(16))) uint8_t iimg_lds[];
    auto &sm = *reinterpret_cast<IimgSmem<T, RPL> *>(iimg_lds);
    auto *ctl = reinterpret_cast<IimgCtl *>(ws);
    auto *words = reinterpret_cast<uint64_t *>(ws + IM_CTL_BYTES);
    const IimgGeom g = iimg_geom(H, W, 16 * RPL);
    const int total = g.ntx * g.nty;
    const bool rows16 = iimg_rows16(in, ip);
    iimg_begin(sm, ctl);
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const IimgTileRef d{0, H, W, ip, op, t % g.ntx, t / g.ntx, g.ntx};
        iimg_tile<T, RPL>(sm, d, in, rows16, out, ctl, words);
    }
    iimg_finish(ctl, sm.tag);
}

// Zeroes a workspace the call allocated itself (a kernel on the call's stream).
__global__ void iimg_zero_kernel(uint4 *__restrict__ p, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = uint4{0, 0, 0, 0};
}

// RPL for an image: the tallest tile that still gives a tile per CU (256),
// else RPL 1 (1920 x 1080: RPL 2, 272 tiles, measured fastest of the three).
int iimg_pick_rpl(int H, int W) {
    const int want = g_iimg_rpl.load();
    if (want) return want;
    for (int rpl : {4, 2})
        if ((long long)ceil_div(W, IM_TW) * ceil_div(H, 16 * rpl) >= 256) return rpl;
    return 1;
}

template <class T>
size_t iimg_ws_bytes(int H, int W) {
    // enough for any tile height the call may pick
    size_t m = 0;
    for (int rpl : {1, 2, 4}) {
        const IimgGeom g = iimg_geom(H, W, 16 * rpl);
        m = std::max(m, iimg_ws_bytes_tiles<typename T::Col>((long long)g.ntx * g.nty, 16 * rpl));
    }
    return m;
}

template <class T, int RPL>
umi_status iimg_launch_rpl(const typename T::In *in, int H, int W, int ip, typename T::Out *out, int op, uint8_t *ws,
                           hipStream_t st, const char *name) {
    const size_t lds = sizeof(IimgSmem<T, RPL>);
    static const hipError_t attr =
        hipFuncSetAttribute((const void *)iimg_kernel<T, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const IimgGeom g = iimg_geom(H, W, 16 * RPL);
    const int grid = iimg_grid(attr, g.ntx * g.nty, g_iimg_per_cu);
    if (grid < 1) return set_error(UMI_ERR_HIP, "%s: launch setup failed", name);
    hipLaunchKernelGGL((iimg_kernel<T, RPL>), dim3(grid), dim3(IM_THREADS), lds, st, in, H, W, ip, out, op, ws);
    return launch_check(name);
}

template <class T>
umi_status iimg_launch(const typename T::In *in, int H, int W, int ip, typename T::Out *out, int op, void *ws,
                       size_t ws_bytes, umi_stream s, const char *name) {
    if (!in || !out) return set_error(UMI_ERR_INVALID, "%s: null pointer", name);
    if (H < 0 || W < 0 || ip < W || op < W)
        return set_error(UMI_ERR_INVALID, "%s: H=%d W=%d in_pitch=%d out_pitch=%d", name, H, W, ip, op);
    if (H == 0 || W == 0) return UMI_OK;
    const size_t need = iimg_ws_bytes<T>(H, W);
    hipStream_t st = as_stream(s);
    void *tmp = ws;
    bool own = false;
    if (!tmp) {
        UMI_TRY(hip_check(hipMalloc(&tmp, need), "hipMalloc (integral image workspace)"));
        own = true;
        // a fresh allocation may hold an earlier workspace's words: zero all of it
        const size_t n16 = need / 16;
        hipLaunchKernelGGL(iimg_zero_kernel, dim3((unsigned)std::min<size_t>((n16 + 255) / 256, 4096)), dim3(256), 0, st,
                           static_cast<uint4 *>(tmp), n16);
        umi_status z = launch_check("iimg_zero_kernel");
        if (z != UMI_OK) {
            (void)hipFree(tmp);
            return z;
        }
    } else if (ws_bytes < need) {
        return set_error(UMI_ERR_INVALID, "%s: workspace %zu B < %zu B", name, ws_bytes, need);
    }
    auto *w = static_cast<uint8_t *>(tmp);
    umi_status rc;
    switch (iimg_pick_rpl(H, W)) {
        case 4: rc = iimg_launch_rpl<T, 4>(in, H, W, ip, out, op, w, st, name); break;
        case 2: rc = iimg_launch_rpl<T, 2>(in, H, W, ip, out, op, w, st, name); break;
        default: rc = iimg_launch_rpl<T, 1>(in, H, W, ip, out, op, w, st, name); break;
    }
    if (own) {
        hipError_t e = hipStreamSynchronize(st);
        (void)hipFree(tmp);  // the synchronisation's status is the one reported
        if (rc == UMI_OK) rc = hip_check(e, "hipStreamSynchronize (integral image)");
    }
    return rc;
}

__global__ __launch_bounds__(256) void box_sum_kernel(const uint32_t *__restrict__ S, int H, int W, int ip, int ry,
                                                      int rx, uint32_t *__restrict__ out, int op) {
    box_sum_tile(S, H, W, ip, ry, rx, out, op, blockIdx.x * BX_TX, blockIdx.y * BX_TY);
}


// Rectangle (Haar-like) features from an inclusive integral image: thread =
// window position (x, y) (64 consecutive x per wave); per feature, up to
// three rectangles (x, y, w, h, weight) relative to the window, each sum
// from four corners; unsigned arithmetic, exact mod 2^32.  Feature
// parameters are the same for every lane and travel by value in the kernel
// arguments (scalar loads from the argument segment), at most HF_CHUNK
// features per launch, so the kernel reads exactly the table the host
// checked; the output planes are written coalesced.
constexpr int HF_RECTS = 3, HF_FIELDS = 5, HF_CHUNK = 32;
struct HaarChunk {
    int32_t r[HF_CHUNK * HF_RECTS * HF_FIELDS];
};

// The same from LDS: a workgroup covers 64 x 4 positions and stages the
// integral image's (4 + win_h) x (64 + win_w) values they read, from row y0 - 1
// and column x0 - 1, with zeros standing in for row or column -1 (S(-1, .) =
// S(., -1) = 0), so every corner read is one branch-free LDS read.
__global__ __launch_bounds__(256) void haar_lds_kernel(const uint32_t *__restrict__ S, int ip, int OH, int OW,
                                                       int win_h, int win_w, const HaarChunk feats, int nfeat,
                                                       int32_t *__restrict__ out) {
    extern __shared__ uint32_t sm[];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    const int tw = 64 + win_w, th = 4 + win_h;  // staged columns x0 - 1 .., rows y0 - 1 ..
    const int H = OH + win_h - 1, W = OW + win_w - 1;
    for (int e = threadIdx.x; e < th * tw; e += 256) {
        const int r = e / tw, c = e - r * tw, gy = y0 - 1 + r, gx = x0 - 1 + c;
        sm[e] = (gy >= 0 && gx >= 0 && gy < H && gx < W) ? S[(size_t)gy * ip + gx] : 0u;
    }
    __syncthreads();
    const int lx = (int)(threadIdx.x & 63), ly = (int)(threadIdx.x >> 6);
    const int x = x0 + lx, y = y0 + ly;
    if (x >= OW || y >= OH) return;
    const size_t plane = (size_t)OH * OW;
    const uint32_t *base = sm + ly * tw + lx;  // staged (y - 1, x - 1)
    for (int f = 0; f < nfeat; ++f) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < HF_RECTS; ++k) {
            const int32_t *r = feats.r + (f * HF_RECTS + k) * HF_FIELDS;
            const int rx = r[0], ry = r[1], rw = r[2], rh = r[3], wt = r[4];
            if (wt == 0) continue;  // uniform
            const uint32_t *ta = base + ry * tw + rx, *tb = base + (ry + rh) * tw + rx;
            v += (uint32_t)wt * (tb[rw] - tb[0] - ta[rw] + ta[0]);
        }
        out[f * plane + (size_t)y * OW + x] = (int32_t)v;
    }
}
__global__ __launch_bounds__(256) void haar_kernel(const uint32_t *__restrict__ S, int ip, int OH, int OW,
                                                   const HaarChunk feats, int nfeat, int32_t *__restrict__ out) {
    const int x = blockIdx.x * 64 + (int)(threadIdx.x & 63), y = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (x >= OW || y >= OH) return;
    const size_t plane = (size_t)OH * OW;
    for (int f = 0; f < nfeat; ++f) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < HF_RECTS; ++k) {
            const int32_t *r = feats.r + (f * HF_RECTS + k) * HF_FIELDS;
            const int rx = r[0], ry = r[1], rw = r[2], rh = r[3], wt = r[4];
            if (wt == 0) continue;  // uniform
            const int xa = x + rx - 1, xb = x + rx + rw - 1, ya = y + ry - 1, yb = y + ry + rh - 1;
            const uint32_t *rowb = S + (size_t)yb * ip, *rowa = S + (size_t)max(ya, 0) * ip;
            const uint32_t a = rowb[xb], b = rowb[max(xa, 0)], c = rowa[xb], d = rowa[max(xa, 0)];
            const uint32_t sum = a - (xa >= 0 ? b : 0u) - (ya >= 0 ? c : 0u) + (xa >= 0 && ya >= 0 ? d : 0u);
            v += (uint32_t)wt * sum;
        }
        out[f * plane + (size_t)y * OW + x] = (int32_t)v;
    }
}

}  // namespace

// Allocation-free integral image: the workspace of umi_iimg_u8_u32_noalloc is
// this array, one copy per GPU, zero at load (as a workspace must be before
// its first use); every call leaves it ready for the next, so nothing clears
// it.  Host code reaches it only through the address that
// hipGetSymbolAddress gives for the current device.
constexpr size_t IIMG_SCRATCH_BYTES = 16u << 20;
static __device__ __align__(256) uint8_t scratch[IIMG_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up on its first use there.
umi_status scratch_ptr(void **out) {
    int dev = 0;
    UMI_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_iimg_u8_u32_noalloc: device %d beyond %d", dev, SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        UMI_TRY(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (integral image scratch)"));
        if (!p) return set_error(UMI_ERR_HIP, "hipGetSymbolAddress (integral image scratch): null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return UMI_OK;
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

umi_status umi_iimg_workspace_status(void *workspace, int *timed_out) {
    if (!workspace || !timed_out) return set_error(UMI_ERR_INVALID, "umi_iimg_workspace_status: null pointer");
    uint32_t f = 0;
    auto *flag = reinterpret_cast<uint8_t *>(workspace) + offsetof(IimgCtl, timeout);
    UMI_TRY(hip_check(hipMemcpy(&f, flag, 4, hipMemcpyDeviceToHost), "hipMemcpy (integral image status)"));
    *timed_out = f != 0;
    if (f) {
        const uint32_t z = 0;
        UMI_TRY(hip_check(hipMemcpy(flag, &z, 4, hipMemcpyHostToDevice), "hipMemcpy (integral image status)"));
    }
    return UMI_OK;
}

umi_status umi_iimg_set_grid(int per_cu) {
    if (per_cu < 0 || per_cu > 8) return set_error(UMI_ERR_INVALID, "umi_iimg_set_grid: %d (0 .. 8)", per_cu);
    g_iimg_per_cu = per_cu;
    return UMI_OK;
}

umi_status umi_iimg_set_config(int rows_per_lane) {
    if (rows_per_lane != 0 && rows_per_lane != 1 && rows_per_lane != 2 && rows_per_lane != 4)
        return set_error(UMI_ERR_INVALID, "umi_iimg_set_config: %d (0, 1, 2 or 4)", rows_per_lane);
    g_iimg_rpl.store(rows_per_lane);
    return UMI_OK;
}

umi_status umi_iimg_u8_u32(const uint8_t *in, int H, int W, int in_pitch, uint32_t *out, int out_pitch,
                           void *workspace, size_t workspace_bytes, umi_stream s) {
    return iimg_launch<IU8>(in, H, W, in_pitch, out, out_pitch, workspace, workspace_bytes, s, "umi_iimg_u8_u32");
}

umi_status umi_iimg_u8_u32_noalloc(const uint8_t *in, int H, int W, int in_pitch, uint32_t *out, int out_pitch,
                                   umi_stream s) {
    const char *name = "umi_iimg_u8_u32_noalloc";
    if (!in || !out) return set_error(UMI_ERR_INVALID, "%s: null pointer", name);
    if (H < 0 || W < 0 || in_pitch < W || out_pitch < W)
        return set_error(UMI_ERR_INVALID, "%s: H=%d W=%d in_pitch=%d out_pitch=%d", name, H, W, in_pitch, out_pitch);
    if (H == 0 || W == 0) return UMI_OK;
    const size_t need = iimg_ws_bytes<IU8>(H, W);
    if (need > IIMG_SCRATCH_BYTES)
        return set_error(UMI_ERR_INVALID, "%s: %d x %d needs %zu B of scratch, more than its %zu B", name, H, W, need,
                         IIMG_SCRATCH_BYTES);
    void *ws = nullptr;
    UMI_TRY(scratch_ptr(&ws));
    return iimg_launch<IU8>(in, H, W, in_pitch, out, out_pitch, ws, IIMG_SCRATCH_BYTES, s, name);
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

umi_status umi_haar_features_u32(const uint32_t *iimg, int H, int W, int pitch, const int32_t *feats, int nfeat,
                                 int win_h, int win_w, int32_t *out, umi_stream s) {
    if (!iimg || (nfeat > 0 && !feats) || !out) return set_error(UMI_ERR_INVALID, "umi_haar_features_u32: null pointer");
    if (H < 1 || W < 1 || pitch < W || nfeat < 0 || win_h < 1 || win_w < 1 || win_h > H || win_w > W)
        return set_error(UMI_ERR_INVALID, "umi_haar_features_u32: H=%d W=%d pitch=%d window %dx%d", H, W, pitch,
                         win_h, win_w);
    // every rectangle must lie inside the window (so every corner read stays inside the image)
    for (int f = 0; f < nfeat; ++f)
        for (int k = 0; k < HF_RECTS; ++k) {
            const int32_t *r = feats + (f * HF_RECTS + k) * HF_FIELDS;
            if (r[4] == 0) continue;
            if (r[0] < 0 || r[1] < 0 || r[2] < 1 || r[3] < 1 || r[0] + r[2] > win_w || r[1] + r[3] > win_h)
                return set_error(UMI_ERR_INVALID, "umi_haar_features_u32: feature %d rectangle %d outside the %dx%d window",
                                 f, k, win_h, win_w);
        }
    const int OH = H - win_h + 1, OW = W - win_w + 1;
    const size_t plane = (size_t)OH * OW;
    // the staged tile of (4 + win_h) x (64 + win_w) values fits 48 KB for square windows up to 80 x 80
    const size_t lds = (size_t)(4 + win_h) * (64 + win_w) * 4;
    for (int f0 = 0; f0 < nfeat; f0 += HF_CHUNK) {
        const int nf = min(HF_CHUNK, nfeat - f0);
        HaarChunk c{};
        std::memcpy(c.r, feats + (size_t)f0 * HF_RECTS * HF_FIELDS, sizeof(int32_t) * nf * HF_RECTS * HF_FIELDS);
        int32_t *o = out + (size_t)f0 * plane;
        if (lds <= 48 * 1024) {
            hipLaunchKernelGGL(haar_lds_kernel, dim3(ceil_div(OW, 64), ceil_div(OH, 4)), dim3(256), lds, as_stream(s),
                               iimg, pitch, OH, OW, win_h, win_w, c, nf, o);
            UMI_TRY(launch_check("haar_lds_kernel"));
        } else {
            hipLaunchKernelGGL(haar_kernel, dim3(ceil_div(OW, 64), ceil_div(OH, 4)), dim3(256), 0, as_stream(s), iimg,
                               pitch, OH, OW, c, nf, o);
            UMI_TRY(launch_check("haar_kernel"));
        }
    }
    return UMI_OK;
}

}  // extern "C"
