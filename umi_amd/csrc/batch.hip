// Batches of images of uneven sizes, one persistent launch per operator.
//
// The reference launches one UMI kernel per image (one Execute per call,
// src/UMI_gpu.h:409-623; main/wfilter3x3.cu:39, main/bilateral.cu:54,
// main/iimg.cu:37,60, johnjohnlin/UMI).  Here a batch object holds, per
// image, its size and its offsets into one input and one output buffer, plus
// prefix sums of the work items (tiles, bands, columns, rows) of every image
// for every tile configuration and an owner table giving the image of every
// work item, all resident in device memory.  A batched operator is then one
// persistent kernel whose workgroups stride over the concatenated work list,
// read their image from the owner table (one load, where a binary search
// over the prefix array took log2(n) dependent ones) and run the same tile
// body as the one-image kernel.  No
// launch allocates or copies, so the launches can be captured in a hipGraph.
// Tile configurations are selectable so a tuner can pick the fastest.
#include "common.h"
#include "iimg_dev.h"
#include "stencil_dev.h"

#include <algorithm>
#include <vector>

namespace {
using namespace umi;

struct BatchImage {
    long long in_off, out_off;  // element offsets from the base pointers
    int H, W, in_pitch, out_pitch;
};

// One integral-image tile of the batch, indexed by its ticket: everything the
// tile needs in one (wave-uniform) descriptor load.  tbase is the ticket of
// the image's tile (0, 0), so the published words of tile (a, b) of the same
// image sit at tbase + a * ntx + b.
struct IimgBatchTile {
    long long in_off, out_off, tbase;
    int H, W, in_pitch, out_pitch, tx, ty, ntx, pad;
};
constexpr int NI = 3;                  // integral-image configurations: rows per lane 1, 2, 4
constexpr int IRPL[NI] = {1, 2, 4};

// 3x3 filter tile shapes (TX, TY) and bilateral tile shapes (TX x TY = 256).
constexpr int NF = 4, NB = 4;
constexpr int FTX[NF] = {64, 64, 128, 32}, FTY[NF] = {32, 16, 16, 32};
constexpr int BTX[NB] = {32, 16, 64, 128}, BTY[NB] = {8, 16, 4, 2};

constexpr int PERSIST_WG = 256 * 8;  // workgroups of a persistent launch (8 per CU)

template <int TX, int TY>
__global__ __launch_bounds__(256) void batch_wfilter_kernel(const BatchImage *__restrict__ imgs, int n,
                                                            const int *__restrict__ prefix,
                                                            const int *__restrict__ owner, const float *__restrict__ in,
                                                            const float *__restrict__ wts, float init, float pad,
                                                            int mode, float *__restrict__ out) {
    __shared__ float tile[(TY + 2) * (TX + 2)];
    float w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) w[k] = wts[k];
    const int total = prefix[n];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = owner[t];
        const BatchImage im = imgs[i];
        const int local = t - prefix[i], tiles_x = ceil_div(im.W, TX);
        const int ty = local / tiles_x, tx = local - ty * tiles_x;
        filter3x3_tile<float, true, TX, TY>(tile, in + im.in_off, im.H, im.W, im.in_pitch, w, init, 0.f, pad, mode,
                                            out + im.out_off, im.out_pitch, tx * TX, ty * TY);
    }
}

template <int TX, int TY, int RT>
__global__ __launch_bounds__(256) void batch_bilateral_kernel(const BatchImage *__restrict__ imgs, int n,
                                                              const int *__restrict__ prefix,
                                                              const int *__restrict__ owner,
                                                              const float *__restrict__ in, int r,
                                                              const float *__restrict__ ws, float norm, float pad,
                                                              int mode, float *__restrict__ out) {
    extern __shared__ float bl_lds[];
    float *wsh = bl_lds;
    if ((int)threadIdx.x < 2 * r + 1) wsh[threadIdx.x] = ws[threadIdx.x];
    __syncthreads();
    const int total = prefix[n];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = owner[t];
        const BatchImage im = imgs[i];
        const int local = t - prefix[i], tiles_x = ceil_div(im.W, TX);
        const int ty = local / tiles_x, tx = local - ty * tiles_x;
        bilateral_tile<TX, TY, RT>(bl_lds + 32, wsh, in + im.in_off, im.H, im.W, im.in_pitch, r, norm, pad, mode,
                               out + im.out_off, im.out_pitch, tx * TX, ty * TY);
    }
}

// Box sums from the batch's integral images: input and output both use the
// output offsets and pitches (the layout the integral images were written
// in).  Work items: BX_TX x BX_TY tiles.
__global__ __launch_bounds__(256) void batch_box_sum_kernel(const BatchImage *__restrict__ imgs, int n,
                                                            const int *__restrict__ prefix,
                                                            const int *__restrict__ owner,
                                                            const uint32_t *__restrict__ S, int ry, int rx,
                                                            uint32_t *__restrict__ out) {
    const int total = prefix[n];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = owner[t];
        const BatchImage im = imgs[i];
        const int local = t - prefix[i], tiles_x = ceil_div(im.W, BX_TX);
        const int ty = local / tiles_x, tx = local - ty * tiles_x;
        box_sum_tile(S + im.out_off, im.H, im.W, im.out_pitch, ry, rx, out + im.out_off, im.out_pitch, tx * BX_TX,
                     ty * BX_TY);
    }
}

// Integral image: the single pass of iimg_dev.h over the tiles of every
// image, a workgroup per tile in ticket order.
template <class T, int RPL>
__global__ __launch_bounds__(IM_THREADS) void batch_iimg_kernel(const IimgBatchTile *__restrict__ tiles, int total,
                                                                const typename T::In *__restrict__ in,
                                                                typename T::Out *__restrict__ out,
                                                                uint8_t *__restrict__ ws) {
    extern __shared__ __attribute__((aligned(16))) uint8_t biimg_lds[];
    auto &sm = *reinterpret_cast<IimgSmem<T, RPL> *>(biimg_lds);
    auto *ctl = reinterpret_cast<IimgCtl *>(ws);
    auto *words = reinterpret_cast<uint64_t *>(ws + IM_CTL_BYTES);
    iimg_begin(sm, ctl);
    for (int k = blockIdx.x; k < total; k += gridDim.x) {
        const IimgBatchTile t = tiles[k];
        const IimgTileRef d{t.tbase, t.H, t.W, t.in_pitch, t.out_pitch, t.tx, t.ty, t.ntx};
        const typename T::In *src = in + t.in_off;
        typename T::Out *dst = out + t.out_off;
        iimg_tile<T, RPL>(sm, d, src, iimg_rows16(src, t.in_pitch), dst, ctl, words);
    }
    iimg_finish(ctl, sm.tag);
}

}  // namespace

struct umi_batch_s {
    int n = 0;
    std::vector<BatchImage> host;
    BatchImage *imgs = nullptr;
    // device prefix arrays (n + 1 ints each) and their totals
    int *ftiles[NF] = {}, *btiles[NB] = {};
    int *fown[NF] = {}, *bown[NB] = {};  // owner tables: the image of every tile
    int ftotal[NF] = {}, btotal[NB] = {};
    int *xtiles = nullptr, *xown = nullptr;    // box-sum tiles and their images
    IimgBatchTile *idesc[NI] = {};            // integral-image tile descriptors, per configuration
    int nxtiles = 0;
    int nitiles[NI] = {};
    std::vector<void *> allocs;
};

namespace {

template <class V>
umi_status upload(umi_batch_s *b, const std::vector<V> &v, V **dst) {
    void *p = nullptr;
    UMI_TRY(hip_check(hipMalloc(&p, sizeof(V) * v.size()), "hipMalloc (batch tables)"));
    b->allocs.push_back(p);
    UMI_TRY(hip_check(hipMemcpy(p, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice), "hipMemcpy (batch tables)"));
    *dst = static_cast<V *>(p);
    return UMI_OK;
}

umi_status batch_check(umi_batch b, const void *in, const void *out, const char *who) {
    if (!b) return set_error(UMI_ERR_INVALID, "%s: null batch", who);
    if (!in || !out) return set_error(UMI_ERR_INVALID, "%s: null buffer", who);
    return UMI_OK;
}

int grid_for(int items) { return items < PERSIST_WG ? (items > 0 ? items : 1) : PERSIST_WG; }

template <int C>
umi_status launch_wfilter(umi_batch b, const float *in, const float *w, float init, float pad, int mode, float *out,
                          hipStream_t st) {
    hipLaunchKernelGGL((batch_wfilter_kernel<FTX[C], FTY[C]>), dim3(grid_for(b->ftotal[C])), dim3(256), 0, st, b->imgs,
                       b->n, b->ftiles[C], b->fown[C], in, w, init, pad, mode, out);
    return launch_check("batch_wfilter_kernel");
}

template <int C>
umi_status launch_bilateral(umi_batch b, const float *in, int r, const float *ws, float norm, float pad, int mode,
                            float *out, hipStream_t st) {
    const size_t lds = sizeof(float) * ((size_t)(BTX[C] + 2 * r) * (BTY[C] + 2 * r) + 32);
    if (r == 2)  // the 5 x 5 window: unrolled
        hipLaunchKernelGGL((batch_bilateral_kernel<BTX[C], BTY[C], 2>), dim3(grid_for(b->btotal[C])), dim3(256), lds,
                           st, b->imgs, b->n, b->btiles[C], b->bown[C], in, r, ws, norm, pad, mode, out);
    else
        hipLaunchKernelGGL((batch_bilateral_kernel<BTX[C], BTY[C], 0>), dim3(grid_for(b->btotal[C])), dim3(256), lds,
                           st, b->imgs, b->n, b->btiles[C], b->bown[C], in, r, ws, norm, pad, mode, out);
    return launch_check("batch_bilateral_kernel");
}

template <class T, int C>
umi_status launch_iimg_c(umi_batch b, const typename T::In *in, typename T::Out *out, uint8_t *ws, hipStream_t st) {
    constexpr int RPL = IRPL[C];
    const size_t lds = sizeof(IimgSmem<T, RPL>);
    static const hipError_t attr =
        hipFuncSetAttribute((const void *)batch_iimg_kernel<T, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int grid = iimg_grid(attr, b->nitiles[C], g_iimg_per_cu);
    if (grid < 1) return set_error(UMI_ERR_HIP, "batch integral image: launch setup failed");
    hipLaunchKernelGGL((batch_iimg_kernel<T, RPL>), dim3(grid), dim3(IM_THREADS), lds, st,
                       b->idesc[C], b->nitiles[C], in, out, ws);
    return launch_check("batch_iimg_kernel");
}

template <class T>
size_t batch_iimg_ws(umi_batch b) {
    size_t m = 0;
    for (int c = 0; c < NI; ++c) m = std::max(m, iimg_ws_bytes_tiles<typename T::Col>(b->nitiles[c], 16 * IRPL[c]));
    return m;
}

template <class T>
umi_status launch_iimg(umi_batch b, const typename T::In *in, typename T::Out *out, void *ws, size_t ws_bytes,
                       int config, hipStream_t st, const char *who) {
    UMI_TRY(batch_check(b, in, out, who));
    const size_t need = batch_iimg_ws<T>(b);
    if (!ws || ws_bytes < need) return set_error(UMI_ERR_INVALID, "%s: workspace %zu B < %zu B", who, ws_bytes, need);
    auto *w = static_cast<uint8_t *>(ws);
    switch (config < 0 ? 2 : config) {  // default: 64-row tiles, measured fastest on 256 images of 64..1024
        case 0: return launch_iimg_c<T, 0>(b, in, out, w, st);
        case 1: return launch_iimg_c<T, 1>(b, in, out, w, st);
        case 2: return launch_iimg_c<T, 2>(b, in, out, w, st);
        default: return set_error(UMI_ERR_INVALID, "%s: config %d", who, config);
    }
}

}  // namespace

extern "C" {

umi_status umi_batch_create(int n, const int *H, const int *W, const int64_t *in_off, const int *in_pitch,
                            const int64_t *out_off, const int *out_pitch, umi_batch *out) {
    if (!out) return set_error(UMI_ERR_INVALID, "umi_batch_create: null result pointer");
    *out = nullptr;
    if (n < 1 || !H || !W || !in_off || !in_pitch || !out_off || !out_pitch)
        return set_error(UMI_ERR_INVALID, "umi_batch_create: n=%d or null arrays", n);
    auto *b = new umi_batch_s;
    b->n = n;
    std::vector<int> fp[NF], bp[NB], xtiles(n + 1, 0);
    for (int c = 0; c < NF; ++c) fp[c].assign(n + 1, 0);
    for (int c = 0; c < NB; ++c) bp[c].assign(n + 1, 0);
    std::vector<IimgBatchTile> idesc[NI];
    for (int i = 0; i < n; ++i) {
        if (H[i] < 1 || W[i] < 1 || in_pitch[i] < W[i] || out_pitch[i] < W[i] || in_off[i] < 0 || out_off[i] < 0) {
            delete b;
            return set_error(UMI_ERR_INVALID, "umi_batch_create: image %d: H=%d W=%d pitches %d %d", i, H[i], W[i],
                             in_pitch[i], out_pitch[i]);
        }
        b->host.push_back({in_off[i], out_off[i], H[i], W[i], in_pitch[i], out_pitch[i]});
        for (int c = 0; c < NF; ++c) fp[c][i + 1] = fp[c][i] + ceil_div(W[i], FTX[c]) * ceil_div(H[i], FTY[c]);
        for (int c = 0; c < NB; ++c) bp[c][i + 1] = bp[c][i] + ceil_div(W[i], BTX[c]) * ceil_div(H[i], BTY[c]);
        xtiles[i + 1] = xtiles[i] + ceil_div(W[i], BX_TX) * ceil_div(H[i], BX_TY);
        for (int c = 0; c < NI; ++c) {
            const IimgGeom g = iimg_geom(H[i], W[i], 16 * IRPL[c]);
            const long long base = (long long)idesc[c].size();
            for (int ty = 0; ty < g.nty; ++ty)
                for (int tx = 0; tx < g.ntx; ++tx)
                    idesc[c].push_back({in_off[i], out_off[i], base, H[i], W[i], in_pitch[i], out_pitch[i], tx, ty,
                                        g.ntx, 0});
        }
    }
    umi_status st = upload(b, b->host, &b->imgs);
    for (int c = 0; c < NF && st == UMI_OK; ++c) {
        st = upload(b, fp[c], &b->ftiles[c]);
        b->ftotal[c] = fp[c][n];
    }
    for (int c = 0; c < NB && st == UMI_OK; ++c) {
        st = upload(b, bp[c], &b->btiles[c]);
        b->btotal[c] = bp[c][n];
    }
    // owner tables: item t of a list belongs to image owner[t]
    auto owners = [&](const std::vector<int> &pre, int per) {
        std::vector<int> o((size_t)pre[n] / per + 1, 0);
        for (int i = 0; i < n; ++i)
            for (int t = pre[i] / per; t < pre[i + 1] / per; ++t) o[t] = i;
        return o;
    };
    for (int c = 0; c < NF && st == UMI_OK; ++c) st = upload(b, owners(fp[c], 1), &b->fown[c]);
    for (int c = 0; c < NB && st == UMI_OK; ++c) st = upload(b, owners(bp[c], 1), &b->bown[c]);
    if (st == UMI_OK) st = upload(b, xtiles, &b->xtiles);
    if (st == UMI_OK) st = upload(b, owners(xtiles, 1), &b->xown);
    b->nxtiles = xtiles[n];
    for (int c = 0; c < NI && st == UMI_OK; ++c) {
        st = upload(b, idesc[c], &b->idesc[c]);
        b->nitiles[c] = (int)idesc[c].size();
    }
    if (st != UMI_OK) {
        for (void *p : b->allocs) (void)hipFree(p);  // the first failure is the status reported
        delete b;
        return st;
    }
    *out = b;
    return UMI_OK;
}

umi_status umi_batch_destroy(umi_batch b) {
    if (!b) return UMI_OK;
    umi_status rc = UMI_OK;
    for (void *p : b->allocs)
        if (rc == UMI_OK) rc = hip_check(hipFree(p), "hipFree (batch tables)");
        else (void)hipFree(p);
    delete b;
    return rc;
}

umi_status umi_batch_configs(int op, int *count) {
    if (!count) return set_error(UMI_ERR_INVALID, "umi_batch_configs: null pointer");
    if (op == 0) *count = NF;
    else if (op == 1) *count = NB;
    else if (op == 2) *count = NI;
    else return set_error(UMI_ERR_INVALID, "umi_batch_configs: op %d", op);
    return UMI_OK;
}

umi_status umi_batch_wfilter3x3_f32(umi_batch b, const float *in, const float *weights, float init, float pad,
                                    int mode, float *out, int config, umi_stream s) {
    UMI_TRY(batch_check(b, in, out, "umi_batch_wfilter3x3_f32"));
    if (!weights) return set_error(UMI_ERR_INVALID, "umi_batch_wfilter3x3_f32: null weights");
    if (mode != 0 && mode != 1) return set_error(UMI_ERR_INVALID, "umi_batch_wfilter3x3_f32: mode %d", mode);
    hipStream_t st = as_stream(s);
    switch (config < 0 ? 0 : config) {
        case 0: return launch_wfilter<0>(b, in, weights, init, pad, mode, out, st);
        case 1: return launch_wfilter<1>(b, in, weights, init, pad, mode, out, st);
        case 2: return launch_wfilter<2>(b, in, weights, init, pad, mode, out, st);
        case 3: return launch_wfilter<3>(b, in, weights, init, pad, mode, out, st);
        default: return set_error(UMI_ERR_INVALID, "umi_batch_wfilter3x3_f32: config %d", config);
    }
}

umi_status umi_batch_bilateral_f32(umi_batch b, const float *in, int radius, const float *spatial, float norm,
                                   float pad, int mode, float *out, int config, umi_stream s) {
    UMI_TRY(batch_check(b, in, out, "umi_batch_bilateral_f32"));
    if (!spatial) return set_error(UMI_ERR_INVALID, "umi_batch_bilateral_f32: null weights");
    if (radius < 0 || radius > BL_MAXR) return set_error(UMI_ERR_UNSUPPORTED, "umi_batch_bilateral_f32: radius %d", radius);
    if (mode != 0 && mode != 1) return set_error(UMI_ERR_INVALID, "umi_batch_bilateral_f32: mode %d", mode);
    hipStream_t st = as_stream(s);
    switch (config < 0 ? 0 : config) {
        case 0: return launch_bilateral<0>(b, in, radius, spatial, norm, pad, mode, out, st);
        case 1: return launch_bilateral<1>(b, in, radius, spatial, norm, pad, mode, out, st);
        case 2: return launch_bilateral<2>(b, in, radius, spatial, norm, pad, mode, out, st);
        case 3: return launch_bilateral<3>(b, in, radius, spatial, norm, pad, mode, out, st);
        default: return set_error(UMI_ERR_INVALID, "umi_batch_bilateral_f32: config %d", config);
    }
}

umi_status umi_batch_iimg_workspace_size(umi_batch b, int kind, size_t *bytes) {
    if (!b || !bytes) return set_error(UMI_ERR_INVALID, "umi_batch_iimg_workspace_size: null pointer");
    if (kind != 0 && kind != 2) return set_error(UMI_ERR_INVALID, "umi_batch_iimg_workspace_size: kind %d", kind);
    *bytes = kind == 0 ? batch_iimg_ws<IU8>(b) : batch_iimg_ws<IF32>(b);
    return UMI_OK;
}

umi_status umi_batch_iimg_u8_u32(umi_batch b, const uint8_t *in, uint32_t *out, void *workspace,
                                 size_t workspace_bytes, int config, umi_stream s) {
    return launch_iimg<IU8>(b, in, out, workspace, workspace_bytes, config, as_stream(s), "umi_batch_iimg_u8_u32");
}

umi_status umi_batch_iimg_f32(umi_batch b, const float *in, float *out, void *workspace, size_t workspace_bytes,
                              int config, umi_stream s) {
    return launch_iimg<IF32>(b, in, out, workspace, workspace_bytes, config, as_stream(s), "umi_batch_iimg_f32");
}

umi_status umi_batch_box_sum_u32(umi_batch b, const uint32_t *iimg, int ry, int rx, uint32_t *out, umi_stream s) {
    UMI_TRY(batch_check(b, iimg, out, "umi_batch_box_sum_u32"));
    if (ry < 0 || rx < 0) return set_error(UMI_ERR_INVALID, "umi_batch_box_sum_u32: radii %d %d", ry, rx);
    hipLaunchKernelGGL(batch_box_sum_kernel, dim3(grid_for(b->nxtiles)), dim3(256), 0, as_stream(s), b->imgs, b->n,
                       b->xtiles, b->xown, iimg, ry, rx, out);
    return launch_check("batch_box_sum_kernel");
}

}  // extern "C"
