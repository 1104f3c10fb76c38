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

#include <vector>

namespace {
using namespace umi;

struct BatchImage {
    long long in_off, out_off;  // element offsets from the base pointers
    int H, W, in_pitch, out_pitch;
};

// One integral-image tile of the batch: everything its passes need, so that
// a tile costs one (wave-uniform) descriptor load instead of a chain of
// owner -> image -> prefix -> workspace-offset loads.
struct IimgTile {
    long long in_off, out_off, ws_off;
    int H, W, in_pitch, out_pitch, tx, ty;
};

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

// Integral image (the passes of iimg_dev.h).  Passes A and C: a workgroup
// per tile of every image at a time; pass B: one thread per carry item (rows,
// columns, tile columns) of every image.

// One wave of carry items (an image's items are whole waves): the image's
// workspace offset and size, and the wave's first item within the image.
struct CarryWave {
    long long ws_off;
    int H, W, k0, pad;
};

// Persistent loop over the tile descriptors: the next tile's descriptor is
// loaded before the current tile is processed.
template <class F>
__device__ inline void for_each_tile(const IimgTile *__restrict__ tiles, int total, F &&body) {
    int t = blockIdx.x;
    if (t >= total) return;
    IimgTile d = tiles[t];
    for (;;) {
        const int tn = t + (int)gridDim.x;
        const IimgTile dn = tiles[tn < total ? tn : t];
        body(d);
        if (tn >= total) return;
        t = tn;
        d = dn;
    }
}

template <class T>
__global__ __launch_bounds__(IM_THREADS) void batch_iimg_totals_kernel(const IimgTile *__restrict__ tiles, int total,
                                                                       const typename T::In *__restrict__ in,
                                                                       typename T::Col *__restrict__ ws) {
    __shared__ typename T::Row lds[IM_TH * IM_TW];
    __shared__ typename T::Col tmp[4];
    for_each_tile(tiles, total, [&](const IimgTile &d) {
        const typename T::In *src = in + d.in_off;
        iimg_tile_totals<T>(lds, tmp, src, d.H, d.W, d.in_pitch, iimg_rows16(src, d.in_pitch), d.tx, d.ty,
                            iimg_ws<T>(ws + d.ws_off, d.H, d.W));
    });
}

template <class T>
__global__ __launch_bounds__(256) void batch_iimg_carry_kernel(const CarryWave *__restrict__ waves, int total,
                                                               typename T::Col *__restrict__ ws) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const CarryWave cw = waves[t >> 6];  // one (wave-uniform) load instead of owner -> image -> prefix
        iimg_carry_item<T>(cw.H, cw.W, cw.k0 + (t & 63), iimg_ws<T>(ws + cw.ws_off, cw.H, cw.W));
    }
}

template <class T>
__global__ __launch_bounds__(IM_THREADS) void batch_iimg_out_kernel(const IimgTile *__restrict__ tiles, int total,
                                                                    const typename T::In *__restrict__ in,
                                                                    typename T::Col *__restrict__ ws,
                                                                    typename T::Out *__restrict__ out) {
    __shared__ typename T::Row lds[IM_TH * IM_TW];
    __shared__ typename T::Col rl[IM_TH];
    for_each_tile(tiles, total, [&](const IimgTile &d) {
        const typename T::In *src = in + d.in_off;
        iimg_tile_out<T>(lds, rl, src, d.H, d.W, d.in_pitch, iimg_rows16(src, d.in_pitch), d.tx, d.ty,
                         iimg_ws<T>(ws + d.ws_off, d.H, d.W), out + d.out_off, d.out_pitch);
    });
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
    CarryWave *cwaves = nullptr;               // integral image carry items, per wave of 64
    int *xtiles = nullptr, *xown = nullptr;    // box-sum tiles and their images
    IimgTile *idesc = nullptr;                 // integral-image tile descriptors
    int nxtiles = 0;
    int nitiles = 0, nicarry = 0;
    long long ws_elems = 0;
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

template <class T>
umi_status launch_iimg(umi_batch b, const typename T::In *in, typename T::Out *out, void *ws, size_t ws_bytes,
                       hipStream_t st, const char *who) {
    UMI_TRY(batch_check(b, in, out, who));
    const size_t need = (size_t)b->ws_elems * sizeof(typename T::Col);
    if (need && (!ws || ws_bytes < need)) return set_error(UMI_ERR_INVALID, "%s: workspace %zu B < %zu B", who, ws_bytes, need);
    auto *w = static_cast<typename T::Col *>(ws);
    hipLaunchKernelGGL(batch_iimg_totals_kernel<T>, dim3(grid_for(b->nitiles)), dim3(IM_THREADS), 0, st, b->idesc,
                       b->nitiles, in, w);
    UMI_TRY(launch_check("batch_iimg_totals_kernel"));
    hipLaunchKernelGGL(batch_iimg_carry_kernel<T>, dim3(grid_for(ceil_div(b->nicarry, 256))), dim3(256), 0, st,
                       b->cwaves, b->nicarry, w);
    UMI_TRY(launch_check("batch_iimg_carry_kernel"));
    hipLaunchKernelGGL(batch_iimg_out_kernel<T>, dim3(grid_for(b->nitiles)), dim3(IM_THREADS), 0, st, b->idesc,
                       b->nitiles, in, w, out);
    UMI_TRY(launch_check("batch_iimg_out_kernel"));
    return UMI_OK;
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
    std::vector<int> fp[NF], bp[NB], itiles(n + 1, 0), icarry(n + 1, 0), xtiles(n + 1, 0);
    for (int c = 0; c < NF; ++c) fp[c].assign(n + 1, 0);
    for (int c = 0; c < NB; ++c) bp[c].assign(n + 1, 0);
    std::vector<long long> wso(n, 0);
    std::vector<IimgTile> idesc;
    long long ws = 0;
    for (int i = 0; i < n; ++i) {
        if (H[i] < 1 || W[i] < 1 || in_pitch[i] < W[i] || out_pitch[i] < W[i] || in_off[i] < 0 || out_off[i] < 0) {
            delete b;
            return set_error(UMI_ERR_INVALID, "umi_batch_create: image %d: H=%d W=%d pitches %d %d", i, H[i], W[i],
                             in_pitch[i], out_pitch[i]);
        }
        const IimgGeom g = iimg_geom(H[i], W[i]);
        b->host.push_back({in_off[i], out_off[i], H[i], W[i], in_pitch[i], out_pitch[i]});
        for (int c = 0; c < NF; ++c) fp[c][i + 1] = fp[c][i] + ceil_div(W[i], FTX[c]) * ceil_div(H[i], FTY[c]);
        for (int c = 0; c < NB; ++c) bp[c][i + 1] = bp[c][i] + ceil_div(W[i], BTX[c]) * ceil_div(H[i], BTY[c]);
        itiles[i + 1] = itiles[i] + g.ntx * g.nty;
        xtiles[i + 1] = xtiles[i] + ceil_div(W[i], BX_TX) * ceil_div(H[i], BX_TY);
        icarry[i + 1] = icarry[i] + iimg_carry_items(H[i], W[i]);
        wso[i] = ws;
        for (int ty = 0; ty < g.nty; ++ty)
            for (int tx = 0; tx < g.ntx; ++tx)
                idesc.push_back({in_off[i], out_off[i], ws, H[i], W[i], in_pitch[i], out_pitch[i], tx, ty});
        ws += iimg_ws_elems(H[i], W[i]);
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
    {
        std::vector<CarryWave> cw;
        for (int i = 0; i < n; ++i)
            for (int k = icarry[i]; k < icarry[i + 1]; k += 64) cw.push_back({wso[i], H[i], W[i], k - icarry[i], 0});
        if (st == UMI_OK) st = upload(b, cw, &b->cwaves);
    }
    if (st == UMI_OK) st = upload(b, xtiles, &b->xtiles);
    if (st == UMI_OK) st = upload(b, owners(xtiles, 1), &b->xown);
    b->nxtiles = xtiles[n];
    if (st == UMI_OK) st = upload(b, idesc, &b->idesc);
    b->nitiles = itiles[n];
    b->nicarry = icarry[n];
    b->ws_elems = ws;
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
    *bytes = (size_t)b->ws_elems * (kind == 0 ? sizeof(uint32_t) : sizeof(double));
    return UMI_OK;
}

umi_status umi_batch_iimg_u8_u32(umi_batch b, const uint8_t *in, uint32_t *out, void *workspace,
                                 size_t workspace_bytes, umi_stream s) {
    return launch_iimg<IU8>(b, in, out, workspace, workspace_bytes, as_stream(s), "umi_batch_iimg_u8_u32");
}

umi_status umi_batch_iimg_f32(umi_batch b, const float *in, float *out, void *workspace, size_t workspace_bytes,
                              umi_stream s) {
    return launch_iimg<IF32>(b, in, out, workspace, workspace_bytes, as_stream(s), "umi_batch_iimg_f32");
}

umi_status umi_batch_box_sum_u32(umi_batch b, const uint32_t *iimg, int ry, int rx, uint32_t *out, umi_stream s) {
    UMI_TRY(batch_check(b, iimg, out, "umi_batch_box_sum_u32"));
    if (ry < 0 || rx < 0) return set_error(UMI_ERR_INVALID, "umi_batch_box_sum_u32: radii %d %d", ry, rx);
    hipLaunchKernelGGL(batch_box_sum_kernel, dim3(grid_for(b->nxtiles)), dim3(256), 0, as_stream(s), b->imgs, b->n,
                       b->xtiles, b->xown, iimg, ry, rx, out);
    return launch_check("batch_box_sum_kernel");
}

}  // extern "C"
