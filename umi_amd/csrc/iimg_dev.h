// Device-side bodies of the integral image, shared by the one-image kernels
// (iimg.hip) and the batch kernels (batch.hip).  Semantics:
// main/iimg.cu:11-18,37-82,84-112 (johnjohnlin/UMI): out[y][x] is the sum of
// in[y'][x'] over y' <= y, x' <= x.
//
// Tiled reduce-then-scan, three passes over tiles of IM_TH x IM_TW:
//  A (totals)  per tile: the row totals of its IM_TH rows, L[y][tx]; the
//              within-tile prefix over x of its column totals, C[ty][x]; the
//              tile total, T[ty][tx].  Reads the input once.
//  B (carries) per row y: L becomes the exclusive prefix over tx (the sum of
//              the row left of the tile); per column x: C becomes the
//              exclusive prefix over ty (the sum, above the tile, of the
//              tile-local row prefixes); per tile column tx: K[ty][tx] = the
//              sum of T over tiles above and to the left.
//  C (output)  per tile: the within-tile row prefix lrp(y, x), then down each
//              column acc = K + C + sum over the tile's rows so far of
//              (L[y] + lrp(y, x)).  Reads the input again, writes the output
//              once.
// out(y, x) = K(ty, tx) + C(ty, x) + sum_{y' = ty0 .. y} (L(y', tx) + lrp(y', x))
// is the summed-area table split into the part above-left of the tile, above
// it, left of it and inside it.  Integer forms are exact (mod 2^32) in any
// order; the float form sums rows within a tile in fp32 and everything
// vertical and every carry in fp64.
#pragma once
#include "common.h"

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
constexpr int IM_TW = 256, IM_TH = 16;  // thread (r = t / 16, s = t % 16) holds row r, columns 16 s .. 16 s + 15

struct IimgGeom {
    int ntx, nty;
};
__host__ __device__ inline IimgGeom iimg_geom(int H, int W) { return {ceil_div(W, IM_TW), ceil_div(H, IM_TH)}; }

// Workspace of one image, in Col elements: L[H][ntx], C[nty][W], T[nty][ntx], K[nty][ntx].
__host__ __device__ inline long long iimg_ws_elems(int H, int W) {
    const IimgGeom g = iimg_geom(H, W);
    return (long long)H * g.ntx + (long long)g.nty * W + 2LL * g.nty * g.ntx;
}

template <class T>
struct IimgWs {
    typename T::Col *L, *C, *T_, *K;
};
template <class T>
__device__ inline IimgWs<T> iimg_ws(typename T::Col *ws, int H, int W) {
    const IimgGeom g = iimg_geom(H, W);
    IimgWs<T> w;
    w.L = ws;
    w.C = w.L + (size_t)H * g.ntx;
    w.T_ = w.C + (size_t)g.nty * W;
    w.K = w.T_ + (size_t)g.nty * g.ntx;
    return w;
}

// True when every row of `in` starts 16-byte aligned.
template <class In>
__device__ inline bool iimg_rows16(const In *in, int pitch) {
    return ((reinterpret_cast<uintptr_t>(in) | ((uintptr_t)pitch * sizeof(In))) & 15) == 0;
}

// The thread's 16 elements of row y (zeros past the image).
template <class T>
__device__ inline void iimg_load16(typename T::Row (&v)[16], const typename T::In *__restrict__ in, int H, int W,
                                   int ip, bool rows16, int y, int x) {
    using In = typename T::In;
    using Row = typename T::Row;
    if (y < H && rows16 && x + 16 <= W) {
        constexpr int NQ = 16 * (int)sizeof(In) / 16;
        uint4 q[NQ];
        const uint4 *src = reinterpret_cast<const uint4 *>(in + (size_t)y * ip + x);
#pragma unroll
        for (int i = 0; i < NQ; ++i) q[i] = src[i];
        const In *e = reinterpret_cast<const In *>(q);
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = (Row)e[j];
    } else if constexpr (sizeof(In) == 1) {
        // bytes of a row that is not 16-byte aligned: the five aligned dwords
        // around them, each clamped to the dword holding the row's last byte
        // (so every load holds at least one byte of the row), then v_alignbyte;
        // bytes past W and rows past H are zero
        uint32_t w[4] = {0, 0, 0, 0};
        if (y < H && x < W) {
            const uintptr_t p = reinterpret_cast<uintptr_t>(in + (size_t)y * ip + x);
            const uintptr_t last = reinterpret_cast<uintptr_t>(in + (size_t)y * ip + W - 1) & ~(uintptr_t)3;
            const uintptr_t a0 = p & ~(uintptr_t)3;
            const uint32_t sh = (uint32_t)(p & 3);
            uint32_t d[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) d[k] = *reinterpret_cast<const uint32_t *>(min(a0 + 4 * k, last));
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const Row e = (Row)__builtin_amdgcn_ubfe(w[j >> 2], 8 * (j & 3), 8);
            v[j] = x + j < W ? e : (Row)0;
        }
    } else if constexpr (sizeof(In) == 4) {
        // 4-byte elements of a row that is not 16-byte aligned: the five
        // aligned 16-byte chunks around them, each clamped to the chunk
        // holding the row's last element, then a per-lane select of the
        // starting element; elements past W and rows past H are zero
        uint32_t w[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) w[j] = 0;
        if (y < H && x < W) {
            const uintptr_t p = reinterpret_cast<uintptr_t>(in + (size_t)y * ip + x);
            const uintptr_t last = reinterpret_cast<uintptr_t>(in + (size_t)y * ip + W - 1) & ~(uintptr_t)15;
            const uintptr_t a0 = p & ~(uintptr_t)15;
            const int e = (int)((p & 15) >> 2);
            uint32_t dw[20];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const uint4 q = *reinterpret_cast<const uint4 *>(min(a0 + 16 * k, last));
                dw[4 * k] = q.x;
                dw[4 * k + 1] = q.y;
                dw[4 * k + 2] = q.z;
                dw[4 * k + 3] = q.w;
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) w[j] = e == 0 ? dw[j] : e == 1 ? dw[j + 1] : e == 2 ? dw[j + 2] : dw[j + 3];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const Row el = (Row)__builtin_bit_cast(In, w[j]);
            v[j] = x + j < W ? el : (Row)0;
        }
    } else {
        // clamped indices keep every load inside the image; the select zeroes the rest
        const In *row = in + (size_t)min(y, H - 1) * ip;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const Row e = (Row)row[min(x + j, W - 1)];
            v[j] = (y < H && x + j < W) ? e : (Row)0;
        }
    }
}

// Inclusive prefix over the 256 threads of the workgroup (tmp: 4 values of
// LDS).  Contains a barrier; the caller separates reuses of tmp by another.
template <class V>
__device__ inline V iimg_block_scan(V v, V *tmp) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const V n = __shfl_up(v, d);
        if (lane >= d) v += n;
    }
    if (lane == 63) tmp[wv] = v;
    __syncthreads();
    V off = 0;
    for (int i = 0; i < wv; ++i) off += tmp[i];
    return v + off;
}

// Pass A for tile (tx, ty).  lds: IM_TH x IM_TW Row values; tmp: 4 Col.
template <class T>
__device__ inline void iimg_tile_totals(typename T::Row *lds, typename T::Col *tmp,
                                        const typename T::In *__restrict__ in, int H, int W, int ip, bool rows16,
                                        int tx, int ty, const IimgWs<T> &w) {
    using Row = typename T::Row;
    using Col = typename T::Col;
    const int t = threadIdx.x, r = t >> 4, s = t & 15;
    const int ntx = iimg_geom(H, W).ntx;
    const int y = ty * IM_TH + r, x = tx * IM_TW + 16 * s;
    Row v[16];
    iimg_load16<T>(v, in, H, W, ip, rows16, y, x);
    Row rs = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) rs += v[j];
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) rs += __shfl_xor(rs, m, 16);
    if (s == 0 && y < H) w.L[(size_t)y * ntx + tx] = (Col)rs;
#pragma unroll
    for (int j = 0; j < 16; ++j) lds[r * IM_TW + 16 * s + j] = v[j];
    __syncthreads();
    Col c = 0;
#pragma unroll
    for (int rr = 0; rr < IM_TH; ++rr) c += (Col)lds[rr * IM_TW + t];
    const Col incl = iimg_block_scan(c, tmp);
    const int xc = tx * IM_TW + t;
    if (xc < W) w.C[(size_t)ty * W + xc] = incl;
    if (t == IM_THREADS - 1) w.T_[ty * ntx + tx] = incl;  // columns past W hold zeros
    __syncthreads();
}

// Pass B items of one image, in whole waves: 64 * ntx corner items (a wave
// per tile column; lanes over tile rows), H row items padded to a multiple
// of 64, then 64 * ceil(W / 8) column items (a wave per 8 columns; lane =
// column c + 8 * segment q, the 8 segments splitting the tile rows).
__host__ __device__ inline int iimg_carry_items(int H, int W) {
    return 64 * iimg_geom(H, W).ntx + round_up(H, 64) + 64 * ceil_div(W, 8);
}

// Pass B, item k of [0, iimg_carry_items(H, W)).  Loads are issued in
// batches ahead of the dependent sums.
template <class T>
__device__ inline void iimg_carry_item(int H, int W, int k, const IimgWs<T> &w) {
    using Col = typename T::Col;
    const IimgGeom g = iimg_geom(H, W);
    constexpr int U = 16;
    const int nc = 64 * g.ntx, nr = round_up(H, 64);
    const int lane = k & 63;
    if (k < nc) {
        // K[ty][tx] = sum over ty' < ty of P(ty'), P(ty') = sum over tx' < tx of T[ty'][tx']
        const int tx = k >> 6;
        const Col *__restrict__ Tp = w.T_;
        Col *__restrict__ Kp = w.K;
        Col off = 0;
        for (int ty0 = 0; ty0 < g.nty; ty0 += 64) {
            const int ty = ty0 + lane;
            Col p = 0;
            if (ty < g.nty)
                for (int i = 0; i < tx; ++i) p += Tp[ty * g.ntx + i];
            Col incl = p;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const Col n = __shfl_up(incl, d);
                if (lane >= d) incl += n;
            }
            Col excl = __shfl_up(incl, 1);
            if (lane == 0) excl = 0;
            if (ty < g.nty) Kp[ty * g.ntx + tx] = off + excl;
            off += __shfl(incl, 63);
        }
    } else if (k < nc + nr) {
        const int y = k - nc;
        if (y >= H) return;
        Col *__restrict__ p = w.L + (size_t)y * g.ntx;
        Col run = 0;
        for (int i0 = 0; i0 < g.ntx; i0 += U) {
            Col v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = i0 + u < g.ntx ? p[i0 + u] : (Col)0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (i0 + u < g.ntx) {
                    p[i0 + u] = run;
                    run += v[u];
                }
        }
    } else {
        // C[ty][x] becomes the exclusive prefix over ty: segment totals, a
        // scan of the 8 segments across lanes, then the segments rewritten
        const int q = lane >> 3, x = ((k - nc - nr) >> 6) * 8 + (lane & 7);
        const int seg = ceil_div(g.nty, 8), j0 = q * seg, j1 = min(g.nty, j0 + seg);
        Col *__restrict__ p = w.C + min(x, W - 1);
        const bool own = x < W;
        Col tot = 0;
        // segments of at most U tile rows (every height up to 16 * 8 * U) keep
        // their values in registers between the two passes: one load round trip
        Col keep[U];
#pragma unroll
        for (int u = 0; u < U; ++u) keep[u] = 0;
        const bool one = seg <= U;  // uniform: then every segment fits one batch of loads
        for (int j = j0; j < j1; j += U) {
            Col v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = own && j + u < j1 ? p[(size_t)(j + u) * W] : (Col)0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                tot += v[u];
                keep[u] = v[u];
            }
        }
        Col incl = tot;
#pragma unroll
        for (int d = 8; d < 64; d <<= 1) {
            const Col n = __shfl_up(incl, d);
            if (lane >= d) incl += n;
        }
        Col run = __shfl_up(incl, 8);
        if (q == 0) run = 0;
        if (one) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (own && j0 + u < j1) {
                    p[(size_t)(j0 + u) * W] = run;
                    run += keep[u];
                }
            return;
        }
        for (int j = j0; j < j1; j += U) {
            Col v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = own && j + u < j1 ? p[(size_t)(j + u) * W] : (Col)0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (own && j + u < j1) {
                    p[(size_t)(j + u) * W] = run;
                    run += v[u];
                }
        }
    }
}

// Pass C for tile (tx, ty).  lds: IM_TH x IM_TW Row values; rl: IM_TH Col.
template <class T>
__device__ inline void iimg_tile_out(typename T::Row *lds, typename T::Col *rl, const typename T::In *__restrict__ in,
                                     int H, int W, int ip, bool rows16, int tx, int ty, const IimgWs<T> &w,
                                     typename T::Out *__restrict__ out, int op) {
    using Row = typename T::Row;
    using Col = typename T::Col;
    const int t = threadIdx.x, r = t >> 4, s = t & 15;
    const int ntx = iimg_geom(H, W).ntx;
    const int y = ty * IM_TH + r, x = tx * IM_TW + 16 * s;
    Row v[16];
    iimg_load16<T>(v, in, H, W, ip, rows16, y, x);
#pragma unroll
    for (int j = 1; j < 16; ++j) v[j] += v[j - 1];
    Row incl = v[15];
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
        const Row n = __shfl_up(incl, d, 16);
        if (s >= d) incl += n;
    }
    Row excl = __shfl_up(incl, 1, 16);
    if (s == 0) excl = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) lds[r * IM_TW + 16 * s + j] = excl + v[j];
    if (t < IM_TH) {
        const int yy = ty * IM_TH + t;
        rl[t] = yy < H ? w.L[(size_t)yy * ntx + tx] : (Col)0;
    }
    __syncthreads();
    const int xc = tx * IM_TW + t;
    Col acc = w.K[ty * ntx + tx] + (xc < W ? w.C[(size_t)ty * W + xc] : (Col)0);
    const int nr = min(IM_TH, H - ty * IM_TH);
    for (int rr = 0; rr < nr; ++rr) {
        acc += rl[rr] + (Col)lds[rr * IM_TW + t];
        if (xc < W) out[(size_t)(ty * IM_TH + rr) * op + xc] = (typename T::Out)acc;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- box sums
// A consumer of the integral image: the sum over the (2 ry + 1) x (2 rx + 1)
// window around (y, x), clipped to the image, from the four corners of the
// inclusive integral image S:
//   out = S(y2, x2) - S(y1 - 1, x2) - S(y2, x1 - 1) + S(y1 - 1, x1 - 1),
// y1 = max(y - ry, 0), y2 = min(y + ry, H - 1) (x likewise), S(-1, .) =
// S(., -1) = 0; unsigned arithmetic, exact mod 2^32 (as the integral image).
// A tile of BX_TX x BX_TY outputs per 256 threads: thread (t % 64) takes a
// column, rows t / 64, + 4, ...; the four corner loads of a row are issued
// together and a wave's loads cover 64 consecutive columns.
constexpr int BX_TX = 64, BX_TY = 16;

__device__ inline void box_sum_tile(const uint32_t *__restrict__ S, int H, int W, int ip, int ry, int rx,
                                    uint32_t *__restrict__ out, int op, int x0, int y0) {
    const int x = x0 + (int)(threadIdx.x & 63);
    if (x >= W) return;
    const int xa = max(x - rx, 0) - 1, xb = min(x + rx, W - 1);
    const int xs = max(xa, 0);  // a valid column for the unconditional load
    for (int r = (int)(threadIdx.x >> 6); r < BX_TY; r += 4) {
        const int y = y0 + r;
        if (y >= H) return;
        const int ya = max(y - ry, 0) - 1, yb = min(y + ry, H - 1);
        const uint32_t *rb = S + (size_t)yb * ip, *ra = S + (size_t)max(ya, 0) * ip;
        const uint32_t a = rb[xb], b = rb[xs], c = ra[xb], d = ra[xs];
        const uint32_t v = a - (xa >= 0 ? b : 0u) - (ya >= 0 ? c : 0u) + (xa >= 0 && ya >= 0 ? d : 0u);
        out[(size_t)y * op + x] = v;
    }
}

}  // namespace umi
