// Device-side bodies of the integral image, shared by the one-image kernel
// (iimg.hip) and the batch kernel (batch.hip).  Semantics:
// main/iimg.cu:11-18,37-82,84-112 (johnjohnlin/UMI): out[y][x] is the sum of
// in[y'][x'] over y' <= y, x' <= x.
//
// One launch, input read once.  The image is cut into tiles of TH x 256
// (TH = 16 * RPL, RPL rows per lane), numbered row-major within an image,
// images in order; the grid walks them in that order (a workgroup per tile,
// or a persistent grid taking tiles b, b + G, ...), so a tile waits only for
// tiles of smaller number, which have started before it.  Per tile, with
// (ty, tx) its tile row and column and y0 its first row:
//   1. load (lane: RPL rows x 16 columns), in-tile row prefixes p(y, x) by
//      16-lane DPP scans into LDS; publish the row totals rt(y);
//   2. L(y) = sum_{tx' < tx} rt(ty, tx', y), the sum of row y left of the
//      tile, and its prefix over the tile's rows LP(y);
//   3. decoupled look-back down the tile column: publish the aggregate
//      A(x) = sum over the tile's rows of (L(y) + p(y, x)), walk up through
//      the tiles above adding aggregates until one shows its inclusive value
//      I(x) = S(y0' + TH - 1, x) (the integral image's value on its last row),
//      giving U(x) = S(y0 - 1, x); publish I(x) = U(x) + A(x) over A(x);
//   4. out(y, x) = U(x) + LP(y) + sum_{y'=y0..y} p(y', x): a thread per
//      column walks the tile's rows in LDS (coalesced 4-byte stores).
// A published word is (tag << 33 | state << 32 | value), state 0 = aggregate
// (or row total), 1 = inclusive.  The tag identifies the call (the
// workspace's epoch + 1), so a word left by an earlier call never passes for
// a current one and no fence orders the value with a flag.  fp64 values are
// two words, one per half, each with tag and state.  The stores and loads are
// relaxed atomics at agent scope, coherent across the XCDs' L2 caches.  The
// last workgroup to finish (counted in groups on separate cache lines)
// resets the counters and advances the epoch, so the workspace needs zeroing
// once, before its first use, and no launch clears anything (graph replays
// included).  Every wait is bounded by the wall clock (IM_WAIT_TICKS); a wait
// that runs out sets a flag (umi_iimg_workspace_status) and the kernel still
// ends.
// Integer forms are exact (mod 2^32) in any order; the float form sums rows
// within a tile in fp32 and everything vertical and every carry in fp64.
#pragma once
#include <algorithm>
#include "common.h"
#include "wave.h"

namespace umi {

// In: element; Row: in-tile row sums; Col: vertical sums and carries; Out:
// output; Lds: in-tile row prefixes as held in LDS (u8 rows of 256 sum to
// at most 65280, so 16 bits hold them).
template <typename Tin, typename Trow, typename Tcol, typename Tout, typename Tlds>
struct IimgTypes {
    using In = Tin;
    using Row = Trow;
    using Col = Tcol;
    using Out = Tout;
    using Lds = Tlds;
};
using IU8 = IimgTypes<uint8_t, uint32_t, uint32_t, uint32_t, uint16_t>;
using II32 = IimgTypes<int32_t, uint32_t, uint32_t, int32_t, uint32_t>;
using IF32 = IimgTypes<float, float, double, float, float>;

// Element forms: uint8 -> uint32 and int32 -> int32 (exact, wrapping),
// float -> float with fp32 row sums inside a tile and fp64 vertical sums.
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

// ------------------------------------------------------------ single pass
constexpr int IM_THREADS = 256, IM_TW = 256;
constexpr long long IM_WAIT_TICKS = 20000000;  // 0.2 s of the 100 MHz wall clock per wait

// Control block at the start of the workspace: the epoch, the timeout flag,
// the group counter and the per-group completion counters, each counter on
// its own 128-byte line.
constexpr int IM_GROUP = 32, IM_MAX_GROUPS = 64, IM_LINE_U32 = 32;
constexpr int IM_MAX_GRID = IM_GROUP * IM_MAX_GROUPS;
struct IimgCtl {
    uint32_t epoch, timeout, pad0[IM_LINE_U32 - 2];
    uint32_t top, pad1[IM_LINE_U32 - 1];
    uint32_t grp[IM_MAX_GROUPS][IM_LINE_U32];
};
constexpr size_t IM_CTL_BYTES = sizeof(IimgCtl);
static_assert(IM_CTL_BYTES % 256 == 0, "control block keeps the words 256-byte aligned");

// Grid of a single-pass launch over `tiles` tiles: per_cu > 0 gives a
// persistent grid of per_cu workgroups per CU; 0 a workgroup per tile.
// `attr` is the status of setting the kernel's LDS limit (done once per
// kernel by the caller); < 1 on failure.
inline int iimg_grid(hipError_t attr, int tiles, int per_cu) {
    if (attr != hipSuccess) return 0;
    if (per_cu <= 0) return tiles;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return std::min(tiles, per_cu * cus);
}
extern int g_iimg_per_cu;  // launch shape of the single-pass kernels (umi_iimg_set_grid)

// Published words per value: 1 (32-bit sums) or 2 (fp64 halves).
template <class Col>
constexpr int pub_words() { return sizeof(Col) == 8 ? 2 : 1; }

struct IimgGeom {
    int ntx, nty;
};
__host__ __device__ inline IimgGeom iimg_geom(int H, int W, int th) { return {ceil_div(W, IM_TW), ceil_div(H, th)}; }

// Words published per tile: the column words (aggregate, then inclusive) of
// 256 columns, then the row totals of th rows.
template <class Col>
__host__ __device__ inline long long iimg_tile_words(int th) {
    return (long long)(IM_TW + th) * pub_words<Col>();
}

// Workspace bytes for `tiles` tiles of height th.
template <class Col>
__host__ __device__ inline size_t iimg_ws_bytes_tiles(long long tiles, int th) {
    const size_t b = IM_CTL_BYTES + (size_t)tiles * (size_t)iimg_tile_words<Col>(th) * sizeof(uint64_t);
    return (b + 15) & ~(size_t)15;
}

__device__ inline uint64_t pub_word(uint32_t tag, uint32_t state, uint32_t half) {
    return ((uint64_t)tag << 33) | ((uint64_t)state << 32) | half;
}
__device__ inline void pub_put(uint64_t *p, uint32_t tag, uint32_t state, uint32_t v) {
    __hip_atomic_store(p, pub_word(tag, state, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void pub_put(uint64_t *p, uint32_t tag, uint32_t state, double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    __hip_atomic_store(p, pub_word(tag, state, (uint32_t)b), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, pub_word(tag, state, (uint32_t)(b >> 32)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t pub_ld(const uint64_t *p) {
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One published value: its words loaded once, then polled until they carry
// the call's tag and agree on the state (or the wait runs out, which is
// recorded in ctl->timeout).
template <class Col>
struct PubVal {
    uint64_t w[pub_words<Col>()];
    __device__ inline void load(const uint64_t *p) {
#pragma unroll
        for (int k = 0; k < pub_words<Col>(); ++k) w[k] = pub_ld(p + k);
    }
    __device__ inline bool ready(uint32_t tag) const {
        bool r = (uint32_t)(w[0] >> 33) == tag;
#pragma unroll
        for (int k = 1; k < pub_words<Col>(); ++k) r = r && (w[k] >> 32) == (w[0] >> 32);
        return r;
    }
    __device__ inline uint32_t state() const { return (uint32_t)(w[0] >> 32) & 1u; }
    __device__ inline Col value() const {
        if constexpr (sizeof(Col) == 8)
            return __builtin_bit_cast(Col, (w[0] & 0xffffffffu) | (w[1] << 32));
        else
            return (Col)(uint32_t)w[0];
    }
    __device__ inline void wait(const uint64_t *p, uint32_t tag, IimgCtl *ctl) {
        if (ready(tag)) return;
        const uint64_t t0 = wall_clock64();
        do {
            __builtin_amdgcn_s_sleep(2);
            load(p);
        } while (!ready(tag) && (long long)(wall_clock64() - t0) < IM_WAIT_TICKS);
        if (!ready(tag)) __hip_atomic_store(&ctl->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Sum of n published values at p + i * stride (i < n), loads issued in
// batches of 16 ahead of the waits.
template <class Col>
__device__ inline Col pub_sum(const uint64_t *p, long long stride, int n, uint32_t tag, IimgCtl *ctl) {
    constexpr int B = 16;
    Col acc = 0;
    for (int i0 = 0; i0 < n; i0 += B) {
        PubVal<Col> v[B];
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (i0 + u < n) v[u].load(p + (i0 + u) * stride);
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (i0 + u < n) {
                v[u].wait(p + (i0 + u) * stride, tag, ctl);
                acc += v[u].value();
            }
    }
    return acc;
}

// Decoupled look-back over n predecessors at p - stride, p - 2 stride, ...
// (nearest first): the sum of their aggregates up to and including the
// nearest inclusive value.  prefetch() loads the nearest one ahead of the
// work that precedes resolve() (it usually already shows its inclusive value
// or soon does); further predecessors are read one at a time, LB_WIN = 1:
// wider windows (2, 4, 8, 16 measured on the 1080p image and the 256-image
// batch) read words a nearer tile's inclusive value makes unnecessary, and
// the extra traffic cost more than the round trips they save (1080p u8,
// 32-row tiles: 11.5 us at 1, 12.0 at 2, 13.0 at 8, 13.8 at 16).
template <class Col>
struct Lookback {
    static constexpr int LB_WIN = 1;
    PubVal<Col> v[LB_WIN];
    __device__ inline void prefetch(const uint64_t *p, long long stride, int n) {
        if (n > 0) v[0].load(p - stride);
    }
    __device__ inline Col resolve(const uint64_t *p, long long stride, int n, uint32_t tag, IimgCtl *ctl) {
        if (n <= 0) return 0;
        v[0].wait(p - stride, tag, ctl);
        Col acc = v[0].value();
        if (v[0].state()) return acc;
        for (int i0 = 1; i0 < n; i0 += LB_WIN) {
#pragma unroll
            for (int u = 0; u < LB_WIN; ++u)
                if (i0 + u < n) v[u].load(p - (i0 + u + 1) * stride);
#pragma unroll
            for (int u = 0; u < LB_WIN; ++u)
                if (i0 + u < n) {
                    v[u].wait(p - (i0 + u + 1) * stride, tag, ctl);
                    acc += v[u].value();
                    if (v[u].state()) return acc;
                }
        }
        return acc;
    }
};

// The tile's place in the call: image geometry and the index, among the
// call's published tiles, of the image's tile (0, 0).
struct IimgTileRef {
    long long tbase;
    int H, W, ip, op, tx, ty, ntx;
};

template <class T, int RPL>
struct IimgSmem {
    static constexpr int TH = 16 * RPL;
    typename T::Lds p[TH][IM_TW];  // in-tile row prefixes
    typename T::Col lp[TH];        // prefix of L over the tile's rows
    uint32_t tag;
};

// Reads the call's tag (thread 0; broadcast through LDS).
template <class S>
__device__ inline void iimg_begin(S &sm, IimgCtl *ctl) {
    if (threadIdx.x == 0)
        sm.tag = (__hip_atomic_load(&ctl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u) & 0x7fffffffu;
    __syncthreads();
}

// Counts the workgroup done; the call's last workgroup resets the counters
// and advances the epoch for the next call.  Workgroups count in groups of
// at least IM_GROUP consecutive ids, at most IM_MAX_GROUPS groups.
__device__ inline void iimg_finish(IimgCtl *ctl, uint32_t tag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const int G = (int)gridDim.x, gs = max(IM_GROUP, ceil_div(G, IM_MAX_GROUPS));
        const int gi = (int)blockIdx.x / gs, ng = ceil_div(G, gs);
        const uint32_t gsize = (uint32_t)min(gs, G - gi * gs);
        uint32_t *gc = &ctl->grp[gi][0];
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(&ctl->top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (uint32_t)ng - 1) {
                __hip_atomic_store(&ctl->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&ctl->epoch, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// One tile of the single pass (the whole workgroup).  ws: the call's
// published words (after the control block).
template <class T, int RPL>
__device__ inline void iimg_tile(IimgSmem<T, RPL> &sm, const IimgTileRef &d, const typename T::In *__restrict__ in,
                                 bool rows16, typename T::Out *__restrict__ out, IimgCtl *ctl,
                                 uint64_t *__restrict__ ws) {
    using Row = typename T::Row;
    using Col = typename T::Col;
    constexpr int TH = 16 * RPL, PW = pub_words<Col>();
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, s = lane & 15;
    const int grp = wv * 4 + g;  // row group: rows grp * RPL .. + RPL - 1
    const int y0 = d.ty * TH, x0 = d.tx * IM_TW;
    const uint32_t tag = sm.tag;
    const long long tw = iimg_tile_words<Col>(TH);
    const long long me = d.tbase + (long long)d.ty * d.ntx + d.tx;
    uint64_t *rec = ws + me * tw;

    // 1. load RPL rows of 16 elements; in-tile row prefixes into LDS; row totals published
    {
        Row v[RPL][16];
#pragma unroll
        for (int i = 0; i < RPL; ++i)
            iimg_load16<T>(v[i], in, d.H, d.W, d.ip, rows16, y0 + grp * RPL + i, x0 + 16 * s);
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = grp * RPL + i;
#pragma unroll
            for (int j = 1; j < 16; ++j) v[i][j] += v[i][j - 1];
            const Row incl = row16_inclusive_scan(v[i][15]);
            const Row excl =
                __builtin_bit_cast(Row, dpp_u32<DPP_ROW_SHR1, 0xf, true>(__builtin_bit_cast(uint32_t, incl)));
            using Lds = typename T::Lds;
            constexpr int PER16 = 16 / (int)sizeof(Lds);  // LDS values per 16-byte store
#pragma unroll
            for (int j = 0; j < 16; j += PER16) {
                Lds q[PER16];
#pragma unroll
                for (int k = 0; k < PER16; ++k) q[k] = (Lds)(excl + v[i][j + k]);
                *reinterpret_cast<uint4 *>(&sm.p[r][16 * s + j]) = __builtin_bit_cast(uint4, q);
            }
            if (s == 15) pub_put(rec + (long long)(IM_TW + r) * PW, tag, 0u, (Col)incl);
        }
    }
    // 2. L(y) from the tiles to the left, and its prefix LP over the tile's rows (one wave);
    // the first look-back window up the tile column is loaded meanwhile
    uint64_t *cw = rec + (long long)t * PW;
    const long long up = (long long)d.ntx * tw;
    Lookback<Col> lb;
    lb.prefetch(cw, up, d.ty);
    if (wv == 0) {
        Col l = 0;
        if (lane < TH)
            l = pub_sum<Col>(ws + (me - d.tx) * tw + (long long)(IM_TW + lane) * PW, tw, d.tx, tag, ctl);
        l = wave_inclusive_scan(l);
        if (lane < TH) sm.lp[lane] = l;
    }
    __syncthreads();
    // 3. column t: aggregate A = sum_y (L(y) + p(y, t)) = LP(TH - 1) + sum_y p(y, t); look-back up the tile column
    Col a = sm.lp[TH - 1];
#pragma unroll 8
    for (int y = 0; y < TH; ++y) a += (Col)sm.p[y][t];
    Col u = 0;
    if (d.ty == 0) {
        pub_put(cw, tag, 1u, a);
    } else {
        pub_put(cw, tag, 0u, a);
        u = lb.resolve(cw, up, d.ty, tag, ctl);
        pub_put(cw, tag, 1u, u + a);
    }
    // 4. out(y, t) = U + LP(y) + sum_{y' <= y} p(y', t)
    const int xc = x0 + t;
    if (xc < d.W) {
        const int nr = min(TH, d.H - y0);
        Col acc = u;
        typename T::Out *o = out + (size_t)y0 * d.op + xc;
        for (int y = 0; y < nr; ++y) {
            acc += (Col)sm.p[y][t];
            o[(size_t)y * d.op] = (typename T::Out)(acc + sm.lp[y]);
        }
    }
    __syncthreads();  // LDS reuse by the workgroup's next tile
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
