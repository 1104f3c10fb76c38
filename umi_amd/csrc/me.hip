// Full-search block motion estimation on uint8 frames (gfx950).
//
// Computes, for every BxB block of `cur`, the SAD against `ref` at every
// integer offset (dy, dx) in [-R, R]^2 with clamp-to-edge reads of `ref`, and
// the first minimum in raster order over (dy, dx).  This is the MotionEst
// strategy of main/full_me.cu:8-15,34-64 (johnjohnlin/UMI) with its clamp
// boundary (src/UMI.h:205-218, "WRAP" clamps) and output layout
// (main/full_me.cu:56-62), generalised to two frames and uint8 pixels.
//
// Design (one workgroup = a TBX x TBY tile of blocks):
//  * the tile's reference window ((TBY*B+2R) rows x (TBX*B+2R+4) bytes,
//    clamped) and the tile's current pixels are staged once in LDS;
//  * a thread owns one block and one group of 4 horizontal offsets
//    dx = 4g..4g+3, and walks the 2R+B window rows once.  Each window row is
//    read from LDS a single time and feeds the B vertical offsets that use it
//    (a ring of B accumulators indexed (row - py) mod B, static after
//    unrolling), so LDS traffic is (B/4+1) dwords per row per thread;
//  * v_qsad_pk_u16_u8 computes the SAD of 4 cur bytes against 4 byte-shifted
//    windows of 8 ref bytes in one instruction (16 |a-b| per lane-op), with
//    16-bit accumulators (B <= 16 keeps 255*B*B < 65536);
//  * the argmin key is (SAD << idx_bits) | raster_index, so a plain unsigned
//    min implements "smallest SAD, then first in raster order"; per-thread
//    minima are merged with an LDS atomic min per block.
#include "common.h"
#include "stage.h"

#include <type_traits>

namespace {
using namespace umi;

typedef __attribute__((ext_vector_type(2))) unsigned short u16x2;

// Largest workgroup per block size: B = 16 holds 64 cur dwords and 16 packed
// accumulators per thread, which needs the 256-VGPR budget of <= 512 threads.
template <int B>
constexpr int me_max_threads() { return B == 8 ? 1024 : 512; }

// Walks the 2R+B window rows of one block for the horizontal offset indices
// 4g .. 4g+NQ-1 and returns the thread's first minimum as the key
// (SAD << idx_bits) | (dyi * S + dxi).  NQ = 4: v_qsad_pk_u16_u8 (or four
// v_sad_u8 over v_alignbyte_b32-shifted words) per cur dword.  NQ = 1: the
// single column 4g, whose window dwords are dword-aligned, by v_sad_u8.
// `wbase` points at the block's first window row, dword g; rows are ws4
// dwords apart.  LFULL says that the 2R + B rows are whole chunks of B rows
// (2R a multiple of B), which makes every skip in the last chunk static.
// (The stride stays a runtime value: with a compile-time stride the compiler
// pairs dwords of consecutive rows in one ds_read2_b32, whose two halves then
// hit overlapping LDS banks.)
template <int B, bool QSAD, bool VOLUME, int NQ, bool LFULL>
__device__ inline uint32_t me_walk(const uint32_t (&c)[B][B / 4], const uint32_t *wbase, int ws4, int R, int g,
                                   int idx_bits, bool valid, size_t blk, int32_t *__restrict__ volume) {
    constexpr int B4 = B / 4;
    constexpr bool Q = QSAD && NQ == 4;
    constexpr int NW = NQ == 4 ? B4 + 1 : B4;  // window dwords read per row
    const int S = 2 * R + 1;
    const int nrows = 2 * R + B;
    // Q: one packed u64 (4 x u16) per ring slot; otherwise NQ x u32.
    uint64_t acc[B];
    uint32_t acc32[B][NQ];
#pragma unroll
    for (int s = 0; s < B; ++s) {
        acc[s] = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc32[s][q] = 0;
    }
    uint32_t bestk = 0xffffffffu;
    // Closings without the volume.  PACKED4 (8x8, four columns): with every
    // column inside the window (LFULL) a closed row's four SADs become their
    // keys (SAD << 5 | j << 2 | q) directly; otherwise they are reduced with
    // packed 16-bit mins: SAD <= 64 x 255 < 2^14, so (SAD << 2 | q) fits 16
    // bits and orders (SAD, column) as the first minimum, and columns past
    // the window are forced to 0xffff.  The thread keeps
    // (SAD << 10 | dyi << 2 | q).  PACKED1 (one column): the thread keeps
    // (SAD << 8 | dyi); dyi <= 2R <= 128.  Inside a chunk of B rows the keys
    // carry the row's index j in the chunk (a constant after unrolling) in
    // place of dyi, and the chunk's base r0 - (B-1) is added once per chunk.
    // Both form the raster key once at the end.
    constexpr bool PACKED4 = B == 8 && Q && !VOLUME;
    // 8x8 blocks without the volume: a row's window dwords are loaded while
    // the previous row's SADs run (row r + 1 before row r's qsads; past the
    // last row the last row is read again and not used).  Otherwise each row
    // loads its own.
    constexpr bool PREFETCH = B == 8 && !VOLUME;
    constexpr bool PACKED1 = NQ == 1 && !VOLUME;
    const uint32_t inv_lo = (4 * g + 0 >= S ? 0x0000ffffu : 0u) | (4 * g + 1 >= S ? 0xffff0000u : 0u);
    const uint32_t inv_hi = (4 * g + 2 >= S ? 0x0000ffffu : 0u) | (4 * g + 3 >= S ? 0xffff0000u : 0u);
    uint32_t bestp = 0xffffffffu, bestc = 0xffffffffu;

    // One window row r (= r0 + j): add its contribution to the open vertical
    // offsets dyi = r - py, then close dyi = r - (B-1).  FIRST (r0 == 0)
    // skips py > j (dyi < 0); LAST skips dyi > 2R and rows past the window.
    // Skipped terms only feed discarded slots.
    auto row = [&](const int r0, const int j, const uint32_t (&wpre)[NW], auto first_tag, auto last_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;
        constexpr bool LAST = decltype(last_tag)::value;
        const int r = r0 + j;
        if (LAST && !LFULL && r >= nrows) return;
        uint32_t w[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = PREFETCH ? wpre[k] : wbase[r * ws4 + k];
        if constexpr (!Q) {
            uint32_t u[NQ][B4];
#pragma unroll
            for (int k = 0; k < B4; ++k) {
                u[0][k] = w[k];
#pragma unroll
                for (int q = 1; q < NQ; ++q) u[q][k] = __builtin_amdgcn_alignbyte(w[k + 1], w[k], q);
            }
#pragma unroll
            for (int py = 0; py < B; ++py) {
                if (FIRST && py > j) continue;
                if (LAST && (LFULL ? py < j : r - py > 2 * R)) continue;
                const int slot = (j - py + B) % B;
#pragma unroll
                for (int q = 0; q < NQ; ++q)
#pragma unroll
                    for (int k = 0; k < B4; ++k)
                        acc32[slot][q] = __builtin_amdgcn_sad_u8(u[q][k], c[py][k], acc32[slot][q]);
            }
        } else {
#pragma unroll
            for (int py = 0; py < B; ++py) {
                if (FIRST && py > j) continue;
                if (LAST && (LFULL ? py < j : r - py > 2 * R)) continue;
                const int slot = (j - py + B) % B;
#pragma unroll
                for (int k = 0; k < B4; ++k) {
                    const uint64_t s0 = (uint64_t)w[k] | ((uint64_t)w[k + 1] << 32);
                    acc[slot] = __builtin_amdgcn_qsad_pk_u16_u8(s0, c[py][k], acc[slot]);
                }
            }
        }
        const int fs = (j + 1) % B;
        const int dyi = r - (B - 1);
        const bool closes = (!FIRST || j == B - 1) && dyi >= 0;
        if (PACKED4 && LFULL && closes) {
            // every column is inside the window (S = 4G + 1 with the light
            // lane): key_q = SAD_q * 32 + (j << 2 | q), one v_mad_u32_u16 per
            // column, op_sel picking the high half for the odd columns
            const uint2 a2 = __builtin_bit_cast(uint2, acc[fs]);
            const int jq = 4 * j;
            uint32_t k0, k1, k2, k3;
            asm("v_mad_u32_u16 %0, %1, 32, %2" : "=v"(k0) : "v"(a2.x), "s"(jq));
            asm("v_mad_u32_u16 %0, %1, 32, %2 op_sel:[1,0,0,0]" : "=v"(k1) : "v"(a2.x), "s"(jq + 1));
            asm("v_mad_u32_u16 %0, %1, 32, %2" : "=v"(k2) : "v"(a2.y), "s"(jq + 2));
            asm("v_mad_u32_u16 %0, %1, 32, %2 op_sel:[1,0,0,0]" : "=v"(k3) : "v"(a2.y), "s"(jq + 3));
            bestc = min(min(bestc, k0), k1);
            bestc = min(min(bestc, k2), k3);
        } else if (PACKED4 && closes) {
            const uint2 a2 = __builtin_bit_cast(uint2, acc[fs]);
            uint32_t lo = a2.x, hi = a2.y;
            // keeps the shift below a 32-bit v_lshl_or_b32 (not a 64-bit funnel shift plus mask)
            asm("" : "+v"(hi));
            // both u16 halves are < 2^14, so one 32-bit shift moves them without carry
            const uint32_t klo = (lo << 2) | (0x00010000u | inv_lo);  // q = 0, 1
            const uint32_t khi = (hi << 2) | (0x00030002u | inv_hi);  // q = 2, 3
            const uint32_t m = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, klo),
                                                                                       __builtin_bit_cast(u16x2, khi)));
            const uint32_t k16 = min(m & 0xffffu, m >> 16);
            // (SAD << 5 | j << 2 | q): SAD, then the row, then the column
            bestc = min(bestc, ((k16 << 3) & ~0x1fu) | ((uint32_t)j << 2) | (k16 & 3u));
        } else if (PACKED1 && closes) {
            bestc = min(bestc, (acc32[fs][0] << 8) | (uint32_t)j);
        } else if (closes) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int dxi = 4 * g + q;
                if (dxi < S) {
                    const uint32_t sad = Q ? (uint32_t)(acc[fs] >> (16 * q)) & 0xffffu : acc32[fs][q];
                    bestk = min(bestk, (sad << idx_bits) | (uint32_t)(dyi * S + dxi));
                    if (VOLUME && valid) volume[(blk * S + dyi) * S + dxi] = (int32_t)sad;
                }
            }
        }
        acc[fs] = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc32[fs][q] = 0;
    };
    using T = std::true_type;
    using F = std::false_type;
    // Every chunk closes at least one offset (the first at j = B-1, the last
    // at its first row), so bestc is a real key when it is folded.
    uint32_t wcur[NW];
    auto load_row = [&](int r, uint32_t (&w)[NW]) {
        const uint32_t *wr = wbase + min(r, nrows - 1) * ws4;
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = wr[k];
    };
    if (PREFETCH) load_row(0, wcur);
    auto chunk = [&](const int r0, auto first_tag, auto last_tag) {
        bestc = 0xffffffffu;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            uint32_t wnext[NW];
            if (PREFETCH) load_row(r0 + j + 1, wnext);
            row(r0, j, wcur, first_tag, last_tag);
            if (PREFETCH) {
#pragma unroll
                for (int k = 0; k < NW; ++k) wcur[k] = wnext[k];
            }
        }
        // chunk key (SAD << 5 | j << 2 | q) -> (SAD << 10 | dyi << 2 | q), dyi = r0 - (B-1) + j
        if constexpr (PACKED4)
            bestp = min(bestp, ((bestc & ~0x1fu) << 5) + (bestc & 0x1fu) + ((uint32_t)(r0 - (B - 1)) << 2));
        if constexpr (PACKED1) bestp = min(bestp, bestc + (uint32_t)(r0 - (B - 1)));
    };
    const int nchunks = (nrows + B - 1) / B;  // >= 2 since R >= 1
    chunk(0, T{}, F{});
    for (int ch = 1; ch < nchunks - 1; ++ch) chunk(ch * B, F{}, F{});
    chunk((nchunks - 1) * B, F{}, T{});
    if constexpr (PACKED4) {
        const uint32_t sad = bestp >> 10, dyi = (bestp >> 2) & 0xffu, q = bestp & 3u;
        bestk = (sad << idx_bits) | (dyi * (uint32_t)S + 4u * g + q);
    } else if constexpr (PACKED1) {
        const uint32_t sad = bestp >> 8, dyi = bestp & 0xffu;
        bestk = (sad << idx_bits) | (dyi * (uint32_t)S + 4u * g);
    }
    return bestk;
}

// One TBX x TBY tile of blocks whose first block is (bx0, by0); all threads
// of the workgroup call it (it contains barriers).  Threads [0, nblk*G) own
// (block, 4-column group g).  With GL = 1 (S = 4G + 1) the next nblk threads
// own (block, the last column 4G): one light lane per block instead of a
// fifth group with three dead columns, so that a 64-block tile is 8 waves of
// four-column work plus one light wave.
template <int B, bool QSAD, bool VOLUME, bool LFULL>
__device__ inline void me_tile(uint32_t *lds, const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref, int H,
                               int W, int pitch, int R, int nby, int nbx, int tbx, int tby, int G, int GL, int ws4,
                               int wrows, int fast16, int idx_bits, int32_t *__restrict__ mv_sad,
                               int32_t *__restrict__ volume, int bx0, int by0) {
    constexpr int B4 = B / 4;
    const int nblk = tbx * tby;
    const int cs4 = tbx * B4;             // cur tile row stride (dwords)
    uint32_t *win = lds;                  // wrows x ws4 dwords
    uint32_t *ctile = win + wrows * ws4;  // (tby*B) x cs4 dwords
    uint32_t *best = ctile + tby * B * cs4;

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int x0 = bx0 * B, y0 = by0 * B;

    // Reference window, clamp-to-edge: window (r, c) <- ref[cl(y0-R+r)][cl(xs+c)]
    // and current tile (blocks past the frame edge read clamped pixels and are
    // never written out).  fast16 (16-byte aligned frames and rows, R % 4 == 0,
    // tile width a multiple of 16): the window starts at xs = floor16(x0 - R)
    // and both are staged in 16-byte chunks; otherwise xs = x0 - R, by dwords.
    const int xr = x0 - R, xs = fast16 ? (xr & ~15) : xr;
    const auto ref_row = [&](int r) { return ref + (size_t)clampi(y0 - R + r, 0, H - 1) * pitch; };
    const auto cur_row = [&](int r) { return cur + (size_t)min(y0 + r, H - 1) * pitch; };
    if (fast16) {
        const int need16 = ((xr - xs) + tbx * B + 2 * R + 4 + 15) >> 4;  // <= ws4 / 4
        stage_rows16<2>(win, wrows, need16, ws4 / 4, xs, W, tid, nthr, ref_row);
        stage_rows16<1>(ctile, tby * B, cs4 / 4, cs4 / 4, x0, W, tid, nthr, cur_row);
    } else {
        stage_clamped<8>(win, wrows * ws4, tid, nthr, W, [&](int i, const uint8_t *&row, int &gx) {
            const int r = i / ws4;
            row = ref_row(r);
            gx = xs + 4 * (i - r * ws4);
        });
        stage_clamped<4>(ctile, tby * B * cs4, tid, nthr, W, [&](int i, const uint8_t *&row, int &gx) {
            const int r = i / cs4;
            row = cur_row(r);
            gx = x0 + 4 * (i - r * cs4);
        });
    }
    for (int i = tid; i < nblk; i += nthr) best[i] = 0xffffffffu;
    __syncthreads();

    const int nmain = nblk * G;
    const bool light = tid >= nmain;
    const int bl = light ? tid - nmain : tid / G;
    const int g = light ? G : tid - bl * G;
    if (bl < nblk && (!light || GL)) {
        const int bxl = bl % tbx, byl = bl / tbx;
        const int bx = bx0 + bxl, by = by0 + byl;
        const bool valid = bx < nbx && by < nby;
        const size_t blk = (size_t)by * nbx + bx;

        uint32_t c[B][B4];
#pragma unroll
        for (int py = 0; py < B; ++py)
#pragma unroll
            for (int k = 0; k < B4; ++k) c[py][k] = ctile[(byl * B + py) * cs4 + bxl * B4 + k];

        const uint32_t *wbase = win + (byl * B) * ws4 + ((xr - xs) >> 2) + bxl * B4 + g;
        const uint32_t bestk =
            light ? me_walk<B, QSAD, VOLUME, 1, LFULL>(c, wbase, ws4, R, g, idx_bits, valid, blk, volume)
                  : me_walk<B, QSAD, VOLUME, 4, LFULL>(c, wbase, ws4, R, g, idx_bits, valid, blk, volume);
        atomicMin(&best[bl], bestk);
    }
    __syncthreads();
    if (mv_sad) {
        const int S = 2 * R + 1;
        for (int i = tid; i < nblk; i += nthr) {
            const int bx = bx0 + i % tbx, by = by0 + i / tbx;
            if (bx < nbx && by < nby) {
                const uint32_t key = best[i];
                const uint32_t idx = key & ((1u << idx_bits) - 1u);
                int32_t *o = mv_sad + ((size_t)by * nbx + bx) * 3;
                o[0] = (int32_t)(idx / S) - R;
                o[1] = (int32_t)(idx % S) - R;
                o[2] = (int32_t)(key >> idx_bits);
            }
        }
    }
}

template <int B, bool QSAD, bool VOLUME, bool LFULL>
__global__ __launch_bounds__(me_max_threads<B>()) void me_full_search_kernel(
    const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref, int H, int W, int pitch, int R, int nby, int nbx,
    int tbx, int tby, int G, int GL, int ws4, int wrows, int fast16, int idx_bits, int32_t *__restrict__ mv_sad,
    int32_t *__restrict__ volume) {
    extern __shared__ __align__(16) uint32_t lds[];
    me_tile<B, QSAD, VOLUME, LFULL>(lds, cur, ref, H, W, pitch, R, nby, nbx, tbx, tby, G, GL, ws4, wrows, fast16, idx_bits,
                             mv_sad, volume, blockIdx.x * tbx, blockIdx.y * tby);
}

// Listed blocks: a persistent grid walks list[0 .. *count) (block indices
// by * nbx + bx), one block per iteration.  Thread t searches the 4 horizontal
// offsets 4 (t % G) .. + 3 over the vertical offsets of chunk t / G (DCH
// offsets each), from the block's staged window; the first minimum in raster
// order wins, as in the full search.  A workgroup reads its next entry
// before searching the current one.
// fast16 (16-byte aligned frames and pitch, W >= 16, R % 4 == 0): the window
// rows start at xs = floor16(x0 - R) and are staged in 16-byte chunks (ws4 a
// multiple of 4); the search then reads them (x0 - R - xs) / 4 dwords in.
template <int B>
__global__ __launch_bounds__(512) void me_listed_kernel(const uint8_t *__restrict__ cur,
                                                        const uint8_t *__restrict__ ref, int H, int W, int pitch,
                                                        int R, int nbx, int G, int DCH, int ws4, int wrows,
                                                        int idx_bits, int fast16, const int32_t *__restrict__ list,
                                                        const int32_t *__restrict__ count,
                                                        int32_t *__restrict__ mv_sad) {
    constexpr int B4 = B / 4;
    extern __shared__ uint32_t lds[];
    uint32_t *win = lds, *ctile = win + wrows * ws4, *best = ctile + B * B4;
    const int S = 2 * R + 1;
    const int tid = threadIdx.x, g = tid % G, ch = tid / G;
    const int n = *count;
    int next = (int)blockIdx.x < n ? list[blockIdx.x] : 0;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int blk = next;
        if (i + (int)gridDim.x < n) next = list[i + gridDim.x];
        const int bx = blk % nbx, by = blk / nbx, x0 = bx * B, y0 = by * B;
        const auto ref_row = [&](int r) { return ref + (size_t)clampi(y0 - R + r, 0, H - 1) * pitch; };
        int woff = 0;
        if (fast16) {
            const int xs = (x0 - R) & ~15;
            woff = (x0 - R - xs) >> 2;
            stage_rows16<2>(win, wrows, ws4 / 4, ws4 / 4, xs, W, tid, blockDim.x, ref_row);
        } else {
            stage_clamped<8>(win, wrows * ws4, tid, blockDim.x, W, [&](int k, const uint8_t *&row, int &gx) {
                const int r = k / ws4;
                row = ref_row(r);
                gx = x0 - R + 4 * (k - r * ws4);
            });
        }
        stage_clamped<1>(ctile, B * B4, tid, blockDim.x, W, [&](int k, const uint8_t *&row, int &gx) {
            row = cur + (size_t)(y0 + k / B4) * pitch;
            gx = x0 + 4 * (k % B4);
        });
        if (tid == 0) *best = 0xffffffffu;
        __syncthreads();
        const int d0 = ch * DCH, d1 = min(S, d0 + DCH);
        if (d0 < S) {
            uint32_t c[B][B4];
#pragma unroll
            for (int py = 0; py < B; ++py)
#pragma unroll
                for (int k = 0; k < B4; ++k) c[py][k] = ctile[py * B4 + k];
            uint32_t bestk = 0xffffffffu;
            for (int dyi = d0; dyi < d1; ++dyi) {
                uint64_t acc = 0;  // 4 x u16 SADs at horizontal offsets 4g .. 4g + 3
#pragma unroll
                for (int py = 0; py < B; ++py) {
                    const uint32_t *wr = win + (dyi + py) * ws4 + woff + g;
                    uint32_t w[B4 + 1];
#pragma unroll
                    for (int k = 0; k <= B4; ++k) w[k] = wr[k];
#pragma unroll
                    for (int k = 0; k < B4; ++k)
                        acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w[k] | ((uint64_t)w[k + 1] << 32), c[py][k],
                                                              acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int dxi = 4 * g + q;
                    if (dxi < S)
                        bestk = min(bestk, ((uint32_t)(acc >> (16 * q)) & 0xffffu) << idx_bits |
                                               (uint32_t)(dyi * S + dxi));
                }
            }
            atomicMin(best, bestk);
        }
        __syncthreads();
        if (tid == 0) {
            const uint32_t key = *best, idx = key & ((1u << idx_bits) - 1u);
            int32_t *o = mv_sad + (size_t)blk * 3;
            o[0] = (int32_t)(idx / S) - R;
            o[1] = (int32_t)(idx % S) - R;
            o[2] = (int32_t)(key >> idx_bits);
        }
        __syncthreads();
    }
}

struct MePlan {
    int G, GL, tbx, tby, threads, ws4, wrows, fast16, stat, idx_bits;
    size_t lds;
};

// S = 2R + 1 columns: S = 4G + 1 takes G four-column groups and one light
// lane per block (unless `groups_only`); S = 4G - 1 takes G groups, the last
// with one dead column.  `aligned16`: both frames and the pitch are 16-byte
// aligned and W >= 16, which with R % 4 == 0 allows 16-byte staging; the
// window row then starts at floor16(x0 - R) and spans whole 16-byte chunks.
//
// `stat`: 8x8 blocks, R % 4 == 0 (the 2R + 8 window rows are whole chunks of
// 8), the light lane (so no group has a column past the window) and no SAD
// volume: the kernel is compiled with a static last chunk and direct keys.
MePlan me_plan(int B, int R, bool groups_only, bool aligned16, bool volume, bool allow_static) {
    MePlan p;
    const int S = 2 * R + 1;
    p.GL = (S % 4 == 1 && !groups_only) ? 1 : 0;
    p.G = p.GL ? S / 4 : umi::ceil_div(S, 4);
    const int max_threads = B == 8 ? me_max_threads<8>() : me_max_threads<16>();
    p.idx_bits = umi::bits_for((uint32_t)(S * S - 1));
    int nblk = 1;
    while (nblk * 2 * (p.G + p.GL) <= max_threads && nblk * 2 <= 256) nblk *= 2;
    // shrink the tile until window + current tile + keys fit 64 KiB of LDS
    for (;; nblk /= 2) {
        p.tbx = nblk < 16 ? nblk : 16;
        p.tby = nblk / p.tbx;
        p.threads = umi::round_up(nblk * (p.G + p.GL), 64);
        p.fast16 = aligned16 && R % 4 == 0 && (p.tbx * B) % 16 == 0;
        // x0 is a multiple of 16, so x0 - R sits (-R) mod 16 bytes past xs
        p.ws4 = p.fast16 ? 4 * umi::ceil_div((16 - R % 16) % 16 + p.tbx * B + 2 * R + 4, 16)
                         : umi::ceil_div(p.tbx * B + 2 * R + 4, 4);
        p.stat = allow_static && B == 8 && R % 4 == 0 && !volume && p.GL == 1;
        p.wrows = p.tby * B + 2 * R;
        p.lds = 4 * ((size_t)p.wrows * p.ws4 + (size_t)p.tby * B * p.tbx * (B / 4) + nblk);
        if (p.lds <= 64 * 1024 || nblk == 1) break;
    }
    return p;
}

// bit 0: 1 = v_qsad_pk_u16_u8, 0 = v_sad_u8 + v_alignbyte_b32;
// bit 1: four-column groups only (no light lanes); bit 2: dword staging only;
// bit 3: no static last chunk
int g_me_variant = 1;

MePlan me_plan_for(int B, int R, const uint8_t *cur, const uint8_t *ref, int W, int pitch, bool volume) {
    const bool aligned16 =
        ((reinterpret_cast<uintptr_t>(cur) | reinterpret_cast<uintptr_t>(ref) | (uintptr_t)pitch) & 15) == 0 &&
        W >= 16 && (g_me_variant & 4) == 0;
    return me_plan(B, R, (g_me_variant & 2) != 0, aligned16, volume, (g_me_variant & 8) == 0);
}

template <int B, bool QSAD>
umi_status me_launch(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R,
                     int32_t *mv_sad, int32_t *volume, hipStream_t st) {
    const MePlan p = me_plan_for(B, R, cur, ref, W, pitch, volume != nullptr);
    const int nby = H / B, nbx = W / B;
    dim3 grid(umi::ceil_div(nbx, p.tbx), umi::ceil_div(nby, p.tby));
#define UMI_ME_ARGS                                                                                                  \
    grid, dim3(p.threads), p.lds, st, cur, ref, H, W, pitch, R, nby, nbx, p.tbx, p.tby, p.G, p.GL, p.ws4, p.wrows,    \
        p.fast16, p.idx_bits, mv_sad, volume
    if (volume)
        hipLaunchKernelGGL((me_full_search_kernel<B, QSAD, true, false>), UMI_ME_ARGS);
    else if constexpr (B == 8) {
        if (p.stat)
            hipLaunchKernelGGL((me_full_search_kernel<B, QSAD, false, true>), UMI_ME_ARGS);
        else
            hipLaunchKernelGGL((me_full_search_kernel<B, QSAD, false, false>), UMI_ME_ARGS);
    } else
        hipLaunchKernelGGL((me_full_search_kernel<B, QSAD, false, false>), UMI_ME_ARGS);
#undef UMI_ME_ARGS
    return umi::launch_check("me_full_search_kernel");
}

}  // namespace

extern "C" {

// Selects the full search's SAD instruction form (bit 0: 1 =
// v_qsad_pk_u16_u8, 0 = v_sad_u8 + v_alignbyte_b32), with bit 1 the layout
// without light lanes, with bit 2 dword staging of aligned frames and with
// bit 3 no static last chunk.  All give identical results; exported for
// measurement and tests.
umi_status umi_me_set_variant(int variant) {
    if (variant < 0 || variant > 15) return umi::set_error(UMI_ERR_INVALID, "umi_me_set_variant: %d", variant);
    g_me_variant = variant;
    return UMI_OK;
}

umi_status umi_me_full_search_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int B, int R,
                                 int32_t *mv_sad, int32_t *volume, umi_stream s) {
    using umi::set_error;
    if (!cur || !ref) return set_error(UMI_ERR_INVALID, "umi_me_full_search_u8: null frame");
    if (B != 8 && B != 16) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_u8: block %d (8 or 16)", B);
    if (R < 1 || R > 64) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_u8: radius %d (1..64)", R);
    if (H < 1 || W < 1 || pitch < W)
        return set_error(UMI_ERR_INVALID, "umi_me_full_search_u8: H=%d W=%d pitch=%d", H, W, pitch);
    if (H / B == 0 || W / B == 0) return UMI_OK;  // no whole block
    if (!mv_sad && !volume) return set_error(UMI_ERR_INVALID, "umi_me_full_search_u8: no output");
    const MePlan p = me_plan_for(B, R, cur, ref, W, pitch, volume != nullptr);
    if (p.lds > 64 * 1024) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_u8: LDS %zu B", p.lds);
    const int sad_bits = umi::bits_for(255u * B * B);
    if (sad_bits + p.idx_bits > 32) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_u8: key overflow");
    hipStream_t st = umi::as_stream(s);
    if (B == 8)
        return (g_me_variant & 1) ? me_launch<8, true>(cur, ref, H, W, pitch, R, mv_sad, volume, st)
                            : me_launch<8, false>(cur, ref, H, W, pitch, R, mv_sad, volume, st);
    // B = 16 always uses v_qsad_pk_u16_u8: the v_sad_u8 form needs 64 more
    // accumulator registers than fit without spilling.
    return me_launch<16, true>(cur, ref, H, W, pitch, R, mv_sad, volume, st);
}

}  // extern "C"

extern "C" umi_status umi_me_full_search_listed_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                                   int B, int R, const int32_t *list, const int32_t *count,
                                                   int max_count, int32_t *mv_sad, umi_stream s) {
    using umi::set_error;
    if (!cur || !ref || !list || !count || !mv_sad)
        return set_error(UMI_ERR_INVALID, "umi_me_full_search_listed_u8: null pointer");
    if (B != 8 && B != 16) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_listed_u8: block %d", B);
    if (R < 1 || R > 64) return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_listed_u8: radius %d", R);
    if (H < 1 || W < 1 || pitch < W) return set_error(UMI_ERR_INVALID, "umi_me_full_search_listed_u8: frame");
    if (max_count <= 0 || H / B == 0 || W / B == 0) return UMI_OK;
    const int G = umi::ceil_div(2 * R + 1, 4);
    const int S = 2 * R + 1;
    const int idx_bits = umi::bits_for((uint32_t)(S * S - 1));
    if (umi::bits_for(255u * B * B) + idx_bits > 32)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_me_full_search_listed_u8: key overflow");
    // (R % 4 == 0 keeps x0 - R - floor16(x0 - R) a whole number of dwords)
    const bool fast16 = ((reinterpret_cast<uintptr_t>(cur) | reinterpret_cast<uintptr_t>(ref) | (uintptr_t)pitch) &
                         15) == 0 &&
                        W >= 16 && R % 4 == 0;
    // fast16: up to 12 bytes before x0 - R and whole 16-byte chunks
    const int ws4 = fast16 ? 4 * umi::ceil_div(12 + B + 2 * R + 4, 16) : umi::ceil_div(B + 2 * R + 4, 4);
    const int wrows = B + 2 * R;
    const size_t lds = 4 * ((size_t)wrows * ws4 + (size_t)B * (B / 4) + 1);
    // split the vertical offsets into chunks so that G x chunks threads fit 512
    // (256 for B = 16, whose cur rows take 64 registers)
    const int max_threads = B == 8 ? 512 : 256;
    const int DCH = umi::ceil_div(S, max_threads / G);
    const int threads = umi::round_up(G * umi::ceil_div(S, DCH), 64);
    const int grid = max_count < 512 ? max_count : 512;  // persistent: each workgroup walks the list
    hipStream_t st = umi::as_stream(s);
    if (B == 8)
        hipLaunchKernelGGL((me_listed_kernel<8>), dim3(grid), dim3(threads), lds, st, cur, ref, H, W, pitch, R, W / B,
                           G, DCH, ws4, wrows, idx_bits, (int)fast16, list, count, mv_sad);
    else
        hipLaunchKernelGGL((me_listed_kernel<16>), dim3(grid), dim3(threads), lds, st, cur, ref, H, W, pitch, R,
                           W / B, G, DCH, ws4, wrows, idx_bits, (int)fast16, list, count, mv_sad);
    return umi::launch_check("me_listed_kernel");
}
