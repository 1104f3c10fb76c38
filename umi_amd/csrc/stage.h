// Staging clamp-to-edge byte rows into LDS as dwords, for gfx950.
//
// A frame row is read four bytes at a time; columns outside [0, W) repeat
// the edge pixel.  stage_clamped keeps U global loads in flight per thread
// before its first LDS store: every thread issues its loads unconditionally
// (indices past the end repeat the last dword, stored twice with the same
// value), and only dwords that are not 4-byte aligned or that cross the
// frame edge take the byte path.
#pragma once

#include "common.h"

namespace umi {

// Four clamped bytes of `row` from column gx, as one dword.
__device__ inline uint32_t load4_clamped(const uint8_t *row, int gx, int W) {
    if (gx >= 0 && gx + 3 < W && ((reinterpret_cast<uintptr_t>(row + gx) & 3) == 0))
        return *reinterpret_cast<const uint32_t *>(row + gx);
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) v |= (uint32_t)row[clampi(gx + b, 0, W - 1)] << (8 * b);
    return v;
}

// dst[i] = load4_clamped(row_i, 4 c_i, W) for i < n, where src(i, row, gx)
// yields the row pointer and first column of dword i.  Rows hold at least 8
// bytes (W >= 8).  All threads of the workgroup (tid < nthr) call it; the
// caller synchronises before reading dst.
template <int U, class Src>
__device__ inline void stage_clamped(uint32_t *dst, int n, int tid, int nthr, int W, Src src) {
    for (int base = 0; base < n; base += U * nthr) {
        uint32_t v[U];
        int idx[U], gx[U];
        const uint8_t *row[U];
        bool fast[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            idx[u] = min(base + u * nthr + tid, n - 1);
            src(idx[u], row[u], gx[u]);
            fast[u] = gx[u] >= 0 && gx[u] + 3 < W && ((reinterpret_cast<uintptr_t>(row[u] + gx[u]) & 3) == 0);
            // a dword-aligned address inside the row stands in for the others
            const uint8_t *safe = row[u] + ((0u - (uint32_t)reinterpret_cast<uintptr_t>(row[u])) & 3u);
            v[u] = *reinterpret_cast<const uint32_t *>(fast[u] ? row[u] + gx[u] : safe);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (!fast[u]) v[u] = load4_clamped(row[u], gx[u], W);
#pragma unroll
        for (int u = 0; u < U; ++u) dst[idx[u]] = v[u];
    }
}

// Sixteen clamped bytes of `row` from column gx, as four dwords.
__device__ inline uint4 load16_clamped(const uint8_t *row, int gx, int W) {
    return make_uint4(load4_clamped(row, gx, W), load4_clamped(row, gx + 4, W), load4_clamped(row, gx + 8, W),
                      load4_clamped(row, gx + 12, W));
}

// dst holds `rows` rows of s16 16-byte chunks at a stride of d16 >= s16
// chunks; chunk (r, c) = the clamped bytes rowp(r)[xs + 16 c .. + 16).
// rowp(r) must be 16-byte aligned with W >= 16 bytes, xs a multiple of 16
// and dst 16-byte aligned.  Every chunk takes one 16-byte load: a chunk
// inside [0, W) loads itself; a chunk left of the row loads the row's first
// chunk and repeats its byte 0; a chunk right of a row whose width is a
// multiple of 16 loads the row's last chunk and repeats its byte 15.  Only a
// chunk across W, or right of a row of another width, is assembled by
// load16_clamped after the loads.  As in stage_clamped, all U loads are
// issued before the first store: indices past the end repeat the last chunk
// (same bytes, same slot).
template <int U, class RowPtr>
__device__ inline void stage_rows16(uint32_t *dst, int rows, int s16, int d16, int xs, int W, int tid, int nthr,
                                    RowPtr rowp) {
    const int n = rows * s16;
    const bool w16 = (W & 15) == 0;
    for (int base = 0; base < n; base += U * nthr) {
        uint4 v[U];
        int idx[U], slot[U], gx[U], kind[U];  // kind: 0 inside, 1 left, 2 right, 3 assembled
        const uint8_t *row[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            idx[u] = min(base + u * nthr + tid, n - 1);
            const int r = idx[u] / s16, c = idx[u] - r * s16;
            row[u] = rowp(r);
            gx[u] = xs + 16 * c;
            slot[u] = r * d16 + c;
            kind[u] = gx[u] >= 0 && gx[u] + 16 <= W ? 0 : gx[u] < 0 ? 1 : gx[u] >= W && w16 ? 2 : 3;
            const int at = kind[u] == 0 ? gx[u] : kind[u] == 2 ? W - 16 : 0;
            v[u] = *reinterpret_cast<const uint4 *>(row[u] + at);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = (kind[u] == 1 ? v[u].x & 0xffu : v[u].w >> 24) * 0x01010101u;
            if (kind[u] == 1 || kind[u] == 2) v[u] = make_uint4(e, e, e, e);
            if (kind[u] == 3) v[u] = load16_clamped(row[u], gx[u], W);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) reinterpret_cast<uint4 *>(dst)[slot[u]] = v[u];
    }
}

}  // namespace umi
