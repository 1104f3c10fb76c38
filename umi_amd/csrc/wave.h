// Wave64 cross-lane scans and reductions built on DPP (gfx9 controls).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace umi {

// DPP control words (GFX9 encoding).
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;
constexpr int DPP_QUAD_XOR1 = 0xb1, DPP_QUAD_XOR2 = 0x4e;  // quad_perm [1,0,3,2], [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;                 // lane i <- lane 7 - i within 8
constexpr int DPP_ROW_ROR8 = 0x128;                        // lane i <- lane (i + 8) mod 16 within a row

template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ inline uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, BOUND_ZERO);
}

// Minimum over each aligned group of 8 lanes, all 8 active: the quad_perm
// swaps give every lane its quad's minimum, row_half_mirror pairs each lane
// with one of the other quad.
__device__ inline uint32_t min8_u32(uint32_t k) {
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp((int)k, (int)k, DPP_QUAD_XOR1, 0xf, 0xf, false));
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp((int)k, (int)k, DPP_QUAD_XOR2, 0xf, 0xf, false));
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp((int)k, (int)k, DPP_ROW_HALF_MIRROR, 0xf, 0xf, false));
    return k;
}

// Inclusive prefix sum within each row of 16 lanes (lane order), 32-bit
// values; row_shr with bound_ctrl feeds zero bits (0 or 0.0f) into the first
// lanes of each row.
template <typename T>
__device__ inline T row16_inclusive_scan(T v) {
    static_assert(sizeof(T) == 4, "32-bit values");
    auto sh = [](T x, auto step) -> T { return __builtin_bit_cast(T, step(__builtin_bit_cast(uint32_t, x))); };
    v += sh(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR1, 0xf, true>(b); });
    v += sh(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR2, 0xf, true>(b); });
    v += sh(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR4, 0xf, true>(b); });
    v += sh(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR8, 0xf, true>(b); });
    return v;
}

// A 4- or 8-byte value moved by one DPP control (8-byte values as two
// dwords, both moved the same way); lanes the control leaves unwritten, or
// that read past their row with bound_ctrl, get zero bits.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO, typename T>
__device__ inline T dpp_move(T x) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_u32<CTRL, ROW_MASK, BOUND_ZERO>(__builtin_bit_cast(uint32_t, x)));
    } else {
        const uint2 w = __builtin_bit_cast(uint2, x);
        return __builtin_bit_cast(T, make_uint2(dpp_u32<CTRL, ROW_MASK, BOUND_ZERO>(w.x),
                                                dpp_u32<CTRL, ROW_MASK, BOUND_ZERO>(w.y)));
    }
}

// Inclusive prefix sum across the 64 lanes of a wave (lane order; all lanes
// active): row_shr 1, 2, 4, 8 scan each row of 16, then row_bcast:15 adds
// row 0's total to row 1 and row 2's to row 3, and row_bcast:31 adds lane
// 31's to rows 2 and 3.  Zero bits are 0 for integers and +0.0 for floats.
template <typename T>
__device__ inline T wave_inclusive_scan(T v) {
    v += dpp_move<DPP_ROW_SHR1, 0xf, true>(v);
    v += dpp_move<DPP_ROW_SHR2, 0xf, true>(v);
    v += dpp_move<DPP_ROW_SHR4, 0xf, true>(v);
    v += dpp_move<DPP_ROW_SHR8, 0xf, true>(v);
    v += dpp_move<DPP_ROW_BCAST15, 0xa, false>(v);
    v += dpp_move<DPP_ROW_BCAST31, 0xc, false>(v);
    return v;
}

}  // namespace umi
