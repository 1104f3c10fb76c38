// Wave64 cross-lane scans built on DPP (gfx9 row_shr / row_bcast controls).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace umi {

// DPP control words (GFX9 encoding).
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ inline uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, BOUND_ZERO);
}

// Inclusive prefix sum across the 64 lanes of a wave (lane order), with
// operator `add` on values moved through DPP as 32-bit words.
template <typename T, typename Add>
__device__ inline T wave_inclusive_scan(T v, Add add) {
    static_assert(sizeof(T) == 4, "32-bit values");
    auto mv = [](T x, auto step) -> T {
        uint32_t b = __builtin_bit_cast(uint32_t, x);
        return __builtin_bit_cast(T, step(b));
    };
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR1, 0xf, true>(b); }));
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR2, 0xf, true>(b); }));
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR4, 0xf, true>(b); }));
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_SHR8, 0xf, true>(b); }));
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_BCAST15, 0xa, false>(b); }));
    v = add(v, mv(v, [](uint32_t b) { return dpp_u32<DPP_ROW_BCAST31, 0xc, false>(b); }));
    return v;
}

}  // namespace umi
