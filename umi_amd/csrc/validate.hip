// Descriptor validation for the generic UMI::Execute<> engine (host code).
//
// Rules kept from the reference planner (src/UMI.cpp:39-66, johnjohnlin/UMI):
// warp_shuf a power of two, block_size a multiple of warp_size * warp_shuf,
// product of warp_size over the four dims == 32.  Differences: warp_shuf 0
// is read as 1 (main/iimg.cu:60 passes 3-element ParallelDims, which the
// reference rejects), and the result is a status code, not abort().
// Added checks: block sizes >= 1, non-negative totals, view dims in 0..3,
// input sizes >= 1, plane counts <= 4, Expand nibbles >= 1, at most 1024
// threads per block, ZeroStrideOpt bits only on inputs whose view does not
// move along the expanded dim.
#include "common.h"

extern "C" umi_status umi_validate_execute(const int32_t *p_dims, const int32_t *a_dims, int n_in, int n_out,
                                           const int32_t *views, const int32_t *phys, const int32_t *planes,
                                           uint64_t expand, uint64_t zero_stride_opt) {
    using umi::set_error;
    if (!p_dims || !a_dims) return set_error(UMI_ERR_INVALID, "Execute: null descriptor");
    if (n_in < 0 || n_out < 0 || n_in + n_out > 64) return set_error(UMI_ERR_INVALID, "Execute: %d inputs, %d outputs", n_in, n_out);
    if ((n_in + n_out) > 0 && (!views || !phys || !planes)) return set_error(UMI_ERR_INVALID, "Execute: null array descriptors");
    int64_t warp_prod = 1;
    for (int d = 0; d < 4; ++d) {
        const int32_t total = p_dims[4 * d], block = p_dims[4 * d + 1], warp = p_dims[4 * d + 2];
        const int32_t shuf = p_dims[4 * d + 3] == 0 ? 1 : p_dims[4 * d + 3];
        if (total < 0) return set_error(UMI_ERR_INVALID, "Execute: ParallelDim[%d].total_size = %d", d, total);
        if (block < 1 || warp < 1) return set_error(UMI_ERR_INVALID, "Execute: ParallelDim[%d] block %d warp %d", d, block, warp);
        if (shuf < 0 || (shuf & (shuf - 1)) != 0)
            return set_error(UMI_ERR_INVALID, "Execute: ParallelDim[%d].warp_shuf = %d is not a power of two", d, shuf);
        if (block % (warp * shuf) != 0)
            return set_error(UMI_ERR_INVALID, "Execute: ParallelDim[%d].block_size %d is not a multiple of warp %d x shuf %d", d, block, warp, shuf);
        warp_prod *= warp;
        if (a_dims[2 * d] < 0 || a_dims[2 * d + 1] < 1)
            return set_error(UMI_ERR_INVALID, "Execute: AccumulationDim[%d] = {%d, %d}", d, a_dims[2 * d], a_dims[2 * d + 1]);
        const uint64_t e = (expand >> (4 * (3 - d))) & 0xf;
        if (e == 0) return set_error(UMI_ERR_INVALID, "Execute: Expand nibble of dim %d is 0", d);
    }
    if (warp_prod != 32) return set_error(UMI_ERR_INVALID, "Execute: product of warp sizes is %lld, not 32", (long long)warp_prod);
    int64_t threads = 1;
    for (int d = 0; d < 4; ++d) threads *= p_dims[4 * d + 1];
    if (threads > 1024) return set_error(UMI_ERR_INVALID, "Execute: %lld threads per block (at most 1024)", (long long)threads);
    // ZeroStrideOpt bit 4 i + d shares input i's operands across the Expand
    // points along dim d: only sound when the input's parallel view does not
    // move along d (the reference would silently read the wrong operands)
    for (int i = 0; i < n_in && i < 16; ++i)
        for (int d = 0; d < 4; ++d) {
            const uint64_t e = (expand >> (4 * (3 - d))) & 0xf;
            if (((zero_stride_opt >> (4 * i + d)) & 1) && e > 1 && views[24 * i + 3 * d + 2] != 0)
                return set_error(UMI_ERR_INVALID,
                                 "Execute: ZeroStrideOpt bit %d set, but input %d's view moves along dim %d (stride %d)",
                                 4 * i + d, i, d, views[24 * i + 3 * d + 2]);
        }
    for (int i = 0; i < n_in + n_out; ++i) {
        if (planes[i] < 0 || planes[i] > 4) return set_error(UMI_ERR_INVALID, "Execute: array %d has %d planes", i, planes[i]);
        for (int v = 0; v < 8; ++v) {
            const int32_t id = views[24 * i + 3 * v];
            if (id < 0 || id > 3) return set_error(UMI_ERR_INVALID, "Execute: array %d view %d maps to physical dim %d", i, v, id);
        }
        if (i < n_in && planes[i] > 0)
            for (int k = 0; k < 4; ++k)
                if (phys[8 * i + 2 * k] < 1)
                    return set_error(UMI_ERR_INVALID, "Execute: input %d physical dim %d has size %d", i, k, phys[8 * i + 2 * k]);
        for (int k = 0; k < 4; ++k) {
            const int32_t mode = phys[8 * i + 2 * k + 1];
            if (mode != 0 && mode != 1) return set_error(UMI_ERR_INVALID, "Execute: array %d dim %d boundary mode %d", i, k, mode);
        }
    }
    return UMI_OK;
}
