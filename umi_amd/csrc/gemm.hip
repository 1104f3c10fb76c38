// MFMA GEMM and implicit-GEMM convolution (gfx950).
//
//   C[m][n] = alpha * sum_k A[m][k] * Bt[n][k]          (A, Bt K-contiguous, C f32)
//
// This is the contraction of the Gemm strategy (accum += a*b) that the
// reference runs as a scalar UMI launch for GEMM (main/gemm.cu:8-15,37-67)
// and direct convolution (main/fprop.cu:10-17,53-82), johnjohnlin/UMI.  Here
// it runs on the matrix cores in one of five forms:
//   F32      v_mfma_f32_32x32x2_f32
//   BF16     v_mfma_f32_32x32x16_bf16, fp32 accumulation
//   BF16X3   the same with A = Ah + Al, B = Bh + Bl (bf16 hi/lo split):
//            Ah Bh + Al Bh + Ah Bl accumulated together (fp32-level accuracy)
//   FP8      v_mfma_scale_f32_32x32x64_f8f6f4 on OCP e4m3 operands scaled
//            per tensor (q = 448 / amax); alpha = 1 / (qa qb)
//   FP8X3    FP8 plus the residual terms Al Bh + Ah Bl, where Al = e4m3((A q -
//            Ah) * 16); the 2^-4 weight is the E8M0 operand scale of the
//            instruction, so all three products share one accumulator.
//
// Design: 256 x 256 output tile per 256-thread workgroup (2 x 2 waves, each
// 128 x 128 = 4 x 4 MFMA 32x32 blocks held in the accumulator registers;
// 128 x 128 and 64 x 128 tiles for smaller or narrow problems);
// K is walked in stages of 64 bytes per operand row (32 bf16, 64 fp8,
// 16 f32), double-buffered in LDS (row stride 80 B so ds_read_b128 of 32
// consecutive rows spreads over the banks), global loads for stage s+1
// issued before the MFMAs of stage s.  Workgroup ids are remapped so that
// the tiles resident on one XCD (ids congruent mod 8) form a compact block
// of C and share A/B panels in that XCD's L2.  The convolution form gathers
// the B panel rows from an NHWC input (implicit im2col, zero padding), the
// A panel is the KRSC filter, and C is written NCHW.
#include "common.h"

#include <algorithm>

namespace {
using namespace umi;

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) int i32x8;
typedef __attribute__((ext_vector_type(4))) int i32x4;

enum GemmMode { GM_F32 = 0, GM_BF16 = 1, GM_BF16X3 = 2, GM_FP8 = 3, GM_FP8X3 = 4 };

constexpr int ROWB = 64;          // bytes of one operand row per stage
constexpr int LDSROW = ROWB + 16; // padded LDS row stride
constexpr int NTHREADS = 256;

template <int MODE>
struct Mfma;
template <>
struct Mfma<GM_F32> {
    static constexpr int ESZ = 4, LANEB = 4, CH = 8, NOP = 2;
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, f32x16 c, int, int) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(*(const float *)a, *(const float *)b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<GM_BF16> {
    static constexpr int ESZ = 2, LANEB = 16, CH = 2, NOP = 2;
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, f32x16 c, int, int) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*(const bf16x8 *)a, *(const bf16x8 *)b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<GM_BF16X3> : Mfma<GM_BF16> {
    static constexpr int NOP = 4;
};
template <>
struct Mfma<GM_FP8> {
    static constexpr int ESZ = 1, LANEB = 32, CH = 1, NOP = 2;
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, f32x16 c, int sa, int sb) {
        const i32x4 *a4 = (const i32x4 *)a, *b4 = (const i32x4 *)b;
        const i32x4 a0 = a4[0], a1 = a4[1], b0 = b4[0], b1 = b4[1];
        const i32x8 av = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        const i32x8 bv = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, sa, 0, sb);
    }
};
template <>
struct Mfma<GM_FP8X3> : Mfma<GM_FP8> {
    static constexpr int NOP = 4;
};

// Row sources for a panel: dense K-contiguous rows, or implicit im2col rows
// of an NHWC tensor (row = output pixel, k = (r, s, c)).  prep(row) is
// evaluated once per tile, kprep(kbyte) once per stage, at() per chunk.
struct DenseSrc {
    const uint8_t *base;
    long ld;      // bytes between rows
    int rows;
    struct P {
        const uint8_t *p;
    };
    struct Kp {
        int kb;
    };
    __device__ P prep(int r, int /*esz*/) const { return {r < rows ? base + (long)r * ld : nullptr}; }
    __device__ Kp kprep(int kb, int /*esz*/) const { return {kb}; }
    __device__ const uint8_t *at(const P &p, const Kp &k) const { return p.p ? p.p + k.kb : nullptr; }
};
struct ConvSrc {
    const uint8_t *base;  // NHWC input
    int N, H, W, C;       // C in elements
    int OH, OW, R, S, stride, pad, dil;
    struct P {
        const uint8_t *img;  // start of image n, or null past the last pixel
        int iy0, ix0;        // oy * stride - pad, ox * stride - pad
    };
    struct Kp {
        int dy, dx;  // tap offset (r * dil, s * dil)
        int cb;      // channel offset in bytes
        int pixb;    // bytes per pixel (C * element size)
    };
    __device__ P prep(int j, int esz) const {
        const int pix = OH * OW;
        if (j >= N * pix) return {nullptr, 0, 0};
        const int n = j / pix, p = j - n * pix, oy = p / OW, ox = p - oy * OW;
        return {base + (long)n * H * W * C * esz, oy * stride - pad, ox * stride - pad};
    }
    __device__ Kp kprep(int kbyte, int esz) const {
        const int k = kbyte / esz;
        const int rs = k / C, c = k - rs * C;
        const int r = rs / S, s = rs - r * S;
        return {r * dil, s * dil, c * esz, C * esz};
    }
    __device__ const uint8_t *at(const P &p, const Kp &k) const {
        if (!p.img) return nullptr;
        const int iy = p.iy0 + k.dy, ix = p.ix0 + k.dx;
        if (iy < 0 || iy >= H || ix < 0 || ix >= W) return nullptr;
        return p.img + ((long)iy * W + ix) * k.pixb + k.cb;
    }
};

// Output sinks: row-major C, or NCHW convolution output (row = out channel,
// column = pixel of the batch).
struct DenseDst {
    float *base;
    long ldc;
    int M, N;
    __device__ void put(int m, int n, float v) const {
        if (m < M && n < N) base[(long)m * ldc + n] = v;
    }
};
// NHWC bf16 output with per-channel bias and optional ReLU (the layer
// output of the motion-proposal network, fed to the next layer as is).
struct ConvDst {
    float *base;  // [N][K][OH][OW]
    int K, npix, total;
    __device__ void put(int m, int j, float v) const {
        if (m < K && j < total) {
            const int n = j / npix, p = j - n * npix;
            base[((long)n * K + m) * npix + p] = v;
        }
    }
};

// 16 bytes of one panel row (zeros where the source row is absent).
__device__ inline i32x4 fetch16(const uint8_t *p) {
    return p ? *(const i32x4 *)p : i32x4{0, 0, 0, 0};
}

// Workgroup tile: WGM x WGN waves, each wave TM x TN MFMA blocks of 32 x 32.
template <int WGM_, int WGN_, int TM_, int TN_>
struct Tile {
    static constexpr int WGM = WGM_, WGN = WGN_, TM = TM_, TN = TN_;
    static constexpr int BM = WGM * TM * 32, BN = WGN * TN * 32;
    static_assert(WGM * WGN * 64 == NTHREADS, "four waves");
};
using TileGemm = Tile<2, 2, 4, 4>;   // 256 x 256: large GEMMs
using TileMid = Tile<2, 2, 2, 2>;    // 128 x 128
using TileNarrow = Tile<1, 4, 2, 1>; // 64 x 128: few output rows (conv with K <= 64)

template <class TL, int MODE, class ASrc, class BSrc, class Dst>
__global__ __launch_bounds__(NTHREADS) void gemm_kernel(ASrc a_hi, ASrc a_lo, BSrc b_hi, BSrc b_lo, Dst dst, int K,
                                                        int tiles_m, int tiles_n, const float *qa, const float *qb) {
    using OP = Mfma<MODE>;
    constexpr bool SPLIT = OP::NOP == 4;
    constexpr int BMt = TL::BM, BNt = TL::BN, TM = TL::TM, TN = TL::TN;
    constexpr int RA = BMt / 64, RB = BNt / 64;          // 16-B chunks per thread per panel
    constexpr int OFF_B = BMt * LDSROW;                   // panel offsets inside one stage
    constexpr int OFF_LO = (BMt + BNt) * LDSROW;
    constexpr int STAGE = (SPLIT ? 2 : 1) * (BMt + BNt) * LDSROW;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv / TL::WGN, wn = wv % TL::WGN;

    // XCD-aware tile order: ids on one XCD (id % 8) take consecutive tiles.
    const int ntiles = tiles_m * tiles_n;
    int pid = blockIdx.x;
    if (ntiles % 8 == 0) pid = (pid % 8) * (ntiles / 8) + pid / 8;
    // grouped raster: 4 tile rows per group keeps A panels hot
    const int GROUP = 4;
    const int group = pid / (GROUP * tiles_n), first_m = group * GROUP;
    const int gsz = min(GROUP, tiles_m - first_m);
    const int tm = first_m + (pid % (GROUP * tiles_n)) % gsz;
    const int tn = (pid % (GROUP * tiles_n)) / gsz;
    const int m0 = tm * BMt, n0 = tn * BNt;

    const int nstages = K * OP::ESZ / ROWB;
    // global -> LDS assignment: thread t copies 16 B chunk (t & 3) of rows t/4 + 64 i
    const int lrow = tid >> 2, lchunk = (tid & 3) * 16;
    i32x4 ra[SPLIT ? 2 : 1][RA], rb[SPLIT ? 2 : 1][RB];
    typename ASrc::P pa_[RA], pal_[RA];
    typename BSrc::P pb_[RB], pbl_[RB];
#pragma unroll
    for (int i = 0; i < RA; ++i) {
        pa_[i] = a_hi.prep(m0 + lrow + 64 * i, OP::ESZ);
        if constexpr (SPLIT) pal_[i] = a_lo.prep(m0 + lrow + 64 * i, OP::ESZ);
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        pb_[i] = b_hi.prep(n0 + lrow + 64 * i, OP::ESZ);
        if constexpr (SPLIT) pbl_[i] = b_lo.prep(n0 + lrow + 64 * i, OP::ESZ);
    }
    auto gload = [&](int st) {
        const int kb = st * ROWB + lchunk;
        const auto ka = a_hi.kprep(kb, OP::ESZ);
        const auto kbp = b_hi.kprep(kb, OP::ESZ);
#pragma unroll
        for (int i = 0; i < RA; ++i) {
            ra[0][i] = fetch16(a_hi.at(pa_[i], ka));
            if constexpr (SPLIT) ra[1][i] = fetch16(a_lo.at(pal_[i], ka));
        }
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            rb[0][i] = fetch16(b_hi.at(pb_[i], kbp));
            if constexpr (SPLIT) rb[1][i] = fetch16(b_lo.at(pbl_[i], kbp));
        }
    };
    auto sstore = [&](int buf) {
        uint8_t *base = smem + buf * STAGE;
#pragma unroll
        for (int h = 0; h < (SPLIT ? 2 : 1); ++h) {
#pragma unroll
            for (int i = 0; i < RA; ++i)
                *(i32x4 *)(base + h * OFF_LO + (lrow + 64 * i) * LDSROW + lchunk) = ra[h][i];
#pragma unroll
            for (int i = 0; i < RB; ++i)
                *(i32x4 *)(base + h * OFF_LO + OFF_B + (lrow + 64 * i) * LDSROW + lchunk) = rb[h][i];
        }
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    constexpr int UNIT = 127;            // E8M0 exponent of 1.0
    constexpr int LO = 127 - 4;          // 2^-4: weight of the fp8 residual terms
    if (nstages > 0) {
        gload(0);
        sstore(0);
    }
    __syncthreads();
    const int lr = lane & 31, lk = (lane >> 5) * OP::LANEB;
    for (int st = 0; st < nstages; ++st) {
        const int buf = st & 1;
        if (st + 1 < nstages) gload(st + 1);
        const uint8_t *base = smem + buf * STAGE;
        const uint8_t *pa = base + (wm * TM * 32 + lr) * LDSROW + lk;
        const uint8_t *pb = base + OFF_B + (wn * TN * 32 + lr) * LDSROW + lk;
#pragma unroll
        for (int kc = 0; kc < OP::CH; ++kc) {
            const int ko = kc * 2 * OP::LANEB;
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const uint8_t *ai = pa + i * 32 * LDSROW + ko;
                    const uint8_t *bj = pb + j * 32 * LDSROW + ko;
                    acc[i][j] = OP::run(ai, bj, acc[i][j], UNIT, UNIT);
                    if constexpr (SPLIT) {
                        constexpr int S1 = (MODE == GM_FP8X3) ? LO : UNIT;
                        acc[i][j] = OP::run(ai + OFF_LO, bj, acc[i][j], S1, UNIT);  // A lo * B hi
                        acc[i][j] = OP::run(ai, bj + OFF_LO, acc[i][j], UNIT, S1);  // A hi * B lo
                    }
                }
        }
        if (st + 1 < nstages) {
            __syncthreads();  // everyone is done reading buffer buf^1 (stage st-1)
            sstore(buf ^ 1);
        }
        __syncthreads();
    }

    const float alpha = (qa ? 1.0f / *qa : 1.0f) * (qb ? 1.0f / *qb : 1.0f);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * TM * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn * TN * 32 + j * 32 + lr;
                dst.put(m, n, acc[i][j][r] * alpha);
            }
}

template <class TL, int MODE, class ASrc, class BSrc, class Dst>
umi_status gemm_launch_t(ASrc ah, ASrc al, BSrc bh, BSrc bl, Dst d, int M, int N, int K, const float *qa,
                         const float *qb, hipStream_t st) {
    using OP = Mfma<MODE>;
    const int tm = ceil_div(M, TL::BM), tn = ceil_div(N, TL::BN);
    const size_t lds = (size_t)2 * (OP::NOP / 2) * (TL::BM + TL::BN) * LDSROW;
    auto kern = gemm_kernel<TL, MODE, ASrc, BSrc, Dst>;
    static bool attr_set = false;
    if (!attr_set) {
        UMI_TRY(hip_check(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                          "hipFuncSetAttribute(gemm LDS)"));
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(tm * tn), dim3(NTHREADS), lds, st, ah, al, bh, bl, d, K, tm, tn, qa, qb);
    return launch_check("gemm_kernel");
}

// Tile choice: the 256 x 256 tile when it yields at least one workgroup per
// CU, otherwise the 128 x 128 one; the 64 x 128 tile when M <= 64.
template <int MODE, class ASrc, class BSrc, class Dst>
umi_status gemm_launch(ASrc ah, ASrc al, BSrc bh, BSrc bl, Dst d, int M, int N, int K, const float *qa,
                       const float *qb, hipStream_t st) {
    if (M <= 64) return gemm_launch_t<TileNarrow, MODE>(ah, al, bh, bl, d, M, N, K, qa, qb, st);
    if ((long)ceil_div(M, 256) * ceil_div(N, 256) >= 256)
        return gemm_launch_t<TileGemm, MODE>(ah, al, bh, bl, d, M, N, K, qa, qb, st);
    return gemm_launch_t<TileMid, MODE>(ah, al, bh, bl, d, M, N, K, qa, qb, st);
}

int esize(int mode) { return mode == GM_F32 ? 4 : (mode <= GM_BF16X3 ? 2 : 1); }

// ---------------------------------------------------------------- packing
__device__ inline uint16_t f32_to_bf16_rne(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ inline uint8_t f32_to_e4m3(float f) {
    // hardware conversion (OCP e4m3 on gfx950), round to nearest even
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(f, 0.0f, 0, false) & 0xff);
}
__device__ inline float e4m3_to_f32(uint8_t b) { return __builtin_amdgcn_cvt_f32_fp8((int)b, 0); }

// dst[r][k] (K-contiguous, ld_dst elements) from src: transpose=0 -> src[r][k],
// transpose=1 -> src[k][r].  MODE selects element type and split.
__global__ void pack_kernel(const float *__restrict__ src, int R, int K, int ld_src, int transpose, int mode,
                            const float *__restrict__ amax, uint8_t *__restrict__ hi, uint8_t *__restrict__ lo,
                            int ld_dst, float *__restrict__ q_out) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)R * K) return;
    const int r = (int)(idx / K), k = (int)(idx - (long)r * K);
    const float v = transpose ? src[(long)k * ld_src + r] : src[(long)r * ld_src + k];
    const long o = (long)r * ld_dst + k;
    if (mode == GM_F32) {
        ((float *)hi)[o] = v;
    } else if (mode == GM_BF16 || mode == GM_BF16X3) {
        const uint16_t h = f32_to_bf16_rne(v);
        ((uint16_t *)hi)[o] = h;
        if (mode == GM_BF16X3) ((uint16_t *)lo)[o] = f32_to_bf16_rne(v - bf16_to_f32(h));
    } else {
        const float a = *amax;
        const float q = a > 0.f ? 448.0f / a : 1.0f;
        if (idx == 0 && q_out) *q_out = q;
        const float x = v * q;
        const uint8_t h = f32_to_e4m3(x);
        hi[o] = h;
        if (mode == GM_FP8X3) lo[o] = f32_to_e4m3((x - e4m3_to_f32(h)) * 16.0f);
    }
}

__global__ void amax_kernel(const float *__restrict__ src, long n, unsigned *__restrict__ out) {
    float m = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(src[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __builtin_bit_cast(unsigned, m));
}

// NCHW f32 -> NHWC packed operand (element type per mode), optional split.
__global__ void pack_nhwc_kernel(const float *__restrict__ src, int N, int C, int H, int W, int mode,
                                 const float *__restrict__ amax, uint8_t *__restrict__ hi, uint8_t *__restrict__ lo,
                                 float *__restrict__ q_out) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // NHWC order
    const long total = (long)N * C * H * W;
    if (idx >= total) return;
    const int c = (int)(idx % C);
    long t = idx / C;
    const int x = (int)(t % W);
    t /= W;
    const int y = (int)(t % H);
    const int n = (int)(t / H);
    const float v = src[(((long)n * C + c) * H + y) * W + x];
    if (mode == GM_F32) {
        ((float *)hi)[idx] = v;
    } else if (mode == GM_BF16 || mode == GM_BF16X3) {
        const uint16_t h = f32_to_bf16_rne(v);
        ((uint16_t *)hi)[idx] = h;
        if (mode == GM_BF16X3) ((uint16_t *)lo)[idx] = f32_to_bf16_rne(v - bf16_to_f32(h));
    } else {
        const float a = *amax;
        const float q = a > 0.f ? 448.0f / a : 1.0f;
        if (idx == 0 && q_out) *q_out = q;
        const float xq = v * q;
        const uint8_t h = f32_to_e4m3(xq);
        hi[idx] = h;
        if (mode == GM_FP8X3) lo[idx] = f32_to_e4m3((xq - e4m3_to_f32(h)) * 16.0f);
    }
}

umi_status check_mode(int mode, const char *who) {
    if (mode < GM_F32 || mode > GM_FP8X3) return set_error(UMI_ERR_INVALID, "%s: mode %d", who, mode);
    return UMI_OK;
}

bool split(int mode) { return mode == GM_BF16X3 || mode == GM_FP8X3; }

__global__ void zero_u32_kernel(unsigned *p) { *p = 0u; }

umi_status pack_common(int mode, const float *amax, void *hi, void *lo, const char *who) {
    UMI_TRY(check_mode(mode, who));
    if (!hi) return set_error(UMI_ERR_INVALID, "%s: null output", who);
    if (split(mode) && !lo) return set_error(UMI_ERR_INVALID, "%s: split mode needs a lo output", who);
    if ((mode == GM_FP8 || mode == GM_FP8X3) && !amax) return set_error(UMI_ERR_INVALID, "%s: fp8 needs amax", who);
    return UMI_OK;
}

}  // namespace

extern "C" {

umi_status umi_gemm_elem_bytes(int mode, int *bytes) {
    UMI_TRY(check_mode(mode, "umi_gemm_elem_bytes"));
    if (!bytes) return set_error(UMI_ERR_INVALID, "umi_gemm_elem_bytes: null pointer");
    *bytes = esize(mode);
    return UMI_OK;
}

umi_status umi_amax_f32(const float *src, size_t n, float *amax, umi_stream s) {
    if (!src || !amax) return set_error(UMI_ERR_INVALID, "umi_amax_f32: null pointer");
    hipStream_t st = as_stream(s);
    // cleared by a kernel rather than a memset so the sequence can be graph-captured
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, (unsigned *)amax);
    UMI_TRY(launch_check("zero_u32_kernel"));
    if (n == 0) return UMI_OK;
    const int blocks = (int)std::min<size_t>(2048, (n + 255) / 256);
    hipLaunchKernelGGL(amax_kernel, dim3(blocks), dim3(256), 0, st, src, (long)n, (unsigned *)amax);
    return launch_check("amax_kernel");
}

umi_status umi_gemm_pack(const float *src, int R, int K, int ld_src, int transpose, int mode, const float *amax,
                         void *dst_hi, void *dst_lo, int ld_dst, float *q_out, umi_stream s) {
    UMI_TRY(pack_common(mode, amax, dst_hi, dst_lo, "umi_gemm_pack"));
    if (!src || R < 0 || K < 0 || ld_dst < K || ld_src < (transpose ? R : K))
        return set_error(UMI_ERR_INVALID, "umi_gemm_pack: R=%d K=%d ld_src=%d ld_dst=%d", R, K, ld_src, ld_dst);
    const long n = (long)R * K;
    if (n == 0) return UMI_OK;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(s), src, R, K, ld_src,
                       transpose, mode, amax, (uint8_t *)dst_hi, (uint8_t *)dst_lo, ld_dst, q_out);
    return launch_check("pack_kernel");
}

umi_status umi_gemm(int mode, int M, int N, int K, const void *a_hi, const void *a_lo, int lda, const void *b_hi,
                    const void *b_lo, int ldb, float *C, int ldc, const float *qa, const float *qb, umi_stream s) {
    UMI_TRY(check_mode(mode, "umi_gemm"));
    if (!a_hi || !b_hi || !C || (split(mode) && (!a_lo || !b_lo)))
        return set_error(UMI_ERR_INVALID, "umi_gemm: null operand");
    if (M < 0 || N < 0 || K < 0 || lda < K || ldb < K || ldc < N)
        return set_error(UMI_ERR_INVALID, "umi_gemm: M=%d N=%d K=%d lda=%d ldb=%d ldc=%d", M, N, K, lda, ldb, ldc);
    const int e = esize(mode);
    if ((K * e) % ROWB != 0 || (lda * e) % 16 != 0 || (ldb * e) % 16 != 0)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_gemm: K*%d and lda/ldb*%d must be multiples of 64 / 16 bytes", e, e);
    if (M == 0 || N == 0) return UMI_OK;
    if ((mode == GM_FP8 || mode == GM_FP8X3) && (!qa || !qb))
        return set_error(UMI_ERR_INVALID, "umi_gemm: fp8 modes need the quantisation scales");
    const DenseSrc ah{(const uint8_t *)a_hi, (long)lda * e, M}, al{(const uint8_t *)a_lo, (long)lda * e, M};
    const DenseSrc bh{(const uint8_t *)b_hi, (long)ldb * e, N}, bl{(const uint8_t *)b_lo, (long)ldb * e, N};
    const DenseDst d{C, ldc, M, N};
    hipStream_t st = as_stream(s);
    switch (mode) {
        case GM_F32: return gemm_launch<GM_F32>(ah, al, bh, bl, d, M, N, K, nullptr, nullptr, st);
        case GM_BF16: return gemm_launch<GM_BF16>(ah, al, bh, bl, d, M, N, K, nullptr, nullptr, st);
        case GM_BF16X3: return gemm_launch<GM_BF16X3>(ah, al, bh, bl, d, M, N, K, nullptr, nullptr, st);
        case GM_FP8: return gemm_launch<GM_FP8>(ah, al, bh, bl, d, M, N, K, qa, qb, st);
        default: return gemm_launch<GM_FP8X3>(ah, al, bh, bl, d, M, N, K, qa, qb, st);
    }
}

umi_status umi_conv_pack_input(const float *src, int N, int C, int H, int W, int mode, const float *amax, void *hi,
                               void *lo, float *q_out, umi_stream s) {
    UMI_TRY(pack_common(mode, amax, hi, lo, "umi_conv_pack_input"));
    if (!src || N < 0 || C < 1 || H < 0 || W < 0)
        return set_error(UMI_ERR_INVALID, "umi_conv_pack_input: N=%d C=%d H=%d W=%d", N, C, H, W);
    const long n = (long)N * C * H * W;
    if (n == 0) return UMI_OK;
    hipLaunchKernelGGL(pack_nhwc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(s), src, N, C, H,
                       W, mode, amax, (uint8_t *)hi, (uint8_t *)lo, q_out);
    return launch_check("pack_nhwc_kernel");
}

umi_status umi_conv_pack_filter(const float *src, int K, int C, int R, int S, int mode, const float *amax, void *hi,
                                void *lo, float *q_out, umi_stream s) {
    // KCRS -> KRSC: the same NCHW -> NHWC reordering with (n, c, h, w) = (k, c, r, s)
    return umi_conv_pack_input(src, K, C, R, S, mode, amax, hi, lo, q_out, s);
}

umi_status umi_conv_fprop(int mode, int N, int C, int H, int W, int K, int R, int S, int stride, int pad,
                          const void *in_hi, const void *in_lo, const void *w_hi, const void *w_lo, float *out,
                          const float *qi, const float *qw, umi_stream s) {
    UMI_TRY(check_mode(mode, "umi_conv_fprop"));
    if (!in_hi || !w_hi || !out || (split(mode) && (!in_lo || !w_lo)))
        return set_error(UMI_ERR_INVALID, "umi_conv_fprop: null operand");
    if (N < 0 || C < 1 || H < 1 || W < 1 || K < 1 || R < 1 || S < 1 || stride < 1 || pad < 0)
        return set_error(UMI_ERR_INVALID, "umi_conv_fprop: bad shape");
    const int e = esize(mode);
    const int OH = (H + 2 * pad - R) / stride + 1, OW = (W + 2 * pad - S) / stride + 1;
    const int KG = R * S * C;
    if ((C * e) % 16 != 0 || (KG * e) % ROWB != 0)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_conv_fprop: C*%d must be a multiple of 16 B and R*S*C*%d of 64 B", e, e);
    if (OH < 1 || OW < 1 || N == 0) return UMI_OK;
    if ((mode == GM_FP8 || mode == GM_FP8X3) && (!qi || !qw))
        return set_error(UMI_ERR_INVALID, "umi_conv_fprop: fp8 modes need the quantisation scales");
    const DenseSrc ah{(const uint8_t *)w_hi, (long)KG * e, K}, al{(const uint8_t *)w_lo, (long)KG * e, K};
    const ConvSrc bh{(const uint8_t *)in_hi, N, H, W, C, OH, OW, R, S, stride, pad, 1};
    const ConvSrc bl{(const uint8_t *)in_lo, N, H, W, C, OH, OW, R, S, stride, pad, 1};
    const ConvDst d{out, K, OH * OW, N * OH * OW};
    const int NP = N * OH * OW;
    hipStream_t st = as_stream(s);
    switch (mode) {
        case GM_F32: return gemm_launch<GM_F32>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_BF16: return gemm_launch<GM_BF16>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_BF16X3: return gemm_launch<GM_BF16X3>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_FP8: return gemm_launch<GM_FP8>(ah, al, bh, bl, d, K, NP, KG, qw, qi, st);
        default: return gemm_launch<GM_FP8X3>(ah, al, bh, bl, d, K, NP, KG, qw, qi, st);
    }
}

}  // extern "C"
