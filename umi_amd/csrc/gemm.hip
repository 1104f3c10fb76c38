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
#include <type_traits>
#include <utility>

namespace {
using namespace umi;

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>), in order.
template <class F, int... I>
__device__ inline void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ inline void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) int i32x8;
typedef __attribute__((ext_vector_type(4))) int i32x4;

enum GemmMode { GM_F32 = 0, GM_BF16 = 1, GM_BF16X3 = 2, GM_FP8 = 3, GM_FP8X3 = 4 };

constexpr int ROWB = 64;    // bytes of one operand row per stage
constexpr int LDSROW = ROWB; // LDS row stride; the 16-byte chunks are swizzled instead of padded
constexpr int NTHREADS = 256;

// LDS position of byte `off` of a stage row whose swizzle is sw = (row >> 2) & 3:
// 16-byte chunk c lives at chunk c ^ sw.  Sixteen lanes that write 4 rows x
// 4 chunks, or read one chunk of 16 consecutive rows, then touch 16 distinct
// groups of 4 banks.
__device__ inline int swz(int off, int sw) { return (((off >> 4) ^ sw) << 4) | (off & 15); }

template <int MODE>
struct Mfma;
template <>
struct Mfma<GM_F32> {
    static constexpr int ESZ = 4, LANEB = 4, CH = 8, NOP = 2;
    // operand fragment of one lane and its LDS reads; the MFMA on fragments
    using Frag = float;
    static constexpr int NRD = 1;
    static __device__ Frag load(const uint8_t *row, int off, int sw) { return *(const float *)(row + swz(off, sw)); }
    static __device__ f32x16 mma(Frag a, Frag b, f32x16 c, int, int) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // a, b: the lane's LDS rows; off: byte offset in the row; sw: the rows' swizzle
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, int off, int sw, f32x16 c, int, int) {
        const int o = swz(off, sw);
        return __builtin_amdgcn_mfma_f32_32x32x2f32(*(const float *)(a + o), *(const float *)(b + o), c, 0, 0, 0);
    }
};
template <>
struct Mfma<GM_BF16> {
    static constexpr int ESZ = 2, LANEB = 16, CH = 2, NOP = 2;
    using Frag = bf16x8;
    static constexpr int NRD = 1;
    static __device__ Frag load(const uint8_t *row, int off, int sw) { return *(const bf16x8 *)(row + swz(off, sw)); }
    static __device__ f32x16 mma(Frag a, Frag b, f32x16 c, int, int) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, int off, int sw, f32x16 c, int, int) {
        const int o = swz(off, sw);
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*(const bf16x8 *)(a + o), *(const bf16x8 *)(b + o), c, 0, 0,
                                                       0);
    }
};
template <>
struct Mfma<GM_BF16X3> : Mfma<GM_BF16> {
    static constexpr int NOP = 4;
};
template <>
struct Mfma<GM_FP8> {
    static constexpr int ESZ = 1, LANEB = 32, CH = 1, NOP = 2;
    using Frag = i32x8;
    static constexpr int NRD = 2;
    static __device__ Frag load(const uint8_t *row, int off, int sw) {
        const i32x4 q0 = *(const i32x4 *)(row + swz(off, sw)), q1 = *(const i32x4 *)(row + swz(off + 16, sw));
        return i32x8{q0[0], q0[1], q0[2], q0[3], q1[0], q1[1], q1[2], q1[3]};
    }
    static __device__ f32x16 mma(Frag a, Frag b, f32x16 c, int sa, int sb) {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
    }
    // 32 bytes = two chunks, which the swizzle may place in either order
    static __device__ f32x16 run(const uint8_t *a, const uint8_t *b, int off, int sw, f32x16 c, int sa, int sb) {
        const int o0 = swz(off, sw), o1 = swz(off + 16, sw);
        const i32x4 a0 = *(const i32x4 *)(a + o0), a1 = *(const i32x4 *)(a + o1);
        const i32x4 b0 = *(const i32x4 *)(b + o0), b1 = *(const i32x4 *)(b + o1);
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
// floor(x / d) = umulhi(x, recip32(d)) for d >= 2 and x * d < 2^32: the
// multiplier is ceil(2^32 / d) rounded from below, so x * m / 2^32 exceeds
// x / d by less than x / 2^32 <= 1 / d.
inline uint32_t recip32(int d) { return d > 1 ? 0xffffffffu / (uint32_t)d + 1u : 0u; }

struct ConvSrc {
    const uint8_t *base;  // NHWC input
    int N, H, W, C;       // C in elements
    int OH, OW, R, S, stride, pad, dil;
    uint32_t mC, mS;      // recip32(C), recip32(S): the per-stage divisions by multiply-high
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
        const int rs = (int)__umulhi((uint32_t)k, mC), c = k - rs * C;  // C >= 4
        const int r = S == 1 ? rs : (int)__umulhi((uint32_t)rs, mS), s = rs - r * S;
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
    struct Col {
        int n;
        bool ok;
    };
    __device__ Col col(int n) const { return {n, n < N}; }
    __device__ void put(int m, const Col &c, float v) const {
        if (m < M && c.ok) base[(long)m * ldc + c.n] = v;
    }
};
// NHWC bf16 output with per-channel bias and optional ReLU (the layer
// output of the motion-proposal network, fed to the next layer as is).
struct ConvDst {
    float *base;  // [N][K][OH][OW]
    int K, npix, total;
    struct Col {
        long off;  // element offset of (image n, channel 0, pixel p)
        bool ok;
    };
    __device__ Col col(int j) const {
        const int n = j / npix, p = j - n * npix;
        return {(long)n * K * npix + p, j < total};
    }
    __device__ void put(int m, const Col &c, float v) const {
        if (m < K && c.ok) base[c.off + (long)m * npix] = v;
    }
};

// 16 bytes of one panel row.  ZERO_NOW: a load only for present rows, zeros
// otherwise (the form for one stage in flight).  Otherwise the load is
// unconditional (`safe`, the operand's own base, stands in for an absent
// row) and the caller zeroes absent rows only when it stores the chunk to
// LDS: a branch or a select right after the load makes the wait counters
// drain every load in flight, which defeats a deeper prefetch.
template <bool ZERO_NOW>
__device__ inline i32x4 fetch16(const uint8_t *p, const uint8_t *safe) {
    if constexpr (ZERO_NOW) return p ? *(const i32x4 *)p : i32x4{0, 0, 0, 0};
    return *(const i32x4 *)(p ? p : safe);
}

// Workgroup tile: WGM x WGN waves, each wave TM x TN MFMA blocks of 32 x 32;
// PF K stages of global loads are kept in flight in registers (a stage's
// MFMA work is short next to a global load's latency, so one stage of
// prefetch leaves the loop waiting on every load).
template <int WGM_, int WGN_, int TM_, int TN_, int PF_>
struct Tile {
    static constexpr int WGM = WGM_, WGN = WGN_, TM = TM_, TN = TN_, PF = PF_;
    static constexpr int BM = WGM * TM * 32, BN = WGN * TN * 32;
    static_assert(WGM * WGN * 64 == NTHREADS, "four waves");
};
using TileGemm = Tile<2, 2, 4, 4, 1>;   // 256 x 256: large GEMMs (the accumulators fill the register file)
using TileMid = Tile<2, 2, 2, 2, 2>;    // 128 x 128
using TileNarrow = Tile<1, 4, 2, 1, 4>; // 64 x 128: few output rows (conv with K <= 64)

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
    // global -> LDS assignment: thread t copies 16 B chunk (t & 3) of rows t/4 + 64 i;
    // stage s travels in register slot s % PF (at most 2 for the split forms,
    // whose hi and lo panels double the ring's registers)
    constexpr int PF = SPLIT && TL::PF > 2 ? 2 : TL::PF;
    const int lrow = tid >> 2, lchunk = (tid & 3) * 16;
    i32x4 ra[PF][SPLIT ? 2 : 1][RA], rb[PF][SPLIT ? 2 : 1][RB];
    // false: absent row, stored as zeros (PF == 1 zeroes at load time: the
    // 256 x 256 tile has no registers to spare for the flags)
    constexpr bool ZN = PF == 1;
    bool oka[PF][SPLIT ? 2 : 1][RA], okb[PF][SPLIT ? 2 : 1][RB];
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
    auto gload = [&](int st, auto slot_tag) {
        constexpr int SL = decltype(slot_tag)::value;
        const int kb = st * ROWB + lchunk;
        const auto ka = a_hi.kprep(kb, OP::ESZ);
        const auto kbp = b_hi.kprep(kb, OP::ESZ);
#pragma unroll
        for (int i = 0; i < RA; ++i) {
            const uint8_t *p = a_hi.at(pa_[i], ka);
            ra[SL][0][i] = fetch16<ZN>(p, a_hi.base);
            oka[SL][0][i] = ZN || p != nullptr;
            if constexpr (SPLIT) {
                const uint8_t *q = a_lo.at(pal_[i], ka);
                ra[SL][1][i] = fetch16<ZN>(q, a_lo.base);
                oka[SL][1][i] = ZN || q != nullptr;
            }
        }
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const uint8_t *p = b_hi.at(pb_[i], kbp);
            rb[SL][0][i] = fetch16<ZN>(p, b_hi.base);
            okb[SL][0][i] = ZN || p != nullptr;
            if constexpr (SPLIT) {
                const uint8_t *q = b_lo.at(pbl_[i], kbp);
                rb[SL][1][i] = fetch16<ZN>(q, b_lo.base);
                okb[SL][1][i] = ZN || q != nullptr;
            }
        }
    };
    // the thread's chunk, swizzled by its rows (lrow + 64 i share (row >> 2) & 3)
    const int schunk = swz(lchunk, (lrow >> 2) & 3);
    auto sstore = [&](int buf, auto slot_tag) {
        constexpr int SL = decltype(slot_tag)::value;
        uint8_t *base = smem + buf * STAGE;
#pragma unroll
        for (int h = 0; h < (SPLIT ? 2 : 1); ++h) {
#pragma unroll
            for (int i = 0; i < RA; ++i)
                *(i32x4 *)(base + h * OFF_LO + (lrow + 64 * i) * LDSROW + schunk) =
                    oka[SL][h][i] ? ra[SL][h][i] : i32x4{0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < RB; ++i)
                *(i32x4 *)(base + h * OFF_LO + OFF_B + (lrow + 64 * i) * LDSROW + schunk) =
                    okb[SL][h][i] ? rb[SL][h][i] : i32x4{0, 0, 0, 0};
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
    const int lr = lane & 31, lk = (lane >> 5) * OP::LANEB;
    const int rsw = (lr >> 2) & 3;  // swizzle of the lane's rows (tile rows are multiples of 32 apart)
    auto compute = [&](int buf) {
        const uint8_t *base = smem + buf * STAGE;
        const uint8_t *pa = base + (wm * TM * 32 + lr) * LDSROW;
        const uint8_t *pb = base + OFF_B + (wn * TN * 32 + lr) * LDSROW;
#pragma unroll
        for (int kc = 0; kc < OP::CH; ++kc) {
            const int off = kc * 2 * OP::LANEB + lk;
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const uint8_t *ai = pa + i * 32 * LDSROW;
                    const uint8_t *bj = pb + j * 32 * LDSROW;
                    acc[i][j] = OP::run(ai, bj, off, rsw, acc[i][j], UNIT, UNIT);
                    if constexpr (SPLIT) {
                        constexpr int S1 = (MODE == GM_FP8X3) ? LO : UNIT;
                        acc[i][j] = OP::run(ai + OFF_LO, bj, off, rsw, acc[i][j], S1, UNIT);  // A lo * B hi
                        acc[i][j] = OP::run(ai, bj + OFF_LO, off, rsw, acc[i][j], UNIT, S1);  // A hi * B lo
                    }
                }
        }
    };
    // Prologue: stages 0 .. PF-1 in flight (every slot filled, repeating the
    // last stage when there are fewer), stage 0 in LDS buffer 0, stage PF
    // into the freed slot 0.
    if (nstages > 0)
        static_for<PF>([&](auto d) { gload(min((int)decltype(d)::value, nstages - 1), d); });
    if (nstages > 0) {
        sstore(0, std::integral_constant<int, 0>{});
        if (PF > 1 || PF < nstages) gload(min(PF, nstages - 1), std::integral_constant<int, 0>{});
    }
    __syncthreads();
    // Step st: store stage st+1 (slot (st+1) % PF) into the other buffer, which
    // every wave finished reading before the previous barrier; refill the slot
    // with stage st+1+PF; multiply stage st; one barrier.  Whole groups of PF
    // steps run without branches around the loads (a refill past the last
    // stage repeats the last stage, a store past it fills a buffer nobody
    // reads), so the wait counters keep PF stages in flight; the last,
    // partial group checks each step.
    auto step = [&](int st, auto slot_tag) {
        const int buf = st & 1;
        sstore(buf ^ 1, slot_tag);
        gload(min(st + 1 + PF, nstages - 1), slot_tag);
        compute(buf);
        __syncthreads();
    };
    if constexpr (PF == 1) {
        // one stage in flight: a plain loop (the 256 x 256 tile's registers
        // are full, and the grouped form below makes it spill)
        for (int st = 0; st < nstages; ++st) {
            const int buf = st & 1;
            if (st + 1 < nstages) {
                sstore(buf ^ 1, std::integral_constant<int, 0>{});
                if (st + 2 < nstages) gload(st + 2, std::integral_constant<int, 0>{});
            }
            compute(buf);
            __syncthreads();
        }
    } else {
        int st0 = 0;
        for (; st0 + PF <= nstages; st0 += PF)
            static_for<PF>([&](auto d) {
                step(st0 + decltype(d)::value, std::integral_constant<int, (decltype(d)::value + 1) % PF>{});
            });
        static_for<PF>([&](auto d) {
            if (st0 + decltype(d)::value < nstages)
                step(st0 + decltype(d)::value, std::integral_constant<int, (decltype(d)::value + 1) % PF>{});
        });
    }

    const float alpha = (qa ? 1.0f / *qa : 1.0f) * (qb ? 1.0f / *qb : 1.0f);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const auto cj = dst.col(n0 + wn * TN * 32 + j * 32 + lr);  // one column per lane and block
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * TM * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                dst.put(m, cj, acc[i][j][r] * alpha);
            }
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

// Dense 256 x 256 tiles with direct-to-LDS loads (global_load_lds_dwordx4):
// the same operands, LDS layout (64-byte rows, 16-byte chunks swizzled by
// (row >> 2) & 3), MFMA sequence and epilogue as gemm_kernel<TileGemm>, so
// the results are bit-identical; what changes is the staging.  A wave's
// direct load writes 1 KB of LDS in lane order (16 rows x 64 B), so a lane
// fetches the global chunk that its LDS slot holds after the swizzle; no
// register ring, no LDS stores.  NS stages of LDS (4 x 32 KB; 2 x 64 KB for
// the split forms' hi and lo panels) keep NS - 1 stages of loads in flight:
// step st waits (s_waitcnt vmcnt) for its own stage only, one barrier, then
// refills the buffer freed by step st - 1 with stage st + NS - 1.
bool gemm_lds_direct = true;  // umi_gemm_set_lds_direct

template <int N>
__device__ inline void wait_vm() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// WGM x WGN waves of TM x TN MFMA blocks (32 x 32): 2 x 2 waves of 4 x 4
// (one wave per SIMD) or 2 x 4 waves of 4 x 2 (two per SIMD, half the
// accumulators each); both cover 256 x 256.
template <int WGM_, int WGN_, int TM_, int TN_>
struct LdsTile {
    static constexpr int WGM = WGM_, WGN = WGN_, TM = TM_, TN = TN_;
    static constexpr int NT = WGM * WGN * 64, BM = WGM * TM * 32, BN = WGN * TN * 32;
    static_assert(BM == 256 && BN == 256, "256 x 256 tiles");
};
using LdsTile4 = LdsTile<2, 2, 4, 4>;
using LdsTile8 = LdsTile<2, 4, 4, 2>;

template <int MODE, class TL>
__global__ __launch_bounds__(TL::NT) void gemm_lds_kernel(const uint8_t *__restrict__ a_hi,
                                                            const uint8_t *__restrict__ a_lo, long lda,
                                                            const uint8_t *__restrict__ b_hi,
                                                            const uint8_t *__restrict__ b_lo, long ldb, DenseDst dst,
                                                            int M, int N, int K, int tiles_m, int tiles_n,
                                                            const float *qa, const float *qb) {
    using OP = Mfma<MODE>;
    constexpr bool SPLIT = OP::NOP == 4;
    constexpr int NS = SPLIT ? 2 : 4;
    constexpr int BMt = TL::BM, BNt = TL::BN, TM = TL::TM, TN = TL::TN;
    constexpr int OFF_B = BMt * LDSROW, OFF_LO = (BMt + BNt) * LDSROW;
    constexpr int STAGE = (SPLIT ? 2 : 1) * (BMt + BNt) * LDSROW;
    constexpr int NW = TL::NT / 64;           // waves
    constexpr int WL = BMt / 16 / NW;         // direct loads of 16 rows per wave per panel
    constexpr int LPS = (SPLIT ? 4 : 2) * WL; // direct loads per thread per stage
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wv / TL::WGN, wn = wv % TL::WGN;

    // tile order as gemm_kernel: XCD-aware, then grouped raster
    const int ntiles = tiles_m * tiles_n;
    int pid = blockIdx.x;
    if (ntiles % 8 == 0) pid = (pid % 8) * (ntiles / 8) + pid / 8;
    const int GROUP = 4;
    const int group = pid / (GROUP * tiles_n), first_m = group * GROUP;
    const int gsz = min(GROUP, tiles_m - first_m);
    const int tm = first_m + (pid % (GROUP * tiles_n)) % gsz;
    const int tn = (pid % (GROUP * tiles_n)) / gsz;
    const int m0 = tm * BMt, n0 = tn * BNt;
    const int nstages = K * OP::ESZ / ROWB;

    // lane's source rows: load i of the wave covers panel rows (wv * WL + i) * 16 .. + 15;
    // the lane's LDS slot (row r, chunk slot lane & 3) holds global chunk slot ^ sw(r)
    const int lrow = lane >> 2, lslot = lane & 3;
    const uint8_t *srcA[WL], *srcB[WL], *srcAl[WL], *srcBl[WL];
#pragma unroll
    for (int i = 0; i < WL; ++i) {
        const int r = (wv * WL + i) * 16 + lrow;
        const int chunk = (lslot ^ ((r >> 2) & 3)) * 16;
        const long ra = (long)min(m0 + r, M - 1) * lda + chunk, rb = (long)min(n0 + r, N - 1) * ldb + chunk;
        srcA[i] = a_hi + ra;
        srcB[i] = b_hi + rb;
        if constexpr (SPLIT) {
            srcAl[i] = a_lo + ra;
            srcBl[i] = b_lo + rb;
        }
    }
    auto issue = [&](int st, int buf) {
        uint8_t *base = smem + buf * STAGE;
        const long kb = (long)st * ROWB;
#pragma unroll
        for (int i = 0; i < WL; ++i) {
            const int row0 = (wv * WL + i) * 16 * LDSROW;
            __builtin_amdgcn_global_load_lds((const void *)(srcA[i] + kb),
                                             (__attribute__((address_space(3))) void *)(base + row0), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(srcB[i] + kb),
                                             (__attribute__((address_space(3))) void *)(base + OFF_B + row0), 16, 0, 0);
            if constexpr (SPLIT) {
                __builtin_amdgcn_global_load_lds((const void *)(srcAl[i] + kb),
                                                 (__attribute__((address_space(3))) void *)(base + OFF_LO + row0), 16,
                                                 0, 0);
                __builtin_amdgcn_global_load_lds(
                    (const void *)(srcBl[i] + kb),
                    (__attribute__((address_space(3))) void *)(base + OFF_LO + OFF_B + row0), 16, 0, 0);
            }
        }
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    constexpr int UNIT = 127, LO = 127 - 4;
    const int lr = lane & 31, lk = (lane >> 5) * OP::LANEB;
    const int rsw = (lr >> 2) & 3;
    auto compute = [&](int buf) {
        const uint8_t *base = smem + buf * STAGE;
        const uint8_t *pa = base + (wm * TM * 32 + lr) * LDSROW;
        const uint8_t *pb = base + OFF_B + (wn * TN * 32 + lr) * LDSROW;
#pragma unroll
        for (int kc = 0; kc < OP::CH; ++kc) {
            const int off = kc * 2 * OP::LANEB + lk;
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const uint8_t *ai = pa + i * 32 * LDSROW;
                    const uint8_t *bj = pb + j * 32 * LDSROW;
                    acc[i][j] = OP::run(ai, bj, off, rsw, acc[i][j], UNIT, UNIT);
                    if constexpr (SPLIT) {
                        constexpr int S1 = (MODE == GM_FP8X3) ? LO : UNIT;
                        acc[i][j] = OP::run(ai + OFF_LO, bj, off, rsw, acc[i][j], S1, UNIT);
                        acc[i][j] = OP::run(ai, bj + OFF_LO, off, rsw, acc[i][j], UNIT, S1);
                    }
                }
        }
    };
    for (int st = 0; st < NS - 1 && st < nstages; ++st) issue(st, st);
    for (int st = 0; st < nstages; ++st) {
        // stage st has landed once at most the loads of the later stages issued so far are outstanding
        if (st + NS - 2 < nstages) wait_vm<LPS * (NS - 2)>();
        else wait_vm<0>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();  // every wave's stage st is in LDS; buffer (st - 1) % NS is free
        asm volatile("" ::: "memory");
        if (st + NS - 1 < nstages) issue(st + NS - 1, (st + NS - 1) % NS);
        compute(st % NS);
    }

    const float alpha = (qa ? 1.0f / *qa : 1.0f) * (qb ? 1.0f / *qb : 1.0f);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const auto cj = dst.col(n0 + wn * TN * 32 + j * 32 + lr);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * TM * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                dst.put(m, cj, acc[i][j][r] * alpha);
            }
    }
}

// The same tiles and LDS ring with the operand reads software-pipelined
// (single-product forms): a wave holds two register slots of MFMA
// fragments and, while the MFMAs of one K chunk run from one slot, reads
// the next chunk into the other, one LDS read issued after each MFMA
// (sched_group_barrier), so no MFMA waits on an LDS read of its own chunk.
// A stage's LDS buffer is refilled once every wave has read its last chunk
// into registers: at the barrier before the reads of stage st + 1, stage
// st + NS - 1 is issued into the buffer stage st - 1 used.  Stages st + 2
// .. st + NS - 2 stay in flight while a wave waits for stage st + 1.
template <int MODE, class TL, int NS>
__global__ __launch_bounds__(TL::NT) void gemm_pipe_kernel(const uint8_t *__restrict__ a_hi, long lda,
                                                             const uint8_t *__restrict__ b_hi, long ldb, DenseDst dst,
                                                             int M, int N, int K, int tiles_m, int tiles_n,
                                                             const float *qa, const float *qb) {
    using OP = Mfma<MODE>;
    static_assert(OP::NOP == 2, "single-product forms");
    using Frag = typename OP::Frag;
    constexpr int BMt = TL::BM, BNt = TL::BN, TM = TL::TM, TN = TL::TN, CH = OP::CH;
    constexpr int OFF_B = BMt * LDSROW, STAGE = (BMt + BNt) * LDSROW;
    constexpr int NW = TL::NT / 64;
    constexpr int WL = BMt / 16 / NW;  // direct loads of 16 rows per wave per panel
    constexpr int LPS = 2 * WL;        // direct loads per thread per stage
    constexpr int NMF = TM * TN, NRD = (TM + TN) * OP::NRD;  // MFMAs and LDS reads per chunk
    static_assert(NS >= 4, "at least two stages in flight");
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wv / TL::WGN, wn = wv % TL::WGN;

    const int ntiles = tiles_m * tiles_n;
    int pid = blockIdx.x;
    if (ntiles % 8 == 0) pid = (pid % 8) * (ntiles / 8) + pid / 8;
    const int GROUP = 4;
    const int group = pid / (GROUP * tiles_n), first_m = group * GROUP;
    const int gsz = min(GROUP, tiles_m - first_m);
    const int tm = first_m + (pid % (GROUP * tiles_n)) % gsz;
    const int tn = (pid % (GROUP * tiles_n)) / gsz;
    const int m0 = tm * BMt, n0 = tn * BNt;
    const int nstages = K * OP::ESZ / ROWB;

    const int lrow = lane >> 2, lslot = lane & 3;
    const uint8_t *srcA[WL], *srcB[WL];
#pragma unroll
    for (int i = 0; i < WL; ++i) {
        const int r = (wv * WL + i) * 16 + lrow;
        const int chunk = (lslot ^ ((r >> 2) & 3)) * 16;
        srcA[i] = a_hi + (long)min(m0 + r, M - 1) * lda + chunk;
        srcB[i] = b_hi + (long)min(n0 + r, N - 1) * ldb + chunk;
    }
    auto issue = [&](int st) {
        uint8_t *base = smem + (st % NS) * STAGE;
        const long kb = (long)st * ROWB;
#pragma unroll
        for (int i = 0; i < WL; ++i) {
            const int row0 = (wv * WL + i) * 16 * LDSROW;
            __builtin_amdgcn_global_load_lds((const void *)(srcA[i] + kb),
                                             (__attribute__((address_space(3))) void *)(base + row0), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(srcB[i] + kb),
                                             (__attribute__((address_space(3))) void *)(base + OFF_B + row0), 16, 0, 0);
        }
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    constexpr int UNIT = 127;
    const int lr = lane & 31, lk = (lane >> 5) * OP::LANEB;
    const int rsw = (lr >> 2) & 3;
    const int arow = (wm * TM * 32 + lr) * LDSROW, brow = OFF_B + (wn * TN * 32 + lr) * LDSROW;
    Frag fa[2][TM], fb[2][TN];
    // chunk kc of stage st into register slot S
    auto read_chunk = [&](auto S, int st, int kc) {
        const uint8_t *base = smem + (st % NS) * STAGE;
        const int off = kc * 2 * OP::LANEB + lk;
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[S][i] = OP::load(base + arow + i * 32 * LDSROW, off, rsw);
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[S][j] = OP::load(base + brow + j * 32 * LDSROW, off, rsw);
    };
    // MFMAs of slot S, the reads of the next chunk interleaved one per MFMA
    auto mma_slot = [&](auto S) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = OP::mma(fa[S][i], fb[S][j], acc[i][j], UNIT, UNIT);
        static_for<NMF>([&](auto t) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            constexpr int lo = (int)t * NRD / NMF, hi = ((int)t + 1) * NRD / NMF;
            if constexpr (hi > lo) __builtin_amdgcn_sched_group_barrier(0x100, hi - lo, 0);  // LDS reads
        });
    };
    // one stage; S0 = the slot its first chunk is in.  Every read is
    // unconditional: past the last stage the chunk read is a harmless repeat.
    auto stage = [&](auto S0, int st) {
        static_for<CH>([&](auto kc) {
            constexpr int S = ((int)S0 + (int)kc) & 1;
            if constexpr ((int)kc == CH - 1) {
                // stage st + 1 lands; buffer (st - 1) % NS is free in every wave
                if (st + NS - 2 < nstages) wait_vm<LPS * (NS - 3)>();
                else wait_vm<0>();
                asm volatile("" ::: "memory");
                __builtin_amdgcn_s_barrier();
                asm volatile("" ::: "memory");
                if (st + NS - 1 < nstages) issue(st + NS - 1);
                read_chunk(std::integral_constant<int, 1 - S>{}, min(st + 1, nstages - 1), 0);
            } else {
                read_chunk(std::integral_constant<int, 1 - S>{}, st, (int)kc + 1);
            }
            mma_slot(std::integral_constant<int, S>{});
        });
    };

    for (int st = 0; st < NS - 1 && st < nstages; ++st) issue(st);
    if (NS - 2 < nstages) wait_vm<LPS * (NS - 2)>();
    else wait_vm<0>();
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    read_chunk(std::integral_constant<int, 0>{}, 0, 0);
    if constexpr (CH % 2 == 0) {
        for (int st = 0; st < nstages; ++st) stage(std::integral_constant<int, 0>{}, st);
    } else {
        int st = 0;
        for (; st + 1 < nstages; st += 2) {
            stage(std::integral_constant<int, 0>{}, st);
            stage(std::integral_constant<int, 1>{}, st + 1);
        }
        if (st < nstages) stage(std::integral_constant<int, 0>{}, st);
    }

    const float alpha = (qa ? 1.0f / *qa : 1.0f) * (qb ? 1.0f / *qb : 1.0f);
    // buffer stores: rows past M fall outside the resource (M * ldc floats,
    // < 2^31 bytes, checked by the launcher) and are dropped by the
    // hardware; a lane past N stores at 2^31, also outside
    const __amdgpu_buffer_rsrc_t rc =
        __builtin_amdgcn_make_buffer_rsrc((void *)dst.base, 0, (int)((long)M * dst.ldc * 4), 0x00020000);
    const int ldb4 = dst.ldc * 4;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * TN * 32 + j * 32 + lr;
        const int vo = n < N ? (m0 + wm * TM * 32 + 4 * (lane >> 5)) * ldb4 + n * 4 : (int)0x80000000u;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int mr = i * 32 + 8 * (r >> 2) + (r & 3);  // row within the wave's block column
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, acc[i][j][r] * alpha), rc, vo,
                                                      mr * ldb4, 0);
            }
    }
}

// umi_gemm_set_lds_direct: 1 = pipelined operand reads (single-product
// forms), 8 waves; 5 = the same with 4 waves; 2 / 4 = the LDS ring without
// the register pipeline, 8 / 4 waves (the split forms always use the ring)
int gemm_lds_form = 1;

template <int MODE, class TL>
umi_status gemm_lds_launch_t(const uint8_t *ah, const uint8_t *al, long lda, const uint8_t *bh, const uint8_t *bl,
                             long ldb, DenseDst d, int M, int N, int K, const float *qa, const float *qb,
                             hipStream_t st) {
    using OP = Mfma<MODE>;
    constexpr int NS = OP::NOP == 4 ? 2 : 4;
    const size_t lds = (size_t)NS * (OP::NOP / 2) * (TL::BM + TL::BN) * LDSROW;
    static const hipError_t attr = hipFuncSetAttribute((const void *)gemm_lds_kernel<MODE, TL>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    UMI_TRY(hip_check(attr, "hipFuncSetAttribute(gemm LDS)"));
    const int tm = ceil_div(M, TL::BM), tn = ceil_div(N, TL::BN);
    hipLaunchKernelGGL((gemm_lds_kernel<MODE, TL>), dim3(tm * tn), dim3(TL::NT), lds, st, ah, al, lda, bh, bl, ldb, d,
                       M, N, K, tm, tn, qa, qb);
    return launch_check("gemm_lds_kernel");
}

template <int MODE, class TL, int NS>
umi_status gemm_pipe_launch_t(const uint8_t *ah, long lda, const uint8_t *bh, long ldb, DenseDst d, int M, int N,
                              int K, const float *qa, const float *qb, hipStream_t st) {
    const size_t lds = (size_t)NS * (TL::BM + TL::BN) * LDSROW;
    static const hipError_t attr = hipFuncSetAttribute((const void *)gemm_pipe_kernel<MODE, TL, NS>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    UMI_TRY(hip_check(attr, "hipFuncSetAttribute(gemm pipe LDS)"));
    const int tm = ceil_div(M, TL::BM), tn = ceil_div(N, TL::BN);
    hipLaunchKernelGGL((gemm_pipe_kernel<MODE, TL, NS>), dim3(tm * tn), dim3(TL::NT), lds, st, ah, lda, bh, ldb, d, M,
                       N, K, tm, tn, qa, qb);
    return launch_check("gemm_pipe_kernel");
}

template <int MODE>
umi_status gemm_lds_launch(const uint8_t *ah, const uint8_t *al, long lda, const uint8_t *bh, const uint8_t *bl,
                           long ldb, DenseDst d, int M, int N, int K, const float *qa, const float *qb,
                           hipStream_t st) {
    // the pipelined kernel's stores address C with 32-bit offsets
    const bool c32 = (long)d.M * d.ldc * 4 < (1L << 31) && (long)d.ldc * 4 * 256 < (1L << 31);
    if constexpr (Mfma<MODE>::NOP == 2) {
        if (c32 && gemm_lds_form == 1)
            return gemm_pipe_launch_t<MODE, LdsTile8, 4>(ah, lda, bh, ldb, d, M, N, K, qa, qb, st);
        if (c32 && gemm_lds_form == 5)
            return gemm_pipe_launch_t<MODE, LdsTile4, 4>(ah, lda, bh, ldb, d, M, N, K, qa, qb, st);
    }
    if (gemm_lds_form == 4 || gemm_lds_form == 5)
        return gemm_lds_launch_t<MODE, LdsTile4>(ah, al, lda, bh, bl, ldb, d, M, N, K, qa, qb, st);
    return gemm_lds_launch_t<MODE, LdsTile8>(ah, al, lda, bh, bl, ldb, d, M, N, K, qa, qb, st);
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

// Direct 3x3 convolution, stride 1, padding 1, C = K = 64, bf16 or fp8 (the
// implicit GEMM's operands, MFMA order and scaling, so the same bits as
// gemm_kernel<.., MODE, DenseSrc, ConvSrc, ConvDst>).  A workgroup (14 waves)
// owns DC_PX = 224 consecutive output pixels of one image (the batch-16 56x56
// layer is 224 workgroups: one per CU, a single round).  It stages the filter
// (64 rows of 576 k, KRSC order) and the input rows the pixels touch plus the
// padding (at most dc_rows(W) rows of W + 2 pixels of 64 channels) in LDS
// once, each row or pixel followed by 16 spare bytes so that the operand
// reads of consecutive lanes spread over the banks.  Wave w computes pixels
// 32 (w % 7) .. + 31 for output channels 32 (w / 7) .. + 31: 36 bf16 MFMAs
// (k = 16 q .. + 15) or 9 fp8 ones (k = 64 q .. + 63) on LDS operands, no
// barrier in the K walk.
constexpr int DC_GROUPS = 7, DC_PX = 32 * DC_GROUPS, DC_THREADS = 128 * DC_GROUPS;
bool direct_conv_enabled = true;  // umi_conv_set_direct
__host__ __device__ constexpr int dc_rows(int W) { return (DC_PX + W - 1) / W + 3; }
template <bool FP8, bool SPLIT = false>
struct DcGeom {
    static constexpr int ESZ = FP8 ? 1 : 2, PIXB = 64 * ESZ, PXB = PIXB + 16, WROW = 576 * ESZ + 16;
    static constexpr int WBYTES = 64 * WROW, WCHUNKS = 64 * 576 * ESZ / 16, LANEB = FP8 ? 32 : 16;
    static constexpr int NQ = FP8 ? 9 : 36, PARTS = SPLIT ? 2 : 1;  // SPLIT: hi and lo operands
};
template <bool FP8, bool SPLIT>
__host__ inline size_t dc_lds(int W) {
    using G = DcGeom<FP8, SPLIT>;
    return G::PARTS * ((size_t)G::WBYTES + (size_t)dc_rows(W) * (W + 2) * G::PXB);
}

template <bool FP8, bool SPLIT>
__global__ __launch_bounds__(DC_THREADS) void conv3x3_direct_kernel(const uint8_t *__restrict__ in_hi,
                                                                    const uint8_t *__restrict__ in_lo,
                                                                    const uint8_t *__restrict__ wt_hi,
                                                                    const uint8_t *__restrict__ wt_lo, int N, int H,
                                                                    int W, const float *qw, const float *qi,
                                                                    float *__restrict__ out) {
    static_assert(FP8 || !SPLIT, "the split form is staged for fp8 only");
    using G = DcGeom<FP8, SPLIT>;
    extern __shared__ __attribute__((aligned(16))) uint8_t dsm[];
    // LDS: filter hi [, lo], then the input tile hi [, lo]
    uint8_t *wsm = dsm, *xsm = dsm + G::PARTS * G::WBYTES;
    const int HW = H * W, tiles = (HW + DC_PX - 1) / DC_PX;
    const int n = blockIdx.x / tiles, p0 = (blockIdx.x - n * tiles) * DC_PX;
    const int ra = p0 / W, rb = min(p0 + DC_PX - 1, HW - 1) / W;
    const int y0 = ra - 1, nrows = rb - ra + 3, W2 = W + 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 31, h = lane >> 5;
    const int g = wv % DC_GROUPS, hc = wv / DC_GROUPS;
    const int xbytes = nrows * W2 * G::PXB;  // one input part in LDS
#pragma unroll
    for (int part = 0; part < G::PARTS; ++part) {
    const uint8_t *wt = part ? wt_lo : wt_hi, *in = part ? in_lo : in_hi;
    uint8_t *wsp = wsm + part * G::WBYTES, *xsp = xsm + part * xbytes;
    // filter: 64 rows of 576 ESZ bytes in 16-byte chunks; a thread's loads are
    // issued first and written to LDS after the input's first loads are issued
    // too, so that both fetches are in flight together
    constexpr int U = (G::WCHUNKS + DC_THREADS - 1) / DC_THREADS, RC = 576 * G::ESZ / 16;
    i32x4 fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = min(u * DC_THREADS + tid, G::WCHUNKS - 1);  // past the end: the last chunk again
        fv[u] = *(const i32x4 *)(wt + (size_t)e * 16);
    }
    auto put_filter = [&]() {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = min(u * DC_THREADS + tid, G::WCHUNKS - 1), row = e / RC;
            *(i32x4 *)(wsp + row * G::WROW + 16 * (e - row * RC)) = fv[u];
        }
    };
    // input: chunk e = (slot, 16-byte part c) of the nrows x W2 pixel slots
    constexpr int PC = G::PIXB / 16;
    const uint8_t *img = in + (size_t)n * HW * G::PIXB;
    const int nchunks = nrows * W2 * PC;
    for (int e0 = 0; e0 < nchunks; e0 += 4 * DC_THREADS) {
        i32x4 v[4];
        int slot[4], cpart[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * DC_THREADS + tid;
            slot[u] = e / PC;
            cpart[u] = e - slot[u] * PC;
            const int iy = y0 + slot[u] / W2, ix = slot[u] % W2 - 1;
            const bool in_img = e < nchunks && iy >= 0 && iy < H && ix >= 0 && ix < W;
            v[u] = in_img ? *(const i32x4 *)(img + ((size_t)iy * W + ix) * G::PIXB + 16 * cpart[u])
                          : i32x4{0, 0, 0, 0};
        }
        if (e0 == 0) put_filter();
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * DC_THREADS + tid < nchunks) *(i32x4 *)(xsp + slot[u] * G::PXB + 16 * cpart[u]) = v[u];
    }
    }
    __syncthreads();
    const int p = p0 + 32 * g + lr;
    const bool valid = p < HW;
    const int pp = valid ? p : p0;
    const int oy = pp / W, ox = pp - oy * W;
    const uint8_t *xb = xsm + ((oy - y0) * W2 + ox + 1) * G::PXB + G::LANEB * h;
    const uint8_t *wb = wsm + (32 * hc + lr) * G::WROW + G::LANEB * h;
    f32x16 acc = {};
#pragma unroll
    for (int q = 0; q < G::NQ; ++q) {
        if constexpr (FP8) {
            const int dy = q / 3 - 1, dx = q % 3 - 1;
            const uint8_t *aq = wb + 64 * q, *bq = xb + (dy * W2 + dx) * G::PXB;
            const auto ld32 = [](const uint8_t *p) {
                const i32x4 u0 = *(const i32x4 *)p, u1 = *(const i32x4 *)(p + 16);
                return i32x8{u0[0], u0[1], u0[2], u0[3], u1[0], u1[1], u1[2], u1[3]};
            };
            const i32x8 av = ld32(aq), bv = ld32(bq);
            acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, 127, 0, 127);
            if constexpr (SPLIT) {
                // the implicit GEMM's residual terms: A lo * B hi, then A hi * B lo, each
                // with the 2^-4 weight as the lo operand's E8M0 scale (127 - 4)
                const i32x8 al = ld32(aq + G::WBYTES), bl = ld32(bq + xbytes);
                acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(al, bv, acc, 0, 0, 0, 123, 0, 127);
                acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bl, acc, 0, 0, 0, 127, 0, 123);
            }
        } else {
            const int t = q >> 2, dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 a = *(const bf16x8 *)(wb + 32 * q);
            const bf16x8 b = *(const bf16x8 *)(xb + (dy * W2 + dx) * G::PXB + 32 * (q & 3));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    // The MFMAs (and the per-lane scale operands of the fp8 form) must run
    // with every lane active: an MFMA reads all 64 lanes' operands whatever
    // EXEC says.  The empty asm consumes acc here, so the compiler cannot
    // sink the chain into the branch below.
    asm volatile("" : "+v"(acc));
    if (!valid) return;
    // the implicit GEMM's scaling: alpha = 1 / (qa qb) with A = filter, B = input
    const float alpha = FP8 ? (1.0f / *qw) * (1.0f / *qi) : 1.0f;
    // nontemporal stores: the output streams out, this kernel never reads it
    float *o = out + (size_t)n * 64 * HW + p;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = 32 * hc + 8 * (r >> 2) + 4 * h + (r & 3);
        __builtin_nontemporal_store(acc[r] * alpha, &o[(size_t)m * HW]);
    }
}

// The direct 3x3 conv in bf16x3 (the implicit GEMM's three products per
// k chunk: Ah Bh, Al Bh, Ah Bl, in that order, so the same bits).  Both input
// parts of the tile's rows (as conv3x3_direct_kernel, 144 B per pixel) stay
// in LDS; the filter's hi and lo parts do not fit beside them, so they pass
// through a ring of two one-tap LDS buffers (64 output-channel rows of 64
// input channels, 128 B + 16 spare per row and part): while the waves run
// tap t from one buffer, the other holds tap t + 1 and the thread's
// registers hold tap t + 2's loads, written after the next barrier.  One
// barrier per tap.  Waves as conv3x3_direct_kernel: 7 pixel groups x 2
// output-channel halves, 12 MFMAs per tap.
constexpr int DX_ROWB = 128 + 16, DX_PART = 64 * DX_ROWB, DX_TAP = 2 * DX_PART;
constexpr int DX_TCH = 64 * 128 / 16;  // 16-byte chunks of one tap part (64 rows x 8), padding not loaded
__host__ inline size_t dx_lds(int W) { return 2 * (size_t)DX_TAP + 2 * (size_t)dc_rows(W) * (W + 2) * 144; }

__global__ __launch_bounds__(DC_THREADS) void conv3x3_direct_x3_kernel(const uint8_t *__restrict__ in_hi,
                                                                       const uint8_t *__restrict__ in_lo,
                                                                       const uint8_t *__restrict__ wt_hi,
                                                                       const uint8_t *__restrict__ wt_lo, int N, int H,
                                                                       int W, float *__restrict__ out) {
    constexpr int PXB = 144, PIXB = 128;
    extern __shared__ __attribute__((aligned(16))) uint8_t dsm[];
    uint8_t *ring = dsm, *xsm = dsm + 2 * DX_TAP;  // ring[2] of (hi, lo) taps, then input hi, input lo
    const int HW = H * W, tiles = (HW + DC_PX - 1) / DC_PX;
    const int n = blockIdx.x / tiles, p0 = (blockIdx.x - n * tiles) * DC_PX;
    const int ra = p0 / W, rb = min(p0 + DC_PX - 1, HW - 1) / W;
    const int y0 = ra - 1, nrows = rb - ra + 3, W2 = W + 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 31, h = lane >> 5;
    const int g = wv % DC_GROUPS, hc = wv / DC_GROUPS;
    const int xbytes = nrows * W2 * PXB;
    // one tap's filter chunks: chunk e < 2 DX_TCH = (part, row, 16-byte piece)
    constexpr int TU = (2 * DX_TCH + DC_THREADS - 1) / DC_THREADS;
    const auto tap_load = [&](int t, i32x4 (&v)[TU]) {
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const int e = min(u * DC_THREADS + tid, 2 * DX_TCH - 1);  // past the end: the last chunk again
            const int part = e / DX_TCH, r = (e % DX_TCH) / 8, c = e % 8;
            v[u] = *(const i32x4 *)((part ? wt_lo : wt_hi) + ((size_t)r * 576 + t * 64) * 2 + 16 * c);
        }
    };
    const auto tap_put = [&](uint8_t *buf, const i32x4 (&v)[TU]) {
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const int e = min(u * DC_THREADS + tid, 2 * DX_TCH - 1);
            const int part = e / DX_TCH, r = (e % DX_TCH) / 8, c = e % 8;
            *(i32x4 *)(buf + part * DX_PART + r * DX_ROWB + 16 * c) = v[u];
        }
    };
    i32x4 t0[TU], t1[TU], tn[TU];
    tap_load(0, t0);
    tap_load(1, t1);
    // input: chunk e = (part, slot, 16-byte piece) of the nrows x W2 pixel slots
    constexpr int PC = PIXB / 16;
    const int nchunks = nrows * W2 * PC;
    for (int part = 0; part < 2; ++part) {
        const uint8_t *img = (part ? in_lo : in_hi) + (size_t)n * HW * PIXB;
        uint8_t *xsp = xsm + part * xbytes;
        for (int e0 = 0; e0 < nchunks; e0 += 4 * DC_THREADS) {
            i32x4 v[4];
            int slot[4], cpart[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * DC_THREADS + tid;
                slot[u] = e / PC;
                cpart[u] = e - slot[u] * PC;
                const int iy = y0 + slot[u] / W2, ix = slot[u] % W2 - 1;
                const bool in_img = e < nchunks && iy >= 0 && iy < H && ix >= 0 && ix < W;
                v[u] = in_img ? *(const i32x4 *)(img + ((size_t)iy * W + ix) * PIXB + 16 * cpart[u])
                              : i32x4{0, 0, 0, 0};
            }
            if (part == 0 && e0 == 0) {
                tap_put(ring, t0);
                tap_put(ring + DX_TAP, t1);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (e0 + u * DC_THREADS + tid < nchunks) *(i32x4 *)(xsp + slot[u] * PXB + 16 * cpart[u]) = v[u];
        }
    }
    tap_load(2, tn);
    __syncthreads();
    const int p = p0 + 32 * g + lr;
    const bool valid = p < HW;
    const int pp = valid ? p : p0;
    const int oy = pp / W, ox = pp - oy * W;
    const uint8_t *xh = xsm + ((oy - y0) * W2 + ox + 1) * PXB + 16 * h, *xl = xh + xbytes;
    f32x16 acc = {};
    for (int t = 0; t < 9; ++t) {
        const uint8_t *wh = ring + (t & 1) * DX_TAP + (32 * hc + lr) * DX_ROWB + 16 * h, *wl = wh + DX_PART;
        const int dy = t / 3 - 1, dx = t % 3 - 1, off = (dy * W2 + dx) * PXB;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x8 ah = *(const bf16x8 *)(wh + 32 * q), al = *(const bf16x8 *)(wl + 32 * q);
            const bf16x8 bh = *(const bf16x8 *)(xh + off + 32 * q), bl = *(const bf16x8 *)(xl + off + 32 * q);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        }
        __syncthreads();  // every wave is done with tap t's buffer
        if (t + 2 < 9) {
            tap_put(ring + (t & 1) * DX_TAP, tn);  // tap t + 2 into the buffer tap t used
            if (t + 3 < 9) tap_load(t + 3, tn);
        }
        // the buffer of tap t + 1 was written before the barrier above
    }
    asm volatile("" : "+v"(acc));
    if (!valid) return;
    float *o = out + (size_t)n * 64 * HW + p;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = 32 * hc + 8 * (r >> 2) + 4 * h + (r & 3);
        __builtin_nontemporal_store(acc[r], &o[(size_t)m * HW]);
    }
}

}  // namespace

extern "C" {

umi_status umi_gemm_set_lds_direct(int enable) {
    if (enable != 0 && enable != 1 && enable != 2 && enable != 4 && enable != 5)
        return set_error(UMI_ERR_INVALID, "umi_gemm_set_lds_direct: %d (0, 1, 2, 4 or 5)", enable);
    gemm_lds_direct = enable != 0;
    if (enable != 0) gemm_lds_form = enable;
    return UMI_OK;
}

umi_status umi_conv_set_direct(int enable) {
    direct_conv_enabled = enable != 0;
    return UMI_OK;
}

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
    if (gemm_lds_direct && M > 64 && (long)ceil_div(M, 256) * ceil_div(N, 256) >= 256 &&
        ((reinterpret_cast<uintptr_t>(a_hi) | reinterpret_cast<uintptr_t>(b_hi) |
          (split(mode) ? reinterpret_cast<uintptr_t>(a_lo) | reinterpret_cast<uintptr_t>(b_lo) : 0)) & 15) == 0) {
        const auto *pah = (const uint8_t *)a_hi, *pal = (const uint8_t *)a_lo;
        const auto *pbh = (const uint8_t *)b_hi, *pbl = (const uint8_t *)b_lo;
        const long la = (long)lda * e, lb = (long)ldb * e;
        switch (mode) {
            case GM_F32: return gemm_lds_launch<GM_F32>(pah, pal, la, pbh, pbl, lb, d, M, N, K, nullptr, nullptr, st);
            case GM_BF16: return gemm_lds_launch<GM_BF16>(pah, pal, la, pbh, pbl, lb, d, M, N, K, nullptr, nullptr, st);
            case GM_BF16X3:
                return gemm_lds_launch<GM_BF16X3>(pah, pal, la, pbh, pbl, lb, d, M, N, K, nullptr, nullptr, st);
            case GM_FP8: return gemm_lds_launch<GM_FP8>(pah, pal, la, pbh, pbl, lb, d, M, N, K, qa, qb, st);
            default: return gemm_lds_launch<GM_FP8X3>(pah, pal, la, pbh, pbl, lb, d, M, N, K, qa, qb, st);
        }
    }
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
    // the stage prep divides k < R*S*C by C and rs < R*S by S by multiply-high (exact while x * d < 2^32)
    if ((uint64_t)KG * (uint64_t)C >= (1ull << 32))
        return set_error(UMI_ERR_UNSUPPORTED, "umi_conv_fprop: R*S*C*C = %llu >= 2^32", (unsigned long long)KG * C);
    const ConvSrc bh{(const uint8_t *)in_hi, N, H, W, C, OH, OW, R, S, stride, pad, 1, recip32(C), recip32(S)};
    const ConvSrc bl{(const uint8_t *)in_lo, N, H, W, C, OH, OW, R, S, stride, pad, 1, recip32(C), recip32(S)};
    const ConvDst d{out, K, OH * OW, N * OH * OW};
    const int NP = N * OH * OW;
    hipStream_t st = as_stream(s);
    if ((mode == GM_BF16 || mode == GM_FP8 || mode == GM_FP8X3) && C == 64 && K == 64 && R == 3 && S == 3 &&
        stride == 1 && pad == 1 && W >= 8 && W <= 64 &&
        ((reinterpret_cast<uintptr_t>(in_hi) | reinterpret_cast<uintptr_t>(w_hi) |
          (mode == GM_FP8X3 ? reinterpret_cast<uintptr_t>(in_lo) | reinterpret_cast<uintptr_t>(w_lo) : 0)) & 15) == 0 &&
        direct_conv_enabled) {
        // the direct form: same operands, MFMA order and scaling as the implicit GEMM below
        const int form = mode == GM_BF16 ? 0 : mode == GM_FP8 ? 1 : 2;
        const auto kern = form == 0 ? conv3x3_direct_kernel<false, false>
                                    : form == 1 ? conv3x3_direct_kernel<true, false> : conv3x3_direct_kernel<true, true>;
        const auto lds_for = [form](int w) {
            return form == 0 ? dc_lds<false, false>(w) : form == 1 ? dc_lds<true, false>(w) : dc_lds<true, true>(w);
        };
        static bool attr_set[3] = {false, false, false};
        if (!attr_set[form]) {
            size_t most = 0;  // the largest LDS over the widths this form takes
            for (int w = 8; w <= 64; ++w) most = std::max(most, lds_for(w));
            UMI_TRY(hip_check(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)most),
                              "hipFuncSetAttribute(direct conv LDS)"));
            attr_set[form] = true;
        }
        hipLaunchKernelGGL(kern, dim3(N * ceil_div(OH * OW, DC_PX)), dim3(DC_THREADS), lds_for(W), st,
                           (const uint8_t *)in_hi, (const uint8_t *)in_lo, (const uint8_t *)w_hi,
                           (const uint8_t *)w_lo, N, H, W, qw, qi, out);
        return launch_check("conv3x3_direct_kernel");
    }
    if (mode == GM_BF16X3 && C == 64 && K == 64 && R == 3 && S == 3 && stride == 1 && pad == 1 && W >= 8 &&
        dx_lds(W) <= 160 * 1024 &&
        ((reinterpret_cast<uintptr_t>(in_hi) | reinterpret_cast<uintptr_t>(w_hi) | reinterpret_cast<uintptr_t>(in_lo) |
          reinterpret_cast<uintptr_t>(w_lo)) & 15) == 0 &&
        direct_conv_enabled) {
        // the direct form with the filter streamed one tap at a time (bit-identical to the implicit GEMM)
        static bool attr_x3 = false;
        if (!attr_x3) {
            UMI_TRY(hip_check(hipFuncSetAttribute((const void *)conv3x3_direct_x3_kernel,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024),
                              "hipFuncSetAttribute(direct bf16x3 conv LDS)"));
            attr_x3 = true;
        }
        hipLaunchKernelGGL(conv3x3_direct_x3_kernel, dim3(N * ceil_div(OH * OW, DC_PX)), dim3(DC_THREADS), dx_lds(W),
                           st, (const uint8_t *)in_hi, (const uint8_t *)in_lo, (const uint8_t *)w_hi,
                           (const uint8_t *)w_lo, N, H, W, out);
        return launch_check("conv3x3_direct_x3_kernel");
    }
    switch (mode) {
        case GM_F32: return gemm_launch<GM_F32>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_BF16: return gemm_launch<GM_BF16>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_BF16X3: return gemm_launch<GM_BF16X3>(ah, al, bh, bl, d, K, NP, KG, nullptr, nullptr, st);
        case GM_FP8: return gemm_launch<GM_FP8>(ah, al, bh, bl, d, K, NP, KG, qw, qi, st);
        default: return gemm_launch<GM_FP8X3>(ah, al, bh, bl, d, K, NP, KG, qw, qi, st);
    }
}

}  // extern "C"
