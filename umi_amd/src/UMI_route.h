// UMI_route.h — UMI::Execute calls that a specialised kernel of the library
// computes exactly are sent to it (no reference counterpart: the reference
// runs every program through its one generic kernel, src/UMI_gpu.h:189-358).
//
// A strategy's body is opaque to a library, so a call is routed only when
// both hold:
//  * its descriptors have the exact shape the kernel computes (checked on
//    the affine maps the engine would use), and
//  * the strategy, run on the host over random probe operands with the
//    call's own constants, returns what the kernel's formula returns for the
//    same operands, bit for bit.
// Recognised: the 3x3 filter of main/filter3x3.cu:10-17,36-58
// (out = mul * (init + sum of the 9 taps)) and the weighted 3x3 filter of
// main/wfilter3x3.cu:8-15,39-66 (out = init + sum of tap * weight), int32 and
// float, PAD or clamp boundary.  The taps are summed in the engine's
// block-major order, which is raster order when a whole tap row fits one
// accumulation block; the float weighted form is then the fused
// multiply-add chain of the device compilation (HIP contracts a * b + c),
// which the host probes check on exactly representable operands.  Anything
// else runs on the generic engine.  UMI_ROUTE=0 in the environment, or
// UMI::route::enable(false), turns routing off.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <type_traits>

namespace UMI {
namespace route {

inline bool &enabled_flag() {
    static bool on = [] {
        const char *e = std::getenv("UMI_ROUTE");
        return !(e && e[0] == '0');
    }();
    return on;
}
inline void enable(bool on) { enabled_flag() = on; }
inline bool enabled() { return enabled_flag(); }
// Calls routed so far in this process (tests and benchmarks read it).
inline long &routed_count() {
    static long n = 0;
    return n;
}
inline long routed() { return routed_count(); }

// Map of an image-shaped array: physical (y, x) in dims 2, 3 of an H x W
// row-major array = (p2 + oy + taps * a2, p3 + ox + taps * a3); dims 0, 1
// unused (size <= 1, no coefficient).
inline bool image_map(const engine::ArrayMap &m, int H, int W, int oy, int ox, bool taps) {
    for (int k = 0; k < DIM; ++k)
        for (int d = 0; d < DIM; ++d) {
            const int diag = (k == d && k >= 2) ? 1 : 0;
            if (m.pc[k][d] != diag || m.ac[k][d] != (taps ? diag : 0)) return false;
        }
    return m.base[0] == 0 && m.base[1] == 0 && m.base[2] == oy && m.base[3] == ox && m.size[0] <= 1 &&
           m.size[1] <= 1 && m.size[2] == H && m.size[3] == W;
}
// Map of the 3x3 weights: (a2, a3) of a contiguous 3 x 3 array, no p term.
inline bool weights_map(const engine::ArrayMap &m) {
    for (int k = 0; k < DIM; ++k)
        for (int d = 0; d < DIM; ++d) {
            const int diag = (k == d && k >= 2) ? 1 : 0;
            if (m.pc[k][d] != 0 || m.ac[k][d] != diag) return false;
        }
    for (int k = 0; k < DIM; ++k)
        if (m.base[k] != 0) return false;
    return m.size[0] <= 1 && m.size[1] <= 1 && m.size[2] == 3 && m.size[3] == 3;
}

// Runs the strategy on the host: Pre, nine Loop steps (tap k with operand
// x[k] and, for two Loop inputs, weight w[k]), Post; returns its output.
template <class Umi, class E, class C>
E run_host(const C &c, const E (&x)[9], const E (&w)[9]) {
    Umi u{};
    simple_array<E, Umi::N_PRE_IN * MAX_PTR> pin;
    simple_array<E, Umi::N_PRE_OUT * MAX_PTR> pout;
    u.Pre(c, pin, pout);
    for (int k = 0; k < 9; ++k) {
        simple_array<E, Umi::N_LOOP_IN * MAX_PTR> in;
        simple_array<E, Umi::N_LOOP_OUT * MAX_PTR> out;
        in[0] = x[k];
        if constexpr (Umi::N_LOOP_IN > 1) in[MAX_PTR] = w[k];
        u.Loop(c, in, out);
    }
    simple_array<E, Umi::N_POST_IN * MAX_PTR> qin;
    simple_array<E, Umi::N_POST_OUT * MAX_PTR> qout;
    u.Post(c, qin, qout);
    return qout[0];
}

template <class E>
bool same_bits(E a, E b) {
    return std::memcmp(&a, &b, sizeof(E)) == 0;
}

// The kernel formulas, evaluated on the host in the kernels' order.
template <class E>
E box_formula(E init, E mul, const E (&x)[9]) {
    if (std::is_integral<E>::value) {
        uint32_t s = (uint32_t)init;
        for (int k = 0; k < 9; ++k) s += (uint32_t)x[k];
        return (E)(int32_t)(s * (uint32_t)mul);
    }
    E s = init;
    for (int k = 0; k < 9; ++k) s = s + x[k];
    return s * mul;
}
template <class E>
E weighted_formula(E init, const E (&x)[9], const E (&w)[9]) {
    if (std::is_integral<E>::value) {
        uint32_t s = (uint32_t)init;
        for (int k = 0; k < 9; ++k) s += (uint32_t)x[k] * (uint32_t)w[k];
        return (E)(int32_t)s;
    }
    E s = init;  // rounds as the host run of the strategy does (both unfused here)
    for (int k = 0; k < 9; ++k) s = s + x[k] * w[k];
    return s;
}

// Probe operands: integers in [-1000, 1000] (the host run of the strategy
// must not overflow int); for float, kind 0 small integers (every sum
// exact), kind 1 general values.
template <class E>
void probe(std::mt19937 &g, E (&v)[9], int kind) {
    for (int k = 0; k < 9; ++k) {
        if (std::is_integral<E>::value || kind == 0)
            v[k] = (E)((int)(g() % 2001) - 1000);
        else
            v[k] = (E)(((double)(g() % 2000001) - 1000000.0) / 1024.0 * (1 + (int)(g() % 3)));
    }
}

template <class Umi, uint64_t PtrNum, class E, class C, size_t NI, size_t NO, bool WEIGHTED>
bool try_stencil_typed(umi_stream stream, const NDTuple<ParallelDim> &pd, const NDTuple<AccumulationDim> &ad,
                       const std::array<UnrolledI<E>, NI> &I, const std::array<UnrolledO<E>, NO> &O, const C &c,
                       umi_status &status) {
    if (Umi::N_PRE_IN != 0 || Umi::N_POST_IN != 0 || (Umi::N_LOOP_IN != 1 && Umi::N_LOOP_IN != 2) ||
        Umi::N_PRE_OUT != 0 || Umi::N_LOOP_OUT != 0 || Umi::N_POST_OUT != 1 || NI != (size_t)Umi::N_LOOP_IN || NO != 1)
        return false;
    if (engine::planes<PtrNum>(0) != 1 || (WEIGHTED && engine::planes<PtrNum>(1) != 1) ||
        engine::planes<PtrNum>((int)NI) != 1)
        return false;
    if (sizeof(C) != (WEIGHTED ? 1 : 2) * sizeof(E)) return false;
    const int H = pd[2].total_size, W = pd[3].total_size;
    if (pd[0].total_size != 1 || pd[1].total_size != 1 || H < 1 || W < 1) return false;
    if (ad[0].total_size != 1 || ad[1].total_size != 1 || ad[2].total_size != 3 || ad[3].total_size != 3 ||
        ad[3].block_size < 3)
        return false;
    const engine::ArrayMap mi = engine::make_map(I[0].p_view_dims, I[0].a_view_dims, I[0].physical_dims);
    const engine::ArrayMap mo = engine::make_map(O[0].p_view_dims, O[0].a_view_dims, O[0].physical_dims);
    if (!image_map(mi, H, W, -1, -1, true) || !image_map(mo, H, W, 0, 0, false)) return false;
    const int pad2 = (mi.pad_mask >> 2) & 1, pad3 = (mi.pad_mask >> 3) & 1;
    if (pad2 != pad3) return false;
    if (WEIGHTED) {
        const engine::ArrayMap mw = engine::make_map(I[NI - 1].p_view_dims, I[NI - 1].a_view_dims,
                                                     I[NI - 1].physical_dims);
        if (!weights_map(mw)) return false;
    }
    // the call's constants, read as the kernel's parameters, then checked
    E prm[2] = {};
    std::memcpy(prm, &c, sizeof(C));
    if constexpr (std::is_integral<E>::value) {  // keep the host runs free of int overflow
        const int64_t init = prm[0], mul = WEIGHTED ? 1 : prm[1];
        const int64_t reach = (init < 0 ? -init : init) + (WEIGHTED ? 9000000 : 9000);
        if ((mul < 0 ? -mul : mul) * reach >= ((int64_t)1 << 31) || reach >= ((int64_t)1 << 31)) return false;
    }
    // the probes' verdict for the last constants seen by this thread
    thread_local unsigned char last_c[sizeof(C)];
    thread_local int last_ok = -1;
    if (last_ok < 0 || std::memcmp(last_c, &c, sizeof(C)) != 0) {
        std::mt19937 g(0x5eed);
        int ok = 1;
        for (int t = 0; t < 64 && ok; ++t) {
            E x[9], w[9];
            probe(g, x, t & 1);
            probe(g, w, WEIGHTED && std::is_floating_point<E>::value ? 0 : (t & 1));
            const E got = run_host<Umi, E, C>(c, x, w);
            const E want = WEIGHTED ? weighted_formula(prm[0], x, w) : box_formula(prm[0], prm[1], x);
            ok = same_bits(got, want);
        }
        std::memcpy(last_c, &c, sizeof(C));
        last_ok = ok;
    }
    if (!last_ok) return false;
    const int mode = pad2 ? 0 : 1;  // PAD, else clamp (the reference's WRAP)
    const E pad = I[0].padding[0];
    if constexpr (std::is_floating_point<E>::value) {
        const float *in = reinterpret_cast<const float *>(I[0].ptr[0]);
        float *out = reinterpret_cast<float *>(O[0].ptr[0]);
        if constexpr (WEIGHTED)
            status = umi_wfilter3x3_f32(in, H, W, W, reinterpret_cast<const float *>(I[NI - 1].ptr[0]), prm[0], pad,
                                        mode, out, W, stream);
        else
            status = umi_filter3x3_f32(in, H, W, W, prm[0], prm[1], pad, mode, out, W, stream);
    } else {
        const int32_t *in = reinterpret_cast<const int32_t *>(I[0].ptr[0]);
        int32_t *out = reinterpret_cast<int32_t *>(O[0].ptr[0]);
        if constexpr (WEIGHTED)
            status = umi_wfilter3x3_i32(in, H, W, W, reinterpret_cast<const int32_t *>(I[NI - 1].ptr[0]),
                                        (int32_t)prm[0], (int32_t)pad, mode, out, W, stream);
        else
            status = umi_filter3x3_i32(in, H, W, W, (int32_t)prm[0], (int32_t)prm[1], (int32_t)pad, mode, out, W,
                                       stream);
    }
    if (status == UMI_OK) ++routed_count();
    return true;
}

// 3x3 filter and weighted filter: the descriptor shape of the reference
// programs (any parallel mapping, whole tap rows per accumulation block).
template <class Umi, uint64_t PtrNum, class E, class C, size_t NI, size_t NO>
bool try_stencil(umi_stream stream, const NDTuple<ParallelDim> &pd, const NDTuple<AccumulationDim> &ad,
                 const std::array<UnrolledI<E>, NI> &I, const std::array<UnrolledO<E>, NO> &O, const C &c,
                 umi_status &status) {
    constexpr bool WEIGHTED = Umi::N_LOOP_IN == 2;
    if constexpr (!std::is_same<E, int32_t>::value && !std::is_same<E, float>::value) {
        return false;
    } else {
        return try_stencil_typed<Umi, PtrNum, E, C, NI, NO, WEIGHTED>(stream, pd, ad, I, O, c, status);
    }
}

}  // namespace route
}  // namespace UMI
