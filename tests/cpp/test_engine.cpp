// GPU test of the generic UMI::Execute<> engine against the CPU interpreter
// (oracle/umi_interp.h).  Each case runs the same strategy struct and the
// same descriptors on the GPU (device pointers, through SyncedMemory) and on
// the CPU (host pointers), then compares every output element: bit-exact for
// integer and add-only float strategies, relative tolerance where the GPU
// contracts a*b+c into one fused multiply-add.
//
// The strategies restate the reference examples' (main/*.cu, johnjohnlin/UMI)
// in this project's code, plus cases the examples do not cover: several
// planes per array (PtrNum), Loop outputs, Pre/Post inputs and outputs,
// negative strides, a partially filled accumulation block, clamp vs pad.
// Build: __graft_entry__.build() (hipcc -x hip).  Run: tests/test_engine_gpu.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <utility>
#include <vector>

#include "SyncedMemory.h"
#include "UMI_gpu.h"
#include "umi_interp.h"

using namespace UMI;
using std::vector;
static const auto PAD = PhysicalDim::PAD;
static const auto WRAP = PhysicalDim::WRAP;
static int g_fail = 0;

template <class T>
static void compare(const char *name, const T *gpu, const T *cpu, size_t n, double rtol) {
    size_t bad = 0, first = n;
    double worst = 0;
    for (size_t i = 0; i < n; ++i) {
        if (rtol == 0) {
            if (!(gpu[i] == cpu[i])) {
                if (!bad) first = i;
                ++bad;
            }
        } else {
            const double d = std::fabs((double)gpu[i] - (double)cpu[i]) / std::fmax(std::fabs((double)cpu[i]), 1e-30);
            worst = std::fmax(worst, d);
            if (d > rtol) {
                if (!bad) first = i;
                ++bad;
            }
        }
    }
    if (bad) {
        ++g_fail;
        std::printf("[engine] %s: FAIL %zu/%zu differ, first at %zu (gpu %g cpu %g)\n", name, bad, n, first,
                    (double)gpu[first], (double)cpu[first]);
    } else {
        std::printf("[engine] %s: ok (%zu values%s)\n", name, n, rtol ? ", within tolerance" : ", bit-exact");
    }
    std::fflush(stdout);
}

#include "strategies.h"

// ---------------------------------------------------------------- helpers
template <class T>
struct Buf {
    MemoryBuffer<T> mb;
    SyncedMemory<T> s;
    vector<T> cpu_out;
    size_t n;
    explicit Buf(size_t n_) : mb(n_), s(mb.CreateSync(n_)), cpu_out(n_), n(n_) {}
    T *host() { return s.get_cpu_wo(); }
};

static std::mt19937 rng(12345);
template <class T>
static void fill_int(T *p, size_t n, int lo, int hi) {
    std::uniform_int_distribution<int> d(lo, hi);
    for (size_t i = 0; i < n; ++i) p[i] = (T)d(rng);
}
static void fill_float(float *p, size_t n) {
    std::uniform_real_distribution<float> d(0.f, 1.f);
    for (size_t i = 0; i < n; ++i) p[i] = d(rng);
}

// ---------------------------------------------------------------- cases
static void case_box(int h, int w, int mode) {
    Buf<int> in(w * h), out(w * h);
    fill_int(in.host(), in.n, -1000, 1000);
    const auto M = mode ? WRAP : PAD;
    auto run = [&](const int *ip, int *op, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, 3}}};
        const std::array<UnrolledI<int>, 1> I{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}},
                                                {{{1, M}, {1, M}, {h, M}, {w, M}}}, DEFAULT_SHM, {ip}, {999}}}};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {op}}}};
        if (gpu)
            Execute<BoxFilter, 0x11, 0x1111, 0x1122>(pd, ad, I, O, {-5, 2});
        else
            umi_oracle::Execute<BoxFilter, 0x11>(pd, ad, I, O, {-5, 2});
    };
    run(in.s.get_gpu_ro(), out.s.get_gpu_wo(), true);
    run(in.s.get_cpu_ro(), out.cpu_out.data(), false);
    char name[64];
    std::snprintf(name, sizeof(name), "box3x3 %dx%d mode=%d", h, w, mode);
    compare(name, out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 0);
}

static void case_wfilter(int h, int w) {
    Buf<int> in(w * h), out(w * h), wt(9);
    fill_int(in.host(), in.n, -1000, 1000);
    fill_int(wt.host(), 9, -5, 5);
    auto run = [&](const int *ip, const int *wp, int *op, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, 3}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {ip}, {56}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {3, PAD}, {3, PAD}}}, DEFAULT_SHM, {wp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {op}}}};
        if (gpu)
            Execute<WeightedFilter, 0x111, 0x1122, 0xc0>(pd, ad, I, O, {-5});
        else
            umi_oracle::Execute<WeightedFilter, 0x111>(pd, ad, I, O, {-5});
    };
    run(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), out.s.get_gpu_wo(), true);
    run(in.s.get_cpu_ro(), wt.s.get_cpu_ro(), out.cpu_out.data(), false);
    compare("wfilter3x3 (weights as a second input, Expand 1122)", out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 0);
}

static void case_bilateral(int h, int w, float sigma1, float sigma2) {
    const int ofs = (int)(sigma1 * 3.0f), k = 2 * ofs + 1;
    const float norm1 = -1.0f / (2.0f * sigma1 * sigma1), norm2 = -1.0f / (2.0f * sigma2 * sigma2);
    Buf<float> in(w * h), out(w * h), wt(k);
    fill_float(in.host(), in.n);
    for (int i = 0; i < k; ++i) wt.host()[i] = std::exp((float)((i - ofs) * (i - ofs)) * norm1);
    auto run = [&](const float *ip, const float *wp, float *op, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 2, 1}, {w, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {k, k}, {k, k}}};
        const std::array<UnrolledI<float>, 4> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {}, {}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {ip}, {}},
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -ofs, 1}, {3, -ofs, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {ip}, {0.0f}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {3, 0, 1}, {}}}, {{{1, PAD}, {1, PAD}, {1, PAD}, {k, PAD}}}, DEFAULT_SHM, {wp}, {}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {1, PAD}, {k, PAD}}}, DEFAULT_SHM, {wp}, {}},
        }};
        const std::array<UnrolledO<float>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {op}}}};
        if (gpu)
            Execute<BilateralF, 0x11111>(pd, ad, I, O, {norm2});
        else
            umi_oracle::Execute<BilateralF, 0x11111>(pd, ad, I, O, {norm2});
    };
    run(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), out.s.get_gpu_wo(), true);
    run(in.s.get_cpu_ro(), wt.s.get_cpu_ro(), out.cpu_out.data(), false);
    compare("bilateral (4 inputs, fma on the GPU)", out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 2e-5);
}

static void case_iimg(int h, int w) {
    // two passes: row scan into a transposed temporary, then the same on it;
    // the second pass uses 3-element ParallelDims (warp_shuf 0), as main/iimg.cu:60
    Buf<int> in(w * h), tmp(w * h), out(w * h);
    fill_int(in.host(), in.n, 0, 255);
    vector<int> cpu_tmp(w * h);
    auto run = [&](const int *ip, int *tp, int *op, bool gpu) {
        const NDTuple<ParallelDim> pd1{{{1, 1, 1, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}, {h, 128, 32, 1}}};
        const NDTuple<AccumulationDim> ad1{{{1, 1}, {1, 1}, {1, 1}, {w, 32}}};
        const std::array<UnrolledI<int>, 1> I1{{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {ip}, {}}}};
        const std::array<UnrolledO<int>, 1> O1{{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {tp}}}};
        const NDTuple<ParallelDim> pd2{{{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {w, 128, 32}}};
        const NDTuple<AccumulationDim> ad2{{{1, 1}, {1, 1}, {1, 1}, {h, 32}}};
        const std::array<UnrolledI<int>, 1> I2{{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {}, {tp}, {}}}};
        const std::array<UnrolledO<int>, 1> O2{{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, {op}}}};
        if (gpu) {
            Execute<RunningSum, 0x11>(pd1, ad1, I1, O1);
            Execute<RunningSum, 0x11>(pd2, ad2, I2, O2);
        } else {
            umi_oracle::Execute<RunningSum, 0x11>(pd1, ad1, I1, O1);
            umi_oracle::Execute<RunningSum, 0x11>(pd2, ad2, I2, O2);
        }
    };
    run(in.s.get_gpu_ro(), tmp.s.get_gpu_wo(), out.s.get_gpu_wo(), true);
    run(in.s.get_cpu_ro(), cpu_tmp.data(), out.cpu_out.data(), false);
    compare("integral image, two passes", out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 0);
    // closed check: the interpreter's table equals a direct 2-D prefix sum
    const int *src = in.s.get_cpu_ro();
    vector<int> ref(w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0, rs = 0; x < w; ++x) ref[y * w + x] = (rs += src[y * w + x]) + (y ? ref[(y - 1) * w + x] : 0);
    compare("integral image, interpreter vs direct prefix sum", out.cpu_out.data(), ref.data(), ref.size(), 0);
}

static void case_full_me(int h, int w, int b, int sr) {
    const int sr_ofs = -(sr / 2);
    const size_t n = (size_t)(w / b) * (h / b) * sr * sr;
    Buf<int> frame(w * h), sad(n);
    fill_int(frame.host(), frame.n, 0, 255);
    auto run = [&](const int *fp, int *sp, bool gpu) {
        const NDTuple<ParallelDim> pd{{{h / b, 1, 1, 1}, {w / b, 1, 1, 1}, {sr, 16, 2, 1}, {sr, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {b, b}, {b, b}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{2, 0, b}, {3, 0, b}, {2, sr_ofs, 1}, {3, sr_ofs, 1}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fp}, {}},
            {{{{2, 0, b}, {3, 0, b}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{h / b, WRAP}, {w / b, WRAP}, {sr, WRAP}, {sr, WRAP}}}, {sp}}}};
        if (gpu)
            Execute<Sad, 0x111, 0x2211>(pd, ad, I, O);
        else
            umi_oracle::Execute<Sad, 0x111>(pd, ad, I, O);
    };
    run(frame.s.get_gpu_ro(), sad.s.get_gpu_wo(), true);
    run(frame.s.get_cpu_ro(), sad.cpu_out.data(), false);
    compare("full-search SAD volume (clamp boundary)", sad.s.get_cpu_ro(), sad.cpu_out.data(), n, 0);
}

static void case_nv12_chroma_planes(int h, int w, int sr) {
    // NV12 chroma: U and V interleaved in one [hc][wc][2] array, read as two
    // planes of one input (pointers uv and uv + 1, the channel index left at
    // 0 by the views); SAD volume of the 4x4 chroma blocks over a search of
    // sr x sr offsets, clamp boundary.
    const int hc = h / 2, wc = w / 2, b = 4, ofs = -(sr / 2);
    const size_t n = (size_t)(wc / b) * (hc / b) * sr * sr;
    Buf<int> cur(hc * wc * 2), ref(hc * wc * 2), sad(n);
    fill_int(cur.host(), cur.n, 0, 255);
    fill_int(ref.host(), ref.n, 0, 255);
    auto run = [&](const int *cp, const int *rp, int *sp, bool gpu) {
        const NDTuple<ParallelDim> pd{{{hc / b, 1, 1, 1}, {wc / b, 1, 1, 1}, {sr, 16, 2, 1}, {sr, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {b, b}, {b, b}}};
        const NDTuple<PhysicalDim> phys{{{1, WRAP}, {hc, WRAP}, {wc, WRAP}, {2, WRAP}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{1, 0, b}, {2, 0, b}, {1, ofs, 1}, {2, ofs, 1}}}, {{{}, {}, {1, 0, 1}, {2, 0, 1}}}, phys, DEFAULT_SHM, {rp, rp + 1}, {}},
            {{{{1, 0, b}, {2, 0, b}, {}, {}}}, {{{}, {}, {1, 0, 1}, {2, 0, 1}}}, phys, DEFAULT_SHM, {cp, cp + 1}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{hc / b, WRAP}, {wc / b, WRAP}, {sr, WRAP}, {sr, WRAP}}}, {sp}}}};
        constexpr uint64_t PN = 0x122;  // inputs: 2 planes each; output: 1
        if (gpu)
            Execute<Sad2, PN, 0x2211>(pd, ad, I, O);
        else
            umi_oracle::Execute<Sad2, PN>(pd, ad, I, O);
    };
    run(cur.s.get_gpu_ro(), ref.s.get_gpu_ro(), sad.s.get_gpu_wo(), true);
    run(cur.s.get_cpu_ro(), ref.s.get_cpu_ro(), sad.cpu_out.data(), false);
    compare("NV12 chroma planes: engine vs interpreter", sad.s.get_cpu_ro(), sad.cpu_out.data(), n, 0);
    // direct loop over the interleaved array
    std::vector<int> direct(n);
    const int *c = cur.s.get_cpu_ro(), *r = ref.s.get_cpu_ro();
    size_t o = 0;
    for (int by = 0; by < hc / b; ++by)
        for (int bx = 0; bx < wc / b; ++bx)
            for (int sy = 0; sy < sr; ++sy)
                for (int sx = 0; sx < sr; ++sx) {
                    int sum = 0;
                    for (int j = 0; j < b; ++j)
                        for (int i = 0; i < b; ++i) {
                            const int y = by * b + j, x = bx * b + i;
                            const int yy = std::min(std::max(y + sy + ofs, 0), hc - 1);
                            const int xx = std::min(std::max(x + sx + ofs, 0), wc - 1);
                            for (int ch = 0; ch < 2; ++ch)
                                sum += std::abs(c[(y * wc + x) * 2 + ch] - r[(yy * wc + xx) * 2 + ch]);
                        }
                    direct[o++] = sum;
                }
    compare("NV12 chroma planes: engine vs direct loop", sad.s.get_cpu_ro(), direct.data(), n, 0);
}

static void case_cost_volume(int h, int w, int k, int sr) {
    const int k_ofs = k / 2, sr_ofs = sr / 2;
    Buf<int> frame(w * h), cv((size_t)w * h * sr);
    fill_int(frame.host(), frame.n, 0, 255);
    auto run = [&](const int *fp, int *cp, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {h, 4, 1, 1}, {w, 8, 2, 1}, {sr, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {k, k}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {2, 0, 1}, {3, 0, 1}, {3, -sr_ofs, 1}}}, {{{}, {}, {}, {3, -k_ofs, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fp}, {}},
            {{{{}, {2, 0, 1}, {3, 0, 1}, {}}}, {{{}, {}, {}, {3, -k_ofs, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{1, WRAP}, {h, WRAP}, {w, WRAP}, {sr, WRAP}}}, {cp}}}};
        if (gpu)
            Execute<Sad, 0x111>(pd, ad, I, O);
        else
            umi_oracle::Execute<Sad, 0x111>(pd, ad, I, O);
    };
    run(frame.s.get_gpu_ro(), cv.s.get_gpu_wo(), true);
    run(frame.s.get_cpu_ro(), cv.cpu_out.data(), false);
    compare("1-D cost volume", cv.s.get_cpu_ro(), cv.cpu_out.data(), cv.n, 0);
}

static void case_fprop(int bs, int ich, int och, int ih, int iw, int f, int s) {
    const int ofs = -f / 2, ow = iw / s, oh = ih / s;
    Buf<int> in((size_t)bs * ich * ih * iw), wt((size_t)och * ich * f * f), out((size_t)bs * och * oh * ow);
    fill_int(in.host(), in.n, 0, 10);
    fill_int(wt.host(), wt.n, 0, 10);
    auto run = [&](const int *ip, const int *wp, int *op, bool gpu) {
        const NDTuple<ParallelDim> pd{{{bs, 1, 1, 1}, {och, 4, 1, 1}, {oh, 8, 4, 1}, {ow, 8, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {ich, 6}, {f, f}, {f, f}}};  // ich not a multiple of 6
        const std::array<UnrolledI<int>, 2> I{{
            {{{{0, 0, 1}, {}, {2, 0, s}, {3, 0, s}}}, {{{}, {1, 0, 1}, {2, ofs, 1}, {3, ofs, 1}}}, {{{bs, PAD}, {ich, PAD}, {ih, PAD}, {iw, PAD}}}, DEFAULT_SHM, {ip}, {}},
            {{{{}, {0, 0, 1}, {}, {}}}, {{{}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {{{och, PAD}, {ich, PAD}, {f, PAD}, {f, PAD}}}, DEFAULT_SHM, {wp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{bs, PAD}, {och, PAD}, {oh, PAD}, {ow, PAD}}}, {op}}}};
        if (gpu)
            Execute<Dot>(pd, ad, I, O);
        else
            umi_oracle::Execute<Dot>(pd, ad, I, O);
    };
    run(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), out.s.get_gpu_wo(), true);
    run(in.s.get_cpu_ro(), wt.s.get_cpu_ro(), out.cpu_out.data(), false);
    compare("direct convolution 7x7/2 (partial accumulation block)", out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 0);
}

static void case_gemm(int m, int k, int n) {
    Buf<int> A(m * k), B(k * n), C(m * n);
    std::iota(A.host(), A.host() + A.n, 0);  // main/gemm.cu: iota operands, int32 wrap
    std::iota(B.host(), B.host() + B.n, 0);
    auto run = [&](const int *ap, const int *bp, int *cp, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {m, 16, 2, 1}, {n, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {k, 16}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {m, PAD}, {k, PAD}}}, DEFAULT_SHM, {ap}, {}},
            {{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {k, PAD}, {n, PAD}}}, DEFAULT_SHM, {bp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {m, PAD}, {n, PAD}}}, {cp}}}};
        if (gpu)
            Execute<Dot, 0x111, 0x1124, 0x48>(pd, ad, I, O);
        else
            umi_oracle::Execute<Dot, 0x111>(pd, ad, I, O);
    };
    run(A.s.get_gpu_ro(), B.s.get_gpu_ro(), C.s.get_gpu_wo(), true);
    run(A.s.get_cpu_ro(), B.s.get_cpu_ro(), C.cpu_out.data(), false);
    compare("gemm (Expand 1124, ZeroStrideOpt 0x48)", C.s.get_cpu_ro(), C.cpu_out.data(), C.n, 0);
}

static void case_planes(int h, int w, int depth) {
    // p = (y, x), a = z walked in blocks of 3 (last block partial); the
    // 3-plane input is read with a negative x stride (mirror) and clamp.
    Buf<float> p0(h * w * depth), p1(h * w * depth), p2(h * w * depth), bias(h * w);
    Buf<float> pre(h * w), step((size_t)h * w * depth), post0(h * w), post1(h * w), post2(h * w);
    fill_float(p0.host(), p0.n);
    fill_float(p1.host(), p1.n);
    fill_float(p2.host(), p2.n);
    fill_float(bias.host(), bias.n);
    auto run = [&](const float *a0, const float *a1, const float *a2, const float *bp, float *pr, float *st, float *q0,
                   float *q1, float *q2, bool gpu) {
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 4, 4, 1}, {w, 8, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {depth, 3}}};
        const std::array<UnrolledI<float>, 3> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {bp}, {}},
            {{{{}, {}, {1, 0, 1}, {2, w, -1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, WRAP}, {h, WRAP}, {w, WRAP}, {depth, WRAP}}}, DEFAULT_SHM, {a0, a1, a2}, {}},
            {{{{}, {}, {2, 1, 1}, {3, -1, 1}}}, {}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {bp}, {7.5f}},
        }};
        const std::array<UnrolledO<float>, 4> O{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {pr}},
            {{{{}, {}, {1, 0, 1}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{}, {h, PAD}, {w, PAD}, {depth, PAD}}}, {st}},
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {q0, q1}},
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {q2}},
        }};
        constexpr uint64_t PN = 0x1211131;  // inputs 1,3,1 planes; outputs 1,1,2,1
        if (gpu)
            Execute<Planes, PN>(pd, ad, I, O, {0.5f});
        else
            umi_oracle::Execute<Planes, PN>(pd, ad, I, O, {0.5f});
    };
    vector<float> c_pre(h * w), c_step((size_t)h * w * depth), c0(h * w), c1(h * w), c2(h * w);
    run(p0.s.get_gpu_ro(), p1.s.get_gpu_ro(), p2.s.get_gpu_ro(), bias.s.get_gpu_ro(), pre.s.get_gpu_wo(),
        step.s.get_gpu_wo(), post0.s.get_gpu_wo(), post1.s.get_gpu_wo(), post2.s.get_gpu_wo(), true);
    run(p0.s.get_cpu_ro(), p1.s.get_cpu_ro(), p2.s.get_cpu_ro(), bias.s.get_cpu_ro(), c_pre.data(), c_step.data(),
        c0.data(), c1.data(), c2.data(), false);
    compare("planes: Pre output", pre.s.get_cpu_ro(), c_pre.data(), c_pre.size(), 0);
    compare("planes: Loop output per step", step.s.get_cpu_ro(), c_step.data(), c_step.size(), 0);
    compare("planes: Post output plane 0", post0.s.get_cpu_ro(), c0.data(), c0.size(), 0);
    compare("planes: Post output plane 1", post1.s.get_cpu_ro(), c1.data(), c1.size(), 0);
    compare("planes: second Post output", post2.s.get_cpu_ro(), c2.data(), c2.size(), 0);
}

static void case_tuned_wfilter_gemm(int h, int w, int m, int k, int n) {
    // Autotune over Expand x block shapes, then ExecuteTuned: the same outputs as the interpreter
    Buf<int> in(w * h), out(w * h), wt(9);
    fill_int(in.host(), in.n, -1000, 1000);
    fill_int(wt.host(), 9, -5, 5);
    const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
    const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, 3}}};
    auto wf_io = [&](const int *ip, const int *wp, int *op) {
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {ip}, {56}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {3, PAD}, {3, PAD}}}, DEFAULT_SHM, {wp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {op}}}};
        return std::make_pair(I, O);
    };
    {
        auto io = wf_io(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), out.s.get_gpu_wo());
        const TuneResult t = Autotune<WeightedFilter, 0x111, 0xc0, 0x1111, 0x1122, 0x1124, 0x1142>(nullptr, pd, ad, io.first, io.second, {-5}, 2);
        const umi_status st = ExecuteTuned<WeightedFilter, 0x111, 0xc0, 0x1111, 0x1122, 0x1124, 0x1142>(t, nullptr, ad, io.first, io.second, {-5});
        umi_stream_synchronize(nullptr);
        std::printf("[engine] tuned wfilter: %d candidates, Expand %04llx, block %d x %d, %.1f us (status %d)\n",
                    t.candidates, (unsigned long long)t.expand, t.p_dims[2].block_size, t.p_dims[3].block_size, t.us, st);
        auto cio = wf_io(in.s.get_cpu_ro(), wt.s.get_cpu_ro(), out.cpu_out.data());
        umi_oracle::Execute<WeightedFilter, 0x111>(pd, ad, cio.first, cio.second, {-5});
        if (st != UMI_OK || t.candidates < 8) ++g_fail;
        compare("wfilter3x3 with the tuned mapping", out.s.get_cpu_ro(), out.cpu_out.data(), out.n, 0);
    }
    Buf<int> A(m * k), B(k * n), C(m * n);
    fill_int(A.host(), A.n, -100, 100);
    fill_int(B.host(), B.n, -100, 100);
    const NDTuple<ParallelDim> gpd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {m, 16, 2, 1}, {n, 16, 16, 1}}};
    const NDTuple<AccumulationDim> gad{{{1, 1}, {1, 1}, {1, 1}, {k, 16}}};
    auto g_io = [&](const int *ap, const int *bp, int *cp) {
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {m, PAD}, {k, PAD}}}, DEFAULT_SHM, {ap}, {}},
            {{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {k, PAD}, {n, PAD}}}, DEFAULT_SHM, {bp}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {m, PAD}, {n, PAD}}}, {cp}}}};
        return std::make_pair(I, O);
    };
    {
        auto io = g_io(A.s.get_gpu_ro(), B.s.get_gpu_ro(), C.s.get_gpu_wo());
        const TuneResult t = Autotune<Dot, 0x111, 0x48, 0x1111, 0x1124, 0x1142, 0x1144>(nullptr, gpd, gad, io.first, io.second, {}, 2);
        const umi_status st = ExecuteTuned<Dot, 0x111, 0x48, 0x1111, 0x1124, 0x1142, 0x1144>(t, nullptr, gad, io.first, io.second, {});
        umi_stream_synchronize(nullptr);
        std::printf("[engine] tuned gemm: %d candidates, Expand %04llx, block %d x %d, %.1f us (status %d)\n",
                    t.candidates, (unsigned long long)t.expand, t.p_dims[2].block_size, t.p_dims[3].block_size, t.us, st);
        auto cio = g_io(A.s.get_cpu_ro(), B.s.get_cpu_ro(), C.cpu_out.data());
        umi_oracle::Execute<Dot, 0x111>(gpd, gad, cio.first, cio.second, {});
        if (st != UMI_OK || t.candidates < 8) ++g_fail;
        compare("gemm with the tuned mapping", C.s.get_cpu_ro(), C.cpu_out.data(), C.n, 0);
    }
}

static void case_invalid_dims() {
    // Product of warp sizes 16 != 32: a status, not an abort, from ExecuteAsync.
    const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {4, 4, 2, 1}, {4, 8, 8, 1}}};
    const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {1, 1}}};
    const std::array<UnrolledI<int>, 1> I{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{1, PAD}, {1, PAD}, {4, PAD}, {4, PAD}}}, DEFAULT_SHM, {nullptr}, {}}}};
    const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {4, PAD}, {4, PAD}}}, {nullptr}}}};
    const umi_status st = ExecuteAsync<BoxFilter, 0x11>(nullptr, pd, ad, I, O, {0, 1});
    if (st == UMI_ERR_INVALID) {
        std::printf("[engine] invalid ParallelDims rejected: ok (%s)\n", umi_last_error());
    } else {
        ++g_fail;
        std::printf("[engine] invalid ParallelDims: FAIL (status %d)\n", st);
    }
}

// Routing (UMI_route.h): the same call with routing on and off.  The routed
// result must equal the generic engine's bit for bit (float included: both
// run the fused device arithmetic) and, for int, the interpreter; the
// look-alike strategies must not be routed.
template <class S, uint64_t PN, class E>
static void case_route(const char *name, int h, int w, int mode, const typename S::UserConstantType &c,
                       bool expect_route, int ablk3 = 3) {
    constexpr bool WEIGHTED = S::N_LOOP_IN == 2;
    Buf<E> in(w * h), out(w * h), oute(w * h), wt(9);
    if constexpr (std::is_same<E, float>::value) {
        fill_float(in.host(), in.n);
        const float wv[9] = {0.0625f, 0.125f, 0.0625f, 0.125f, 0.25f, 0.125f, 0.0625f, 0.125f, 0.0625f};
        std::copy(wv, wv + 9, wt.host());
    } else {
        fill_int(in.host(), in.n, -1000, 1000);
        fill_int(wt.host(), 9, -5, 5);
    }
    const auto M = mode ? WRAP : PAD;
    auto run = [&](const E *ip, const E *wp, E *op, int how) {  // 0 engine, 1 routed, 2 interpreter
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, ablk3}}};
        const UnrolledI<E> img{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}},
                               {{{1, M}, {1, M}, {h, M}, {w, M}}}, DEFAULT_SHM, {ip}, {(E)7}};
        const std::array<UnrolledO<E>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {op}}}};
        auto go = [&](const auto &I) {
            if (how == 2) {
                umi_oracle::Execute<S, PN>(pd, ad, I, O, c);
            } else {
                route::enable(how == 1);
                const long before = route::routed();
                Execute<S, PN>(pd, ad, I, O, c);
                if (how == 1 && (route::routed() - before == 1) != expect_route) {
                    std::printf("[engine] %s: routed %ld, expected %d\n", name, route::routed() - before, (int)expect_route);
                    g_fail = 1;
                }
                route::enable(false);
            }
        };
        if constexpr (WEIGHTED) {
            const UnrolledI<E> wts{{{{}, {}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {3, PAD}, {3, PAD}}},
                                   DEFAULT_SHM, {wp}, {}};
            go(std::array<UnrolledI<E>, 2>{{img, wts}});
        } else {
            go(std::array<UnrolledI<E>, 1>{{img}});
        }
    };
    run(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), oute.s.get_gpu_wo(), 0);
    run(in.s.get_gpu_ro(), wt.s.get_gpu_ro(), out.s.get_gpu_wo(), 1);
    run(in.s.get_cpu_ro(), wt.s.get_cpu_ro(), out.cpu_out.data(), 2);
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s: routing on vs off", name);
    compare(buf, out.s.get_cpu_ro(), oute.s.get_cpu_ro(), out.n, 0);
    std::snprintf(buf, sizeof buf, "%s: vs interpreter", name);
    compare(buf, out.s.get_cpu_ro(), out.cpu_out.data(), out.n, std::is_same<E, float>::value ? 1e-5 : 0);
}

int main() {
    route::enable(false);  // the generic engine; routing has its own cases below
    case_box(300, 500, 0);
    case_box(37, 53, 1);
    case_wfilter(250, 300);
    case_bilateral(60, 70, 2.0f, 0.3f);
    case_bilateral(50, 40, 4.0f, 10.0f);
    case_iimg(300, 500);
    case_full_me(96, 160, 8, 15);
    case_nv12_chroma_planes(72, 104, 9);
    case_cost_volume(40, 200, 7, 15);
    case_fprop(2, 13, 9, 20, 24, 7, 2);
    case_gemm(100, 100, 200);
    case_planes(21, 19, 8);
    case_tuned_wfilter_gemm(123, 301, 96, 80, 200);
    case_invalid_dims();
    case_route<BoxFilter, 0x11, int>("route: 3x3 filter int32 PAD", 300, 500, 0, {-5, 2}, true);
    case_route<BoxFilter, 0x11, int>("route: 3x3 filter int32 clamp", 37, 53, 1, {7, -3}, true);
    case_route<WeightedFilter, 0x111, int>("route: weighted 3x3 int32", 250, 300, 0, {-5}, true);
    case_route<BoxFilterF, 0x11, float>("route: 3x3 filter f32 clamp", 123, 301, 1, {0.5f, 0.25f}, true);
    case_route<WeightedFilterF, 0x111, float>("route: weighted 3x3 f32 (fma)", 129, 257, 0, {-0.125f}, true);
    case_route<BoxFilterSwapped, 0x11, int>("route: constants in the other order stay on the engine", 40, 50, 0, {2, -5}, false);
    case_route<BoxFilterPlusOne, 0x11, int>("route: a different body stays on the engine", 40, 50, 1, {-5, 2}, false);
    case_route<BoxFilterF, 0x11, float>("route: taps split over accumulation blocks stay on the engine", 40, 50, 0, {0.5f, 0.25f}, false, 2);
    std::printf("[engine] %s\n", g_fail ? "FAILED" : "all cases passed");
    return g_fail ? 1 : 0;
}
