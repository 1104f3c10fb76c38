// Timing of the generic UMI::Execute<> engine on the reference examples'
// strategies and descriptors (main/*.cu, johnjohnlin/UMI), at larger sizes:
// hipEvents around repeated ExecuteAsync launches on one stream, inputs
// resident.  One JSON line per case: time per launch, algorithmic bytes
// (inputs read once + outputs written once) and the rate; the tuner's pick;
// for the two stencils also the time of the same call routed to the
// specialised kernel.  Not a test: the results are checked by test_engine
// (same engine and routes, interpreter oracle).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>
#include "SyncedMemory.h"
#include "UMI_gpu.h"
#include "strategies.h"
using namespace UMI;
static const auto PAD = PhysicalDim::PAD;
static const auto WRAP = PhysicalDim::WRAP;

template <class F>
static double time_us(F &&launch, int reps) {
    hipStream_t st = nullptr;
    for (int i = 0; i < 3; ++i) launch();
    (void)hipDeviceSynchronize();
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, st);
    for (int i = 0; i < reps; ++i) launch();
    (void)hipEventRecord(b, st);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms * 1e3 / reps;
}

static void report(const char *name, double us, double bytes, double ops, const char *ops_unit,
                   const TuneResult *t = nullptr, double routed_us = -1) {
    std::printf("{\"case\": \"%s\", \"us\": %.2f, \"bytes\": %.0f, \"gbs\": %.1f, \"ops\": %.0f, \"ops_per_s\": %.4g, "
                "\"ops_unit\": \"%s\"",
                name, us, bytes, bytes / (us * 1e-6) / 1e9, ops, ops / (us * 1e-6), ops_unit);
    if (t)
        std::printf(", \"tuned_us\": %.2f, \"tuned_gbs\": %.1f, \"tuned_ops_per_s\": %.4g, \"tuned_expand\": \"%04llx\", "
                    "\"tuned_block\": [%d, %d, %d, %d], \"candidates\": %d",
                    t->us, bytes / (t->us * 1e-6) / 1e9, ops / (t->us * 1e-6), (unsigned long long)t->expand,
                    t->p_dims[0].block_size, t->p_dims[1].block_size, t->p_dims[2].block_size, t->p_dims[3].block_size,
                    t->candidates);
    if (routed_us >= 0)
        std::printf(", \"routed_us\": %.2f, \"routed_gbs\": %.1f", routed_us, bytes / (routed_us * 1e-6) / 1e9);
    std::printf("}\n");
    std::fflush(stdout);
}

template <class T>
static T *dev(size_t n, int v) {
    T *p = nullptr;
    if (umi_malloc((void **)&p, n * sizeof(T)) != UMI_OK) std::abort();
    std::vector<T> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = (T)((i * 2654435761u >> 7) % (unsigned)v);
    umi_memcpy(p, h.data(), n * sizeof(T), 1);
    return p;
}

static void ok(umi_status s) {
    if (s != UMI_OK) {
        std::fprintf(stderr, "ExecuteAsync: %s\n", umi_last_error());
        std::abort();
    }
}

// Times ExecuteAsync with routing to the specialised kernels on; aborts if
// the call was not routed.
template <class F>
static double routed_time_us(F &&launch, int reps) {
    route::enable(true);
    const long before = route::routed();
    const double us = time_us(launch, reps);
    if (route::routed() == before) {
        std::fprintf(stderr, "expected the call to be routed\n");
        std::abort();
    }
    route::enable(false);
    return us;
}

int main() {
    const int R = 20;
    route::enable(false);  // the generic engine (and its tuner) unless a case asks for routing
    {  // main/wfilter3x3.cu:39-68 at 1024 x 1024
        const int h = 1024, w = 1024;
        int *in = dev<int>(h * w, 1000), *wt = dev<int>(9, 5), *out = dev<int>(h * w, 1);
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, 3}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in}, {56}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {3, PAD}, {3, PAD}}}, DEFAULT_SHM, {wt}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {out}}}};
        const double us = time_us([&] { ok(ExecuteAsync<WeightedFilter, 0x111, 0x1122, 0xc0>(nullptr, pd, ad, I, O, {-5})); }, R);
        const TuneResult t = Autotune<WeightedFilter, 0x111, 0xc0, 0x1111, 0x1112, 0x1122, 0x1124, 0x1142, 0x1144>(nullptr, pd, ad, I, O, {-5}, R);
        const double ru = routed_time_us([&] { ok(ExecuteAsync<WeightedFilter, 0x111, 0x1122, 0xc0>(nullptr, pd, ad, I, O, {-5})); }, R);
        report("wfilter3x3 int32 1024x1024 (Expand 1122, ZeroStrideOpt c0)", us, 8.0 * h * w, 9.0 * h * w, "MAC", &t, ru);
    }
    {  // main/filter3x3.cu:36-58 at 1024 x 1024
        const int h = 1024, w = 1024;
        int *in = dev<int>(h * w, 1000), *out = dev<int>(h * w, 1);
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {3, 3}, {3, 3}}};
        const std::array<UnrolledI<int>, 1> I{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}},
                                                {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in}, {999}}}};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {out}}}};
        const double us = time_us([&] { ok(ExecuteAsync<BoxFilter, 0x11, 0x1111, 0x1122>(nullptr, pd, ad, I, O, {-5, 2})); }, R);
        const TuneResult t = Autotune<BoxFilter, 0x11, 0, 0x1111, 0x1112, 0x1122, 0x1124, 0x1142, 0x1144>(nullptr, pd, ad, I, O, {-5, 2}, R);
        const double ru = routed_time_us([&] { ok(ExecuteAsync<BoxFilter, 0x11, 0x1111, 0x1122>(nullptr, pd, ad, I, O, {-5, 2})); }, R);
        report("filter3x3 int32 1024x1024", us, 8.0 * h * w, 9.0 * h * w, "add", &t, ru);
    }
    {  // main/bilateral.cu:54-99, 5 x 5 window, 1024 x 1024
        const int h = 1024, w = 1024, ofs = 2, k = 5;
        float *in = dev<float>(h * w, 256), *wt = dev<float>(k, 2), *out = dev<float>(h * w, 1);
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 2, 1}, {w, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {k, k}, {k, k}}};
        const std::array<UnrolledI<float>, 4> I{{
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {}, {}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in}, {}},
            {{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -ofs, 1}, {3, -ofs, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in}, {0.0f}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {3, 0, 1}, {}}}, {{{1, PAD}, {1, PAD}, {1, PAD}, {k, PAD}}}, DEFAULT_SHM, {wt}, {}},
            {{{{}, {}, {}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {1, PAD}, {k, PAD}}}, DEFAULT_SHM, {wt}, {}},
        }};
        const std::array<UnrolledO<float>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {out}}}};
        const double us = time_us([&] { ok(ExecuteAsync<BilateralF, 0x11111>(nullptr, pd, ad, I, O, {-50.f})); }, R);
        const TuneResult t = Autotune<BilateralF, 0x11111, 0, 0x1111, 0x1112, 0x1121, 0x1122, 0x1124>(nullptr, pd, ad, I, O, {-50.f}, R);
        report("bilateral f32 1024x1024 5x5", us, 8.0 * h * w, 25.0 * h * w, "tap", &t);
    }
    {  // main/iimg.cu:37-82, both passes, 1080 x 1920
        const int h = 1080, w = 1920;
        int *in = dev<int>(h * w, 256), *tp = dev<int>(h * w, 1), *out = dev<int>(h * w, 1);
        const NDTuple<ParallelDim> pd1{{{1, 1, 1, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}, {h, 128, 32, 1}}};
        const NDTuple<AccumulationDim> ad1{{{1, 1}, {1, 1}, {1, 1}, {w, 32}}};
        const std::array<UnrolledI<int>, 1> I1{{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in}, {}}}};
        const std::array<UnrolledO<int>, 1> O1{{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {tp}}}};
        const NDTuple<ParallelDim> pd2{{{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {w, 128, 32}}};
        const NDTuple<AccumulationDim> ad2{{{1, 1}, {1, 1}, {1, 1}, {h, 32}}};
        const std::array<UnrolledI<int>, 1> I2{{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {}, {tp}, {}}}};
        const std::array<UnrolledO<int>, 1> O2{{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, {out}}}};
        const double us = time_us([&] {
            ok(ExecuteAsync<RunningSum, 0x11>(nullptr, pd1, ad1, I1, O1));
            ok(ExecuteAsync<RunningSum, 0x11>(nullptr, pd2, ad2, I2, O2));
        }, R);
        report("iimg int32 1080x1920, two passes", us, 16.0 * h * w, 2.0 * h * w, "add");
    }
    {  // main/full_me.cu:34-64 at 1080p: 8x8 blocks, 15 x 15 offsets, SAD volume
        const int h = 1080, w = 1920, b = 8, sr = 15, sr_ofs = -(sr / 2);
        int *fr = dev<int>(h * w, 256), *sad = dev<int>((size_t)(h / b) * (w / b) * sr * sr, 1);
        const NDTuple<ParallelDim> pd{{{h / b, 1, 1, 1}, {w / b, 1, 1, 1}, {sr, 16, 2, 1}, {sr, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {b, b}, {b, b}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{2, 0, b}, {3, 0, b}, {2, sr_ofs, 1}, {3, sr_ofs, 1}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fr}, {}},
            {{{{2, 0, b}, {3, 0, b}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fr}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{h / b, WRAP}, {w / b, WRAP}, {sr, WRAP}, {sr, WRAP}}}, {sad}}}};
        const double us = time_us([&] { ok(ExecuteAsync<Sad, 0x111, 0x2211>(nullptr, pd, ad, I, O)); }, 5);
        const TuneResult t = Autotune<Sad, 0x111, 0xc0, 0x1111, 0x2211, 0x1122, 0x2222, 0x4411>(nullptr, pd, ad, I, O, {}, 5);
        const double nsad = (double)(h / b) * (w / b) * sr * sr;
        report("full_me int32 1920x1080, 8x8, 15x15 SAD volume", us, 4.0 * h * w + 4.0 * nsad, nsad * b * b, "absdiff", &t);
    }
    {  // main/cost_volume.cu:35-65 at 1242 x 375, 7-wide window, 15 shifts
        const int h = 375, w = 1242, k = 7, sr = 15, k_ofs = k / 2, sr_ofs = sr / 2;
        int *fr = dev<int>(h * w, 256), *cv = dev<int>((size_t)h * w * sr, 1);
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {h, 4, 1, 1}, {w, 8, 2, 1}, {sr, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {k, k}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {2, 0, 1}, {3, 0, 1}, {3, -sr_ofs, 1}}}, {{{}, {}, {}, {3, -k_ofs, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fr}, {}},
            {{{{}, {2, 0, 1}, {3, 0, 1}, {}}}, {{{}, {}, {}, {3, -k_ofs, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {fr}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{1, WRAP}, {h, WRAP}, {w, WRAP}, {sr, WRAP}}}, {cv}}}};
        const double us = time_us([&] { ok(ExecuteAsync<Sad, 0x111>(nullptr, pd, ad, I, O)); }, R);
        const TuneResult t = Autotune<Sad, 0x111, 0, 0x1111, 0x1211, 0x1121, 0x1221, 0x1411>(nullptr, pd, ad, I, O, {}, R);
        report("cost_volume int32 1242x375, 7-wide, 15 shifts", us, 4.0 * h * w + 4.0 * h * w * sr, (double)h * w * sr * k, "absdiff", &t);
    }
    {  // main/fprop.cu:53-82: N=16, C=64, 56x56, K=64, 3x3, stride 1, pad 1 (int32)
        const int bs = 16, ich = 64, och = 64, ih = 56, iw = 56, f = 3, s = 1, ofs = -1, oh = 56, ow = 56;
        int *in = dev<int>((size_t)bs * ich * ih * iw, 10), *wt = dev<int>((size_t)och * ich * f * f, 10);
        int *out = dev<int>((size_t)bs * och * oh * ow, 1);
        const NDTuple<ParallelDim> pd{{{bs, 1, 1, 1}, {och, 4, 1, 1}, {oh, 8, 4, 1}, {ow, 8, 8, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {ich, 6}, {f, f}, {f, f}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{0, 0, 1}, {}, {2, 0, s}, {3, 0, s}}}, {{{}, {1, 0, 1}, {2, ofs, 1}, {3, ofs, 1}}}, {{{bs, PAD}, {ich, PAD}, {ih, PAD}, {iw, PAD}}}, DEFAULT_SHM, {in}, {}},
            {{{{}, {0, 0, 1}, {}, {}}}, {{{}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {{{och, PAD}, {ich, PAD}, {f, PAD}, {f, PAD}}}, DEFAULT_SHM, {wt}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{bs, PAD}, {och, PAD}, {oh, PAD}, {ow, PAD}}}, {out}}}};
        const double us = time_us([&] { ok(ExecuteAsync<Dot>(nullptr, pd, ad, I, O)); }, 5);
        const TuneResult t = Autotune<Dot, 0x111, 0xc2, 0x1111, 0x1211, 0x1112, 0x1212, 0x1412, 0x1214, 0x1414>(nullptr, pd, ad, I, O, {}, 5);
        const double macs = (double)bs * och * oh * ow * ich * f * f;
        report("fprop int32 16x64x56x56 -> 64, 3x3", us, 4.0 * (bs * ich * ih * iw + och * ich * f * f + (double)bs * och * oh * ow), macs, "MAC", &t);
    }
    {  // main/gemm.cu:37-67 at 1024^3 (int32)
        const int m = 1024, k = 1024, n = 1024;
        int *A = dev<int>((size_t)m * k, 100), *B = dev<int>((size_t)k * n, 100), *C = dev<int>((size_t)m * n, 1);
        const NDTuple<ParallelDim> pd{{{1, 1, 1, 1}, {1, 1, 1, 1}, {m, 16, 2, 1}, {n, 16, 16, 1}}};
        const NDTuple<AccumulationDim> ad{{{1, 1}, {1, 1}, {1, 1}, {k, 16}}};
        const std::array<UnrolledI<int>, 2> I{{
            {{{{}, {}, {2, 0, 1}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {m, PAD}, {k, PAD}}}, DEFAULT_SHM, {A}, {}},
            {{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {k, PAD}, {n, PAD}}}, DEFAULT_SHM, {B}, {}},
        }};
        const std::array<UnrolledO<int>, 1> O{{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {m, PAD}, {n, PAD}}}, {C}}}};
        const double us = time_us([&] { ok(ExecuteAsync<Dot, 0x111, 0x1124, 0x48>(nullptr, pd, ad, I, O)); }, 5);
        const TuneResult t = Autotune<Dot, 0x111, 0x48, 0x1111, 0x1124, 0x1142, 0x1144, 0x1148, 0x1184>(nullptr, pd, ad, I, O, {}, 5);
        report("gemm int32 1024^3 (Expand 1124, ZeroStrideOpt 48)", us, 4.0 * (2.0 * m * k + (double)m * n), (double)m * n * k, "MAC", &t);
    }
    return 0;
}
