// Pins the CPU interpreter of UMI::Execute (oracle/umi_interp.h) on the host,
// against the reference examples' own known answers (johnjohnlin/UMI):
//   filter3x3 on an iota image, init -5, mul 2, pad 999  (main/filter3x3.cu:21-78)
//   full-search SAD volume of an iota frame: interior 64 |W dy + dx|  (main/full_me.cu)
//   1-D cost volume of an iota frame: interior 7 |s|     (main/cost_volume.cu)
//   integral image of all ones: (y+1)(x+1)              (main/iimg.cu:31,84-112)
//   gemm of iota operands against a direct triple loop   (main/gemm.cu:69-78)
// Built with g++ (no HIP): the strategy structs are host-callable.
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "umi_interp.h"

using namespace UMI;
static const auto PAD = PhysicalDim::PAD;
static const auto WRAP = PhysicalDim::WRAP;
static int fails = 0;
#define CHECK(cond, ...)                     \
    do {                                     \
        if (!(cond)) {                       \
            ++fails;                         \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");               \
        }                                    \
    } while (0)

struct Box {
    int accum;
    struct Constant { int init, mul; };
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(1, 0) { accum += IMEM(0, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum * c.mul; }
};
struct Sad {
    int sad;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { sad = 0; }
    STRATEGY_LOOP_STUB(2, 0) { sad += abs(IMEM(0, 0) - IMEM(1, 0)); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = sad; }
};
struct Scan {
    int accum;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = 0; }
    STRATEGY_LOOP_STUB(1, 1) { OMEM(0, 0) = accum += IMEM(0, 0); }
    STRATEGY_POST_STUB_
};
struct Dot {
    int accum;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = 0; }
    STRATEGY_LOOP_STUB(2, 0) { accum += IMEM(0, 0) * IMEM(1, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum; }
};

int main() {
    {  // filter3x3
        const int w = 50, h = 30;
        std::vector<int> in(w * h), out(w * h);
        std::iota(in.begin(), in.end(), 0);
        umi_oracle::Execute<Box, 0x11>(
            {{{1, 1, 1, 1}, {1, 1, 1, 1}, {h, 16, 4, 1}, {w, 16, 8, 1}}}, {{{1, 1}, {1, 1}, {3, 3}, {3, 3}}},
            {{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{}, {}, {2, -1, 1}, {3, -1, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in.data()}, {999}}}},
            {{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {h, PAD}, {w, PAD}}}, {out.data()}}}}, {-5, 2});
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                int s = -5;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int yy = y + dy, xx = x + dx;
                        s += (yy < 0 || yy >= h || xx < 0 || xx >= w) ? 999 : w * yy + xx;
                    }
                CHECK(out[y * w + x] == 2 * s, "filter3x3 at %d,%d", y, x);
            }
    }
    {  // full_me iota: interior SAD = b^2 |w dy + dx|
        const int w = 96, h = 64, b = 8, sr = 15, ofs = -7;
        std::vector<int> f(w * h), vol((w / b) * (h / b) * sr * sr);
        std::iota(f.begin(), f.end(), 0);
        umi_oracle::Execute<Sad, 0x111>(
            {{{h / b, 1, 1, 1}, {w / b, 1, 1, 1}, {sr, 16, 2, 1}, {sr, 16, 16, 1}}}, {{{1, 1}, {1, 1}, {b, b}, {b, b}}},
            {{{{{{2, 0, b}, {3, 0, b}, {2, ofs, 1}, {3, ofs, 1}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {f.data()}, {}},
              {{{{2, 0, b}, {3, 0, b}, {}, {}}}, {{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {f.data()}, {}}}},
            {{{{{{0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{h / b, WRAP}, {w / b, WRAP}, {sr, WRAP}, {sr, WRAP}}}, {vol.data()}}}});
        for (int by = 1; by < h / b - 1; ++by)
            for (int bx = 1; bx < w / b - 1; ++bx)
                for (int sy = 0; sy < sr; ++sy)
                    for (int sx = 0; sx < sr; ++sx) {
                        const int v = vol[((by * (w / b) + bx) * sr + sy) * sr + sx];
                        CHECK(v == b * b * abs(w * (sy + ofs) + (sx + ofs)), "full_me %d %d %d %d", by, bx, sy, sx);
                    }
    }
    {  // 1-D cost volume iota: interior 7 |s|
        const int w = 60, h = 4, k = 7, sr = 15;
        std::vector<int> f(w * h), cv(w * h * sr);
        std::iota(f.begin(), f.end(), 0);
        umi_oracle::Execute<Sad, 0x111>(
            {{{1, 1, 1, 1}, {h, 4, 1, 1}, {w, 8, 2, 1}, {sr, 16, 16, 1}}}, {{{1, 1}, {1, 1}, {1, 1}, {k, k}}},
            {{{{{{}, {2, 0, 1}, {3, 0, 1}, {3, -sr / 2, 1}}}, {{{}, {}, {}, {3, -k / 2, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {f.data()}, {}},
              {{{{}, {2, 0, 1}, {3, 0, 1}, {}}}, {{{}, {}, {}, {3, -k / 2, 1}}}, {{{1, WRAP}, {1, WRAP}, {h, WRAP}, {w, WRAP}}}, DEFAULT_SHM, {f.data()}, {}}}},
            {{{{{{}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{1, WRAP}, {h, WRAP}, {w, WRAP}, {sr, WRAP}}}, {cv.data()}}}});
        for (int y = 0; y < h; ++y)
            for (int x = 12; x < w - 12; ++x)
                for (int s = 0; s < sr; ++s) CHECK(cv[(y * w + x) * sr + s] == 7 * abs(s - 7), "cost volume %d %d %d", y, x, s);
    }
    {  // integral image of ones, two passes (second with warp_shuf 0)
        const int w = 70, h = 40;
        std::vector<int> in(w * h, 1), tmp(w * h), out(w * h);
        umi_oracle::Execute<Scan, 0x11>(
            {{{1, 1, 1, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}, {h, 128, 32, 1}}}, {{{1, 1}, {1, 1}, {1, 1}, {w, 32}}},
            {{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, DEFAULT_SHM, {in.data()}, {}}}},
            {{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {tmp.data()}}}});
        umi_oracle::Execute<Scan, 0x11>(
            {{{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {w, 128, 32}}}, {{{1, 1}, {1, 1}, {1, 1}, {h, 32}}},
            {{{{{{}, {}, {}, {2, 0, 1}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {w, PAD}, {h, PAD}}}, {}, {tmp.data()}, {}}}},
            {{{{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {h, PAD}, {w, PAD}}}, {out.data()}}}});
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) CHECK(out[y * w + x] == (y + 1) * (x + 1), "iimg %d %d", y, x);
    }
    {  // gemm with a partial last accumulation block (k = 100, block 16)
        const int m = 20, k = 100, n = 30;
        std::vector<int> A(m * k), B(k * n), C(m * n);
        std::iota(A.begin(), A.end(), 0);
        std::iota(B.begin(), B.end(), 0);
        umi_oracle::Execute<Dot, 0x111>(
            {{{1, 1, 1, 1}, {1, 1, 1, 1}, {m, 16, 2, 1}, {n, 16, 16, 1}}}, {{{1, 1}, {1, 1}, {1, 1}, {k, 16}}},
            {{{{{{}, {}, {2, 0, 1}, {}}}, {{{}, {}, {}, {3, 0, 1}}}, {{{1, PAD}, {1, PAD}, {m, PAD}, {k, PAD}}}, DEFAULT_SHM, {A.data()}, {}},
              {{{{}, {}, {}, {3, 0, 1}}}, {{{}, {}, {}, {2, 0, 1}}}, {{{1, PAD}, {1, PAD}, {k, PAD}, {n, PAD}}}, DEFAULT_SHM, {B.data()}, {}}}},
            {{{{{{}, {}, {2, 0, 1}, {3, 0, 1}}}, {}, {{{}, {}, {m, PAD}, {n, PAD}}}, {C.data()}}}});
        for (int y = 0; y < m; ++y)
            for (int x = 0; x < n; ++x) {
                unsigned s = 0;
                for (int kk = 0; kk < k; ++kk) s += (unsigned)A[y * k + kk] * (unsigned)B[kk * n + x];
                CHECK(C[y * n + x] == (int)s, "gemm %d %d", y, x);
            }
    }
    std::printf(fails ? "interpreter: %d failures\n" : "interpreter: all known answers match\n", fails);
    return fails ? 1 : 0;
}
