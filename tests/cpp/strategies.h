// Strategy structs shared by the engine test (test_engine.cpp) and the
// engine timing program (engine_bench.cpp).  They restate the reference
// examples' strategies (main/*.cu, johnjohnlin/UMI) in this project's code,
// plus a multi-plane strategy the examples do not have.
#pragma once
#include <cmath>
#include <cstdlib>
#include "UMI.h"

struct BoxFilter {  // main/filter3x3.cu:10-17
    int accum;
    struct Constant {
        int init, mul;
    };
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(1, 0) { accum += IMEM(0, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum * c.mul; }
};

struct WeightedFilter {  // main/wfilter3x3.cu:8-15
    int accum;
    struct Constant {
        int init;
    };
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(2, 0) { accum += IMEM(0, 0) * IMEM(1, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum; }
};

// Float forms of the two filters (the same bodies over float).
struct BoxFilterF {
    float accum;
    struct Constant {
        float init, mul;
    };
    STRATEGY_TYPEDEF_STUB(float, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(1, 0) { accum += IMEM(0, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum * c.mul; }
};
struct WeightedFilterF {
    float accum;
    struct Constant {
        float init;
    };
    STRATEGY_TYPEDEF_STUB(float, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(2, 0) { accum += IMEM(0, 0) * IMEM(1, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum; }
};
// Look-alikes that a specialised kernel does not compute: constants in the
// other order, and one more per tap.  Both must stay on the generic engine.
struct BoxFilterSwapped {
    int accum;
    struct Constant {
        int mul, init;
    };
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(1, 0) { accum += IMEM(0, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum * c.mul; }
};
struct BoxFilterPlusOne {
    int accum;
    struct Constant {
        int init, mul;
    };
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = c.init; }
    STRATEGY_LOOP_STUB(1, 0) { accum += IMEM(0, 0) + 1; }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum * c.mul; }
};

struct BilateralF {  // main/bilateral.cu:9-24
    float wsum, wxsum, center;
    struct Constant {
        float norm;
    };
    STRATEGY_TYPEDEF_STUB(float, Constant)
    STRATEGY_PRE_STUB(1, 0) {
        wsum = wxsum = 0;
        center = IMEM(0, 0);
    }
    STRATEGY_LOOP_STUB(3, 0) {
        float n = IMEM(0, 0), d = n - center;
        float w = expf(d * d * c.norm) * IMEM(1, 0) * IMEM(2, 0);
        wsum += w;
        wxsum += w * n;
    }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = wxsum / wsum; }
};

struct RunningSum {  // main/iimg.cu:11-18: every accumulation step writes the running sum
    int accum;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = 0; }
    STRATEGY_LOOP_STUB(1, 1) { OMEM(0, 0) = accum += IMEM(0, 0); }
    STRATEGY_POST_STUB_
};

struct Sad {  // main/full_me.cu:8-15, main/cost_volume.cu:8-15
    int sad;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { sad = 0; }
    STRATEGY_LOOP_STUB(2, 0) { sad += abs(IMEM(0, 0) - IMEM(1, 0)); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = sad; }
};

// SAD over two planes at once (the U and V planes of a chroma array):
// the loop of main/full_me.cu:8-15 summed over the planes of its inputs.
struct Sad2 {
    int sad;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { sad = 0; }
    STRATEGY_LOOP_STUB(2, 0) { sad += abs(IMEM(0, 0) - IMEM(1, 0)) + abs(IMEM(0, 1) - IMEM(1, 1)); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = sad; }
};

struct Dot {  // main/gemm.cu:8-15, main/fprop.cu:10-17
    int accum;
    struct Constant {};
    STRATEGY_TYPEDEF_STUB(int, Constant)
    STRATEGY_PRE_STUB(0, 0) { accum = 0; }
    STRATEGY_LOOP_STUB(2, 0) { accum += IMEM(0, 0) * IMEM(1, 0); }
    STRATEGY_POST_STUB(0, 1) { OMEM(0, 0) = accum; }
};

// Three planes in, two planes out; Pre reads a bias array, Post writes a
// second output; the Loop output records a per-step value.
struct Planes {
    float s0, s1, bias;
    struct Constant {
        float scale;
    };
    STRATEGY_TYPEDEF_STUB(float, Constant)
    STRATEGY_PRE_STUB(1, 1) {
        bias = IMEM(0, 0);
        s0 = s1 = 0.f;
        OMEM(0, 0) = bias * c.scale;
    }
    STRATEGY_LOOP_STUB(1, 1) {
        s0 += IMEM(0, 0) + IMEM(0, 1);
        s1 += IMEM(0, 2);
        OMEM(0, 0) = s0 - s1;
    }
    STRATEGY_POST_STUB(1, 2) {
        OMEM(0, 0) = s0 + bias;
        OMEM(0, 1) = s1 + IMEM(0, 0);
        OMEM(1, 0) = s0 * c.scale;
    }
};

