// UMI.h — descriptor types and strategy macros of the UMI operator.
//
// Source-compatible with the public part of src/UMI.h of johnjohnlin/UMI
// (names, field order and macro names are kept so strategy structs and
// brace-initialised descriptors written for it compile unchanged):
//   constants                      src/UMI.h:41-50
//   simple_array, NDTuple, CeilDiv, Slice ...   src/UMI.h:54-75
//   AccumulationDim/ParallelDim/ViewDim/PhysicalDim  src/UMI.h:77-93
//   UnrolledI / UnrolledO          src/UMI.h:95-111
//   STRATEGY_* / IMEM / OMEM / DEFAULT_SHM       src/UMI.h:113-137
// The reference's planner internals (Int4, *_packed, Precompute*,
// src/UMI.h:139-386) are not part of the call surface and have no
// counterpart: the MI355X engine addresses memory in closed form.
#pragma once
#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <tuple>
#include <type_traits>
#include <vector>

#if defined(__HIPCC__)
#define DHFunc __host__ __device__
#define DFunc __device__
#else
#define DHFunc
#define DFunc
#endif

namespace UMI {

template <class Int>
inline constexpr bool IsPowOf2(const Int i) {
    return i > 0 && (i & (i - 1)) == 0;
}
static constexpr int MAX_PTR = 4;   // planes per array (e.g. Y/U/V)
static constexpr int DIM = 4;       // every index space is 4-D
static constexpr int MAX_WARP = 32;
static constexpr int WARP_SIZE = 32;  // logical warp of the descriptors (not the 64-lane wavefront)
static constexpr int WORD_SIZE = 4;
static constexpr int MAX_BLOCK_SIZE = MAX_WARP * WARP_SIZE;
static constexpr int LOAD_WARP = 2;
static constexpr int MIN_LOAD_THREAD = WARP_SIZE * LOAD_WARP;
using namespace std;

constexpr DHFunc int SafeArrayLength(int i) { return i >= 1 ? i : 1; }

// Fixed-size array usable in device code; the strategy's in/out operands.
template <class T, size_t N>
struct simple_array {
    DHFunc T &operator[](size_t i) { return data[i]; }
    DHFunc const T &operator[](size_t i) const { return data[i]; }
    DHFunc simple_array() {}
    simple_array(const array<T, N> &a) {
        for (size_t i = 0; i < N; ++i) data[i] = a[i];
    }
    T data[SafeArrayLength(N)];
};

template <class T>
using identity_t = typename enable_if<true, T>::type;
template <class T>
using NDTuple = array<T, DIM>;
template <class Int>
inline constexpr Int CeilDiv(const Int a, const Int b) { return (a - 1) / b + 1; }
template <class Int>
inline constexpr Int CeilDivBit(const Int a, const Int b) { return ((a - 1) >> b) + 1; }
template <class Int>
inline constexpr Int CeilAlign(const Int a, const Int b) { return CeilDiv(a, b) * b; }
template <class Int>
inline constexpr Int CeilAlignBit(const Int a, const Int b) { return CeilDivBit(a, b) << b; }
template <class Int>
inline constexpr Int CeilAlignBitExact(const Int a, const Int b) { return (CeilDivBit(a, b) | Int(1)) << b; }
// Bit field `nth` of width `b` (counting from the least significant end).
constexpr DHFunc uint64_t Slice(const uint64_t i, int b, int nth) { return (i >> (nth * b)) & ((b >= 64) ? ~0ull : ((1ull << b) - 1)); }

// Accumulation space: `total_size` steps walked in chunks of `block_size`.
struct AccumulationDim {
    int32_t total_size, block_size;
};
// Parallel space: `total_size` points; block/warp/shuf describe the mapping
// of points to threads (a performance hint: results never depend on it).
struct ParallelDim {
    int32_t total_size, block_size, warp_size, warp_shuf;
};
// One axis of an affine view: adds ofs + index * stride to physical dim
// `physical_dim_id`.
struct ViewDim {
    int32_t physical_dim_id, ofs, stride;
};
// One axis of a dense row-major array.  Out-of-range reads: PAD -> the
// array's padding value; WRAP -> the coordinate is clamped to the edge.
struct PhysicalDim {
    enum BoundaryMode : int32_t { PAD, WRAP };
    int32_t size;
    BoundaryMode mode;
};

template <typename T>
struct UnrolledI {
    NDTuple<ViewDim> p_view_dims;
    NDTuple<ViewDim> a_view_dims;
    NDTuple<PhysicalDim> physical_dims;
    NDTuple<int32_t> shm_dims;  // accepted, unused (so is the reference's, src/UMI_gpu.h:474-475)
    const T *ptr[MAX_PTR];
    const T padding[MAX_PTR];
};

template <typename T>
struct UnrolledO {
    NDTuple<ViewDim> p_view_dims;
    NDTuple<ViewDim> a_view_dims;
    NDTuple<PhysicalDim> physical_dims;
    T *ptr[MAX_PTR];
};

// Strategy declaration macros.  A strategy struct declares its element and
// user-constant types, then its Pre / Loop / Post stages with the number of
// input and output arrays each stage reads and writes.
#define STRATEGY_TYPEDEF_STUB(e, c) \
    typedef e ElementType;          \
    typedef c UserConstantType;     \
    template <int s>                \
    using A = UMI::simple_array<ElementType, s>;
#define STRATEGY_PRE_STUB(ni, no)            \
    static constexpr int N_PRE_IN = ni;      \
    static constexpr int N_PRE_OUT = no;     \
    DHFunc void Pre(const UserConstantType &c, const A<N_PRE_IN * UMI::MAX_PTR> &in, A<N_PRE_OUT * UMI::MAX_PTR> &out)
#define STRATEGY_LOOP_STUB(ni, no)           \
    static constexpr int N_LOOP_IN = ni;     \
    static constexpr int N_LOOP_OUT = no;    \
    DHFunc void Loop(const UserConstantType &c, const A<N_LOOP_IN * UMI::MAX_PTR> &in, A<N_LOOP_OUT * UMI::MAX_PTR> &out)
#define STRATEGY_POST_STUB(ni, no)                                        \
    static constexpr int N_POST_IN = ni;                                  \
    static constexpr int N_POST_OUT = no;                                 \
    static constexpr int TOTAL_IN = N_PRE_IN + N_LOOP_IN + N_POST_IN;     \
    static constexpr int TOTAL_OUT = N_PRE_OUT + N_LOOP_OUT + N_POST_OUT; \
    static constexpr int TOTAL = TOTAL_IN + TOTAL_OUT;                    \
    DHFunc void Post(const UserConstantType &c, const A<N_POST_IN * UMI::MAX_PTR> &in, A<N_POST_OUT * UMI::MAX_PTR> &out)
// Empty stages.  (The reference spells the empty Pre "STRATEGY_POST_PRE_".)
#define STRATEGY_POST_PRE_ STRATEGY_PRE_STUB(0, 0) {}
#define STRATEGY_PRE_STUB_ STRATEGY_PRE_STUB(0, 0) {}
#define STRATEGY_LOOP_STUB_ STRATEGY_LOOP_STUB(0, 0) {}
#define STRATEGY_POST_STUB_ STRATEGY_POST_STUB(0, 0) {}
#define IMEM(i, j) in[(i) * UMI::MAX_PTR + (j)]
#define OMEM(i, j) out[(i) * UMI::MAX_PTR + (j)]
#define DEFAULT_SHM {-1, -1, -1, -1}

}  // namespace UMI
