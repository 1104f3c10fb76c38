// UMI_gpu.h — the UMI operator on MI355X (gfx950): UMI::Execute<>.
//
// Same template signature and call semantics as UMI::Execute of
// src/UMI_gpu.h:409-422 (johnjohnlin/UMI):
//
//   template <class Umi, uint64_t PtrNum, uint64_t Expand, uint64_t ZeroStrideOpt, EType, CType>
//   void Execute(p_dims, a_dims, imems, omems, c);
//
// For every point p of the parallel space (dim 3 fastest) one strategy
// object runs Pre, then Loop over the accumulation space in block-major
// raster order (blocks of a_dims[d].block_size, dim 3 fastest; steps beyond
// total_size skipped), then Post.  Operand of array X at (p, a):
//   c_k = sum_{d: P[d].id=k} (P[d].ofs + p_d P[d].stride)
//       + sum_{d: A[d].id=k} (A[d].ofs + [Loop stage] a_d A[d].stride)
// read from the row-major array with PAD (-> padding[j]) or clamp (WRAP)
// boundaries; outputs are written at the same affine address, unclamped,
// for every valid p (Loop outputs at every accumulation step).
// PtrNum nibble i = planes of array i (inputs, then outputs).  Expand and
// ZeroStrideOpt are accepted and validated; they are mapping hints in the
// reference and do not change results here.
//
// Differences by design: the call validates descriptors and aborts with a
// message (the reference aborts silently on bad ParallelDims and does not
// check runtime errors); warp_shuf = 0 is read as 1 (the reference aborts,
// which makes main/iimg.cu's second pass fail); any element size is accepted.
//
// Host side: descriptors are validated by libumi_hip (umi_validate_execute)
// and the kernel, instantiated here from the user's strategy struct, is
// launched through umi_launch_kernel.  Execute returns after the kernel
// completes (the reference is synchronous through its cudaFree);
// ExecuteAsync takes a stream and does not wait.
#pragma once
#include <hip/hip_runtime.h>

#include <cassert>  // the examples use assert() through this header, as with the reference's
#include <cstdio>
#include <cstdlib>

#include "UMI.h"
#include "umi_hip.h"

namespace UMI {
namespace engine {

// Affine map (p, a) -> physical coordinates plus the array's geometry.
struct ArrayMap {
    int32_t base[DIM];      // sum of view offsets per physical dim
    int32_t pc[DIM][DIM];   // pc[k][d]: coefficient of p_d in coordinate k
    int32_t ac[DIM][DIM];   // ac[k][d]: coefficient of a_d in coordinate k
    int32_t size[DIM];      // physical sizes
    int64_t stride[DIM];    // row-major element strides
    int32_t pad_mask;       // bit k: dim k uses PAD
};

template <class E>
struct InArray {
    ArrayMap m;
    const E *ptr[MAX_PTR];
    E pad[MAX_PTR];
};
template <class E>
struct OutArray {
    ArrayMap m;
    E *ptr[MAX_PTR];
};

template <class E, int NI, int NO>
struct Args {
    int32_t ptot[DIM];   // parallel totals
    int32_t atot[DIM];   // accumulation totals
    int32_t ablk[DIM];   // accumulation block sizes
    int32_t agrid[DIM];  // accumulation blocks per dim
    int64_t np;          // number of parallel points
    InArray<E> in[SafeArrayLength(NI)];
    OutArray<E> out[SafeArrayLength(NO)];
};

inline ArrayMap make_map(const NDTuple<ViewDim> &pv, const NDTuple<ViewDim> &av, const NDTuple<PhysicalDim> &phy) {
    ArrayMap m{};
    for (int d = 0; d < DIM; ++d) {
        m.base[pv[d].physical_dim_id] += pv[d].ofs;
        m.base[av[d].physical_dim_id] += av[d].ofs;
        m.pc[pv[d].physical_dim_id][d] += pv[d].stride;
        m.ac[av[d].physical_dim_id][d] += av[d].stride;
    }
    int64_t s = 1;
    for (int k = DIM - 1; k >= 0; --k) {
        m.size[k] = phy[k].size;
        m.stride[k] = s;
        s *= phy[k].size > 0 ? phy[k].size : 1;
        if (phy[k].mode == PhysicalDim::PAD) m.pad_mask |= 1 << k;
    }
    return m;
}

// Physical coordinates of (p, a) for map m; a contributes only in the Loop stage.
template <bool LOOP>
__device__ inline void coords(const ArrayMap &m, const int32_t (&p)[DIM], const int32_t (&a)[DIM], int32_t (&c)[DIM]) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        int32_t v = m.base[k];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            v += m.pc[k][d] * p[d];
            if (LOOP) v += m.ac[k][d] * a[d];
        }
        c[k] = v;
    }
}

template <class E>
__device__ inline E load(const InArray<E> &A, int j, const int32_t (&c)[DIM]) {
    int64_t lin = 0;
    bool outside = false;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        const int32_t hi = A.m.size[k] - 1;
        const int32_t cl = c[k] < 0 ? 0 : (c[k] > hi ? hi : c[k]);
        outside |= (cl != c[k]) && ((A.m.pad_mask >> k) & 1);
        lin += (int64_t)cl * A.m.stride[k];
    }
    return outside ? A.pad[j] : A.ptr[j][lin];
}

template <class E>
__device__ inline void store(const OutArray<E> &O, int j, const int32_t (&c)[DIM], const E &v) {
    int64_t lin = 0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) lin += (int64_t)c[k] * O.m.stride[k];
    O.ptr[j][lin] = v;
}

template <uint64_t PtrNum>
__host__ __device__ constexpr int planes(int array_index) {
    return (int)Slice(PtrNum, 4, array_index);
}

// Gathers the inputs of one stage: arrays [I0, I0+N) into in[].
template <bool LOOP, uint64_t PtrNum, int I0, int N, class E, int NI, int NO, size_t S>
__device__ inline void gather(const Args<E, NI, NO> &g, const int32_t (&p)[DIM], const int32_t (&a)[DIM],
                              simple_array<E, S> &in) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int32_t c[DIM];
        coords<LOOP>(g.in[I0 + i].m, p, a, c);
#pragma unroll
        for (int j = 0; j < planes<PtrNum>(I0 + i); ++j) in[i * MAX_PTR + j] = load(g.in[I0 + i], j, c);
    }
}

template <bool LOOP, uint64_t PtrNum, int TOTAL_IN, int O0, int N, class E, int NI, int NO, size_t S>
__device__ inline void scatter(const Args<E, NI, NO> &g, const int32_t (&p)[DIM], const int32_t (&a)[DIM],
                               const simple_array<E, S> &out) {
#pragma unroll
    for (int o = 0; o < N; ++o) {
        int32_t c[DIM];
        coords<LOOP>(g.out[O0 + o].m, p, a, c);
#pragma unroll
        for (int j = 0; j < planes<PtrNum>(TOTAL_IN + O0 + o); ++j) store(g.out[O0 + o], j, c, out[o * MAX_PTR + j]);
    }
}

template <class Umi, uint64_t PtrNum, class E, class C>
__global__ __launch_bounds__(256) void execute_kernel(const Args<E, Umi::TOTAL_IN, Umi::TOTAL_OUT> g, const C uc) {
    constexpr int PRE_I = 0, LOOP_I = Umi::N_PRE_IN, POST_I = LOOP_I + Umi::N_LOOP_IN;
    constexpr int PRE_O = 0, LOOP_O = Umi::N_PRE_OUT, POST_O = LOOP_O + Umi::N_LOOP_OUT;
    const int64_t flat = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (flat >= g.np) return;
    int32_t p[DIM];
    int64_t f = flat;
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) {
        p[d] = (int32_t)(f % g.ptot[d]);
        f /= g.ptot[d];
    }
    int32_t a[DIM] = {0, 0, 0, 0};
    Umi ctx;
    {
        simple_array<E, Umi::N_PRE_IN * MAX_PTR> in;
        simple_array<E, Umi::N_PRE_OUT * MAX_PTR> out;
        gather<false, PtrNum, PRE_I, Umi::N_PRE_IN>(g, p, a, in);
        ctx.Pre(uc, in, out);
        scatter<false, PtrNum, Umi::TOTAL_IN, PRE_O, Umi::N_PRE_OUT>(g, p, a, out);
    }
    // block-major raster walk of the accumulation space
    const int32_t nblk = g.agrid[0] * g.agrid[1] * g.agrid[2] * g.agrid[3];
    const int32_t bsz = g.ablk[0] * g.ablk[1] * g.ablk[2] * g.ablk[3];
    for (int32_t ab = 0; ab < nblk; ++ab) {
        int32_t abi[DIM];
        int32_t t = ab;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            abi[d] = t % g.agrid[d];
            t /= g.agrid[d];
        }
        for (int32_t as = 0; as < bsz; ++as) {
            int32_t s = as;
            bool ok = true;
#pragma unroll
            for (int d = DIM - 1; d >= 0; --d) {
                a[d] = abi[d] * g.ablk[d] + s % g.ablk[d];
                s /= g.ablk[d];
                ok &= a[d] < g.atot[d];
            }
            if (!ok) continue;
            simple_array<E, Umi::N_LOOP_IN * MAX_PTR> in;
            simple_array<E, Umi::N_LOOP_OUT * MAX_PTR> out;
            gather<true, PtrNum, LOOP_I, Umi::N_LOOP_IN>(g, p, a, in);
            ctx.Loop(uc, in, out);
            scatter<true, PtrNum, Umi::TOTAL_IN, LOOP_O, Umi::N_LOOP_OUT>(g, p, a, out);
        }
    }
    {
        simple_array<E, Umi::N_POST_IN * MAX_PTR> in;
        simple_array<E, Umi::N_POST_OUT * MAX_PTR> out;
        gather<false, PtrNum, POST_I, Umi::N_POST_IN>(g, p, a, in);
        ctx.Post(uc, in, out);
        scatter<false, PtrNum, Umi::TOTAL_IN, POST_O, Umi::N_POST_OUT>(g, p, a, out);
    }
}

inline void pack_dims(const NDTuple<ParallelDim> &p_dims, const NDTuple<AccumulationDim> &a_dims, int32_t (&pd)[16],
                      int32_t (&ad)[8]) {
    for (int d = 0; d < DIM; ++d) {
        pd[4 * d + 0] = p_dims[d].total_size;
        pd[4 * d + 1] = p_dims[d].block_size;
        pd[4 * d + 2] = p_dims[d].warp_size;
        pd[4 * d + 3] = p_dims[d].warp_shuf;
        ad[2 * d + 0] = a_dims[d].total_size;
        ad[2 * d + 1] = a_dims[d].block_size;
    }
}

inline void pack_view(const NDTuple<ViewDim> &v, int32_t *o) {
    for (int d = 0; d < DIM; ++d) {
        o[3 * d + 0] = v[d].physical_dim_id;
        o[3 * d + 1] = v[d].ofs;
        o[3 * d + 2] = v[d].stride;
    }
}

inline void pack_phys(const NDTuple<PhysicalDim> &v, int32_t *o) {
    for (int d = 0; d < DIM; ++d) {
        o[2 * d + 0] = v[d].size;
        o[2 * d + 1] = (int32_t)v[d].mode;
    }
}

}  // namespace engine

// Status-returning form of Execute: validates, launches on `stream`, does not wait.
template <class Umi, uint64_t PtrNum = 0x1111111111111111llu, uint64_t Expand = 0x1111, uint64_t ZeroStrideOpt = 0,
          class EType = typename remove_cv<typename Umi::ElementType>::type,
          class CType = typename remove_cv<typename Umi::UserConstantType>::type>
umi_status ExecuteAsync(umi_stream stream, const NDTuple<ParallelDim> &p_dims, const NDTuple<AccumulationDim> &a_dims,
                        const array<UnrolledI<EType>, Umi::TOTAL_IN> &imems,
                        const array<UnrolledO<EType>, Umi::TOTAL_OUT> &omems, const CType &c = CType()) {
    constexpr int NI = Umi::TOTAL_IN, NO = Umi::TOTAL_OUT;
    // descriptor validation in the library (ParallelDim rules of the reference, view ids, sizes)
    int32_t pd[16], ad[8];
    engine::pack_dims(p_dims, a_dims, pd, ad);
    int32_t views[(NI + NO) * 24 + 1], phys[(NI + NO) * 8 + 1];
    int32_t planes[NI + NO + 1];
    for (int i = 0; i < NI; ++i) {
        engine::pack_view(imems[i].p_view_dims, views + 24 * i);
        engine::pack_view(imems[i].a_view_dims, views + 24 * i + 12);
        engine::pack_phys(imems[i].physical_dims, phys + 8 * i);
        planes[i] = engine::planes<PtrNum>(i);
    }
    for (int o = 0; o < NO; ++o) {
        engine::pack_view(omems[o].p_view_dims, views + 24 * (NI + o));
        engine::pack_view(omems[o].a_view_dims, views + 24 * (NI + o) + 12);
        engine::pack_phys(omems[o].physical_dims, phys + 8 * (NI + o));
        planes[NI + o] = engine::planes<PtrNum>(NI + o);
    }
    umi_status st = umi_validate_execute(pd, ad, NI, NO, views, phys, planes, Expand, ZeroStrideOpt);
    if (st != UMI_OK) return st;

    engine::Args<EType, NI, NO> g{};
    int64_t np = 1;
    for (int d = 0; d < DIM; ++d) {
        g.ptot[d] = p_dims[d].total_size;
        g.atot[d] = a_dims[d].total_size;
        g.ablk[d] = a_dims[d].block_size;
        g.agrid[d] = g.atot[d] > 0 ? CeilDiv(g.atot[d], g.ablk[d]) : 0;
        np *= p_dims[d].total_size;
    }
    g.np = np;
    for (int i = 0; i < NI; ++i) {
        g.in[i].m = engine::make_map(imems[i].p_view_dims, imems[i].a_view_dims, imems[i].physical_dims);
        for (int j = 0; j < MAX_PTR; ++j) {
            g.in[i].ptr[j] = imems[i].ptr[j];
            g.in[i].pad[j] = imems[i].padding[j];
        }
    }
    for (int o = 0; o < NO; ++o) {
        g.out[o].m = engine::make_map(omems[o].p_view_dims, omems[o].a_view_dims, omems[o].physical_dims);
        for (int j = 0; j < MAX_PTR; ++j) g.out[o].ptr[j] = omems[o].ptr[j];
    }
    if (np <= 0) return UMI_OK;
    const int64_t blocks = (np + 255) / 256;
    if (blocks > 0x7fffffff) return UMI_ERR_UNSUPPORTED;
    void *args[] = {(void *)&g, (void *)&c};
    return umi_launch_kernel((const void *)&engine::execute_kernel<Umi, PtrNum, EType, CType>, (unsigned)blocks, 1, 1,
                             256, 1, 1, 0, args, stream);
}

// UMI::Execute — drop-in for src/UMI_gpu.h:409-623.  Runs on the default
// stream and returns when the kernel has finished; any failure prints the
// library's message and aborts (the reference aborts on invalid
// ParallelDims, src/UMI.cpp:59-65).
template <class Umi, uint64_t PtrNum = 0x1111111111111111llu, uint64_t Expand = 0x1111, uint64_t ZeroStrideOpt = 0,
          class EType = typename remove_cv<typename Umi::ElementType>::type,
          class CType = typename remove_cv<typename Umi::UserConstantType>::type>
void Execute(const NDTuple<ParallelDim> &p_dims, const NDTuple<AccumulationDim> &a_dims,
             const array<UnrolledI<EType>, Umi::TOTAL_IN> &imems, const array<UnrolledO<EType>, Umi::TOTAL_OUT> &omems,
             const CType &c = CType()) {
    umi_status st = ExecuteAsync<Umi, PtrNum, Expand, ZeroStrideOpt, EType, CType>(nullptr, p_dims, a_dims, imems,
                                                                                   omems, c);
    if (st == UMI_OK) st = umi_stream_synchronize(nullptr);
    if (st != UMI_OK) {
        std::fprintf(stderr, "UMI::Execute failed (status %d): %s\n", st, umi_last_error());
        std::abort();
    }
}

}  // namespace UMI
