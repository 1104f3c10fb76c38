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
// PtrNum nibble i = planes of array i (inputs, then outputs).
//
// The kernel (the role of Execute_gpukernel, src/UMI_gpu.h:189-358):
//  * a workgroup of prod(block_size) threads owns an "eblock" of
//    Expand_d x block_size_d points per dim; a thread owns the Expand_d
//    points t_d + e * block_size_d (interleaved, so neighbouring threads take
//    neighbouring points: conflict-free LDS reads, coalesced writes) and
//    keeps one strategy object per point in registers (CoreExecution's
//    umi_ctx, src/UMI_gpu.h:101-187);
//  * per accumulation block, every Loop input whose footprint pays for it
//    is staged in LDS (FillShm, src/UMI_gpu.h:33-83): the footprint of an
//    affine view over the eblock x accumulation block is a box of the same
//    size for every workgroup; the staging applies PAD / clamp once per
//    element, so the strategy's operands become plain LDS reads (other inputs
//    and the Pre / Post inputs are read from device memory with the boundary
//    rule);
//  * ZeroStrideOpt bit 4 i + d shares input i's operands across the Expand
//    points along dim d (one load for all of them); the library checks that
//    the input's view does not move along dim d, so a hint cannot change
//    results (the reference would read wrong data).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cassert>  // the examples use assert() through this header, as with the reference's
#include <cstdio>
#include <cstdlib>
#include <utility>

#include "UMI.h"
#include "umi_hip.h"

namespace UMI {
namespace engine {

constexpr int MAX_THREADS = 1024;              // threads per workgroup (MAX_BLOCK_SIZE)
constexpr size_t LDS_BUDGET = 64 * 1024;       // bytes of staged operands per workgroup

// Affine map (p, a) -> physical coordinates plus the array's geometry.
struct ArrayMap {
    int32_t base[DIM];      // sum of view offsets per physical dim
    int32_t pc[DIM][DIM];   // pc[k][d]: coefficient of p_d in coordinate k
    int32_t ac[DIM][DIM];   // ac[k][d]: coefficient of a_d in coordinate k
    int32_t size[DIM];      // physical sizes
    int64_t stride[DIM];    // row-major element strides
    int32_t pad_mask;       // bit k: dim k uses PAD
};

// LDS staging of a Loop input: the footprint box of extent ext over the
// eblock x accumulation block, planes one after another from lds_ofs.
struct Staging {
    int32_t on;             // 1: staged
    int32_t lds_ofs;        // element offset of plane 0
    int32_t box;            // elements per plane
    int32_t ext[DIM];       // box extent per physical dim
    int32_t lst[DIM];       // LDS stride per physical dim
    int32_t pcl[DIM];       // LDS-linear coefficient of p_d
    int32_t acl[DIM];       // LDS-linear coefficient of a_d
};

template <class E>
struct InArray {
    ArrayMap m;
    Staging st;
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
    int32_t pblk[DIM];   // threads of a workgroup per dim (block_size)
    int32_t pgrid[DIM];  // workgroups per dim
    int32_t atot[DIM];   // accumulation totals
    int32_t ablk[DIM];   // accumulation block sizes
    int32_t agrid[DIM];  // accumulation blocks per dim
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

// Footprint box of map m over eblock extents eb and accumulation block ab.
inline Staging make_staging(const ArrayMap &m, const int32_t (&eb)[DIM], const int32_t (&ab)[DIM]) {
    Staging st{};
    int64_t box = 1;
    for (int k = 0; k < DIM; ++k) {
        int64_t e = 1;
        for (int d = 0; d < DIM; ++d) {
            e += (int64_t)(m.pc[k][d] < 0 ? -m.pc[k][d] : m.pc[k][d]) * (eb[d] - 1);
            e += (int64_t)(m.ac[k][d] < 0 ? -m.ac[k][d] : m.ac[k][d]) * (ab[d] - 1);
        }
        st.ext[k] = (int32_t)(e < (1 << 30) ? e : (1 << 30));
        box *= st.ext[k];
        if (box > (1 << 30)) box = 1 << 30;
    }
    st.box = (int32_t)box;
    int32_t s = 1;
    for (int k = DIM - 1; k >= 0; --k) {
        st.lst[k] = s;
        s = (int32_t)std::min<int64_t>((int64_t)s * st.ext[k], 1 << 30);
    }
    for (int d = 0; d < DIM; ++d) {
        int64_t pl = 0, al = 0;
        for (int k = 0; k < DIM; ++k) {
            pl += (int64_t)m.pc[k][d] * st.lst[k];
            al += (int64_t)m.ac[k][d] * st.lst[k];
        }
        st.pcl[d] = (int32_t)pl;
        st.acl[d] = (int32_t)al;
    }
    return st;
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

// Expand nibbles (dim d: bits 12 - 4 d .. 15 - 4 d, src/UMI_gpu.h:128-131).
template <uint64_t Expand>
struct Ex {
    static constexpr int e[DIM] = {(int)Slice(Expand, 4, 3), (int)Slice(Expand, 4, 2), (int)Slice(Expand, 4, 1),
                                   (int)Slice(Expand, 4, 0)};
    static constexpr int size = e[0] * e[1] * e[2] * e[3];
    // expansion index -> its coordinate along dim d (dim 3 fastest)
    __host__ __device__ static constexpr int coord(int id, int d) {
        int r = id;
        for (int k = DIM - 1; k > d; --k) r /= e[k];
        return r % e[d];
    }
    // the expansion whose operands input i shares (ZeroStrideOpt bit 4 i + d zeroes dim d)
    template <uint64_t Opt>
    __host__ __device__ static constexpr int canon(int i, int id) {
        int c = 0;
        for (int d = 0; d < DIM; ++d) c = c * e[d] + (Slice(Opt, 1, DIM * i + d) ? 0 : coord(id, d));
        return c;
    }
};

// Lowest coordinate of the footprint box of map m for eblock origin P0 and
// accumulation block origin A0 (Loop stage).
__device__ inline void box_min(const ArrayMap &m, const int32_t (&P0)[DIM], const int32_t (&eb)[DIM],
                               const int32_t (&A0)[DIM], const int32_t (&ab)[DIM], int32_t (&mn)[DIM]) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        int32_t v = m.base[k];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int32_t pc = m.pc[k][d], ac = m.ac[k][d];
            v += pc >= 0 ? pc * P0[d] : pc * (P0[d] + eb[d] - 1);
            v += ac >= 0 ? ac * A0[d] : ac * (A0[d] + ab[d] - 1);
        }
        mn[k] = v;
    }
}

// Stages input A's footprint box (all its planes) into lds; the whole
// workgroup, coalesced along the innermost physical dim.
template <class E>
__device__ inline void stage(const InArray<E> &A, int npl, E *lds, const int32_t (&mn)[DIM], int nthreads) {
    const Staging &st = A.st;
    const int32_t e1 = st.ext[1], e2 = st.ext[2], e3 = st.ext[3];
    // (row, column) of element idx = row * e3 + column, advanced by nthreads
    // without dividing: q whole rows and r columns per step
    const int32_t q = nthreads / e3, r = nthreads - q * e3;
    int32_t row = (int32_t)threadIdx.x / e3, col = (int32_t)threadIdx.x - row * e3;
    const bool flat = e1 * st.ext[0] == 1;  // rows are dim 2 alone (2-D boxes)
    for (int32_t idx = (int32_t)threadIdx.x; idx < st.box; idx += nthreads) {
        int32_t c[DIM];
        c[3] = mn[3] + col;
        if (flat) {
            c[2] = mn[2] + row;
            c[1] = mn[1];
            c[0] = mn[0];
        } else {
            const int32_t r12 = row / e2;
            c[2] = mn[2] + (row - r12 * e2);
            c[1] = mn[1] + r12 % e1;
            c[0] = mn[0] + r12 / e1;
        }
        for (int j = 0; j < npl; ++j) lds[st.lds_ofs + j * st.box + idx] = load(A, j, c);
        row += q;
        col += r;
        if (col >= e3) {
            col -= e3;
            ++row;
        }
    }
}

// Operands of one stage for all Expand points: buf[i][plane][expansion],
// loaded once per shared (ZeroStrideOpt) group.
// LDS offsets of the staged Loop inputs: per Expand point (once per
// accumulation block) and per step (the same for the whole workgroup).
template <int N, int ES>
struct LdsOffsets {
    int32_t pt[SafeArrayLength(N)][ES];
    int32_t step[SafeArrayLength(N)];
};

template <bool LOOP, uint64_t PtrNum, uint64_t Expand, uint64_t Opt, int I0, int N, class E, int NI, int NO, class LO>
__device__ inline void gather_all(const Args<E, NI, NO> &g, const E *lds, const LO &lo, const int32_t (&pt)[DIM],
                                  const int32_t (&a)[DIM], E (&buf)[SafeArrayLength(N)][MAX_PTR][Ex<Expand>::size]) {
    using X = Ex<Expand>;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const InArray<E> &A = g.in[I0 + i];
#pragma unroll
        for (int id = 0; id < X::size; ++id) {
            if (X::template canon<Opt>(I0 + i, id) != id) continue;  // shared with an earlier point
            if (LOOP && A.st.on) {
                const int32_t off = lo.pt[i][id] + lo.step[i];
#pragma unroll
                for (int j = 0; j < planes<PtrNum>(I0 + i); ++j) buf[i][j][id] = lds[off + j * A.st.box];
            } else {
                int32_t p[DIM], c[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) p[d] = pt[d] + X::coord(id, d) * g.pblk[d];
                coords<LOOP>(A.m, p, a, c);
#pragma unroll
                for (int j = 0; j < planes<PtrNum>(I0 + i); ++j) buf[i][j][id] = load(A, j, c);
            }
        }
    }
}

// Runs one stage (0 Pre, 1 Loop, 2 Post) for every Expand point of the thread.
template <int STAGE, uint64_t PtrNum, uint64_t Expand, uint64_t Opt, int I0, int NIN, int O0, int NOUT, int TOTAL_IN,
          class Umi, class E, class C, int NI, int NO, class LO>
__device__ inline void run_stage(const Args<E, NI, NO> &g, const E *lds, const LO &lo, const int32_t (&pt)[DIM],
                                 const int32_t (&a)[DIM], const bool (&valid)[Ex<Expand>::size],
                                 Umi (&ctx)[Ex<Expand>::size], const C &uc) {
    using X = Ex<Expand>;
    E buf[SafeArrayLength(NIN)][MAX_PTR][X::size];
    gather_all<STAGE == 1, PtrNum, Expand, Opt, I0, NIN>(g, lds, lo, pt, a, buf);
#pragma unroll
    for (int id = 0; id < X::size; ++id) {
        simple_array<E, NIN * MAX_PTR> in;
        simple_array<E, NOUT * MAX_PTR> out;
#pragma unroll
        for (int i = 0; i < NIN; ++i) {
            const int src = X::template canon<Opt>(I0 + i, id);
#pragma unroll
            for (int j = 0; j < planes<PtrNum>(I0 + i); ++j) in[i * MAX_PTR + j] = buf[i][j][src];
        }
        if constexpr (STAGE == 0)
            ctx[id].Pre(uc, in, out);
        else if constexpr (STAGE == 1)
            ctx[id].Loop(uc, in, out);
        else
            ctx[id].Post(uc, in, out);
        if (NOUT > 0 && valid[id]) {
            int32_t p[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) p[d] = pt[d] + X::coord(id, d) * g.pblk[d];
#pragma unroll
            for (int o = 0; o < NOUT; ++o) {
                int32_t c[DIM];
                coords<STAGE == 1>(g.out[O0 + o].m, p, a, c);
#pragma unroll
                for (int j = 0; j < planes<PtrNum>(TOTAL_IN + O0 + o); ++j)
                    store(g.out[O0 + o], j, c, out[o * MAX_PTR + j]);
            }
        }
    }
}

template <class Umi, uint64_t PtrNum, uint64_t Expand, uint64_t Opt, class E, class C, int MAXT>
__global__ __launch_bounds__(MAXT) void execute_kernel(const Args<E, Umi::TOTAL_IN, Umi::TOTAL_OUT> g, const C uc) {
    constexpr int NI = Umi::TOTAL_IN;
    constexpr int PRE_I = 0, LOOP_I = Umi::N_PRE_IN, POST_I = LOOP_I + Umi::N_LOOP_IN;
    constexpr int PRE_O = 0, LOOP_O = Umi::N_PRE_OUT, POST_O = LOOP_O + Umi::N_LOOP_OUT;
    using X = Ex<Expand>;
    extern __shared__ __attribute__((aligned(16))) char engine_lds[];
    E *lds = reinterpret_cast<E *>(engine_lds);
    const int nthreads = (int)blockDim.x;
    // workgroup origin, thread offset, eblock extents
    int32_t P0[DIM], pt[DIM], eb[DIM];
    {
        int32_t b = (int32_t)blockIdx.x, t = (int32_t)threadIdx.x;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            eb[d] = X::e[d] * g.pblk[d];
            P0[d] = (b % g.pgrid[d]) * eb[d];
            b /= g.pgrid[d];
            pt[d] = P0[d] + t % g.pblk[d];
            t /= g.pblk[d];
        }
    }
    bool valid[X::size];
#pragma unroll
    for (int id = 0; id < X::size; ++id) {
        bool v = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d) v &= pt[d] + X::coord(id, d) * g.pblk[d] < g.ptot[d];
        valid[id] = v;
    }
    int32_t a[DIM] = {0, 0, 0, 0};
    LdsOffsets<Umi::N_LOOP_IN, X::size> lo{};
    Umi ctx[X::size];
    run_stage<0, PtrNum, Expand, Opt, PRE_I, Umi::N_PRE_IN, PRE_O, Umi::N_PRE_OUT, NI>(g, lds, lo, pt, a, valid, ctx,
                                                                                      uc);
    // Loop: block-major raster walk of the accumulation space
    bool any_staged = false;
#pragma unroll
    for (int i = 0; i < Umi::N_LOOP_IN; ++i) any_staged |= g.in[LOOP_I + i].st.on != 0;
    const int32_t nblk = g.agrid[0] * g.agrid[1] * g.agrid[2] * g.agrid[3];
    const int32_t bsz = g.ablk[0] * g.ablk[1] * g.ablk[2] * g.ablk[3];
    for (int32_t ab = 0; ab < nblk; ++ab) {
        int32_t A0[DIM];
        {
            int32_t r = ab;
#pragma unroll
            for (int d = DIM - 1; d >= 0; --d) {
                A0[d] = (r % g.agrid[d]) * g.ablk[d];
                r /= g.agrid[d];
            }
        }
        if (any_staged) {
            __syncthreads();  // the previous block's operands are no longer read
#pragma unroll
            for (int i = 0; i < Umi::N_LOOP_IN; ++i) {
                const InArray<E> &A = g.in[LOOP_I + i];
                if (!A.st.on) continue;
                int32_t mn[DIM];
                box_min(A.m, P0, eb, A0, g.ablk, mn);
                stage<E>(A, planes<PtrNum>(LOOP_I + i), lds, mn, nthreads);
                // LDS index of (thread's first point, a = 0) = sum_k (c_k - mn_k) lst_k
                int32_t c[DIM];
                coords<false>(A.m, pt, a, c);  // a = 0 here: the Loop term is added per step
                int32_t off = A.st.lds_ofs;
#pragma unroll
                for (int k = 0; k < DIM; ++k) off += (c[k] - mn[k]) * A.st.lst[k];
#pragma unroll
                for (int id = 0; id < X::size; ++id) {
                    int32_t o = off;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) o += A.st.pcl[d] * X::coord(id, d) * g.pblk[d];
                    lo.pt[i][id] = o;
                }
            }
            __syncthreads();
        }
        // steps of the block in raster order (dim 3 fastest), an odometer
        // instead of dividing the step index
        int32_t l[DIM] = {0, 0, 0, 0};
        for (int32_t as = 0; as < bsz; ++as) {
            bool ok = true;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                a[d] = A0[d] + l[d];
                ok &= a[d] < g.atot[d];
            }
#pragma unroll
            for (int i = 0; i < Umi::N_LOOP_IN; ++i) {
                const Staging &st = g.in[LOOP_I + i].st;
                lo.step[i] = st.acl[0] * a[0] + st.acl[1] * a[1] + st.acl[2] * a[2] + st.acl[3] * a[3];
            }
            if (ok)
                run_stage<1, PtrNum, Expand, Opt, LOOP_I, Umi::N_LOOP_IN, LOOP_O, Umi::N_LOOP_OUT, NI>(g, lds, lo, pt, a,
                                                                                                      valid, ctx, uc);
            if (++l[3] == g.ablk[3]) {
                l[3] = 0;
                if (++l[2] == g.ablk[2]) {
                    l[2] = 0;
                    if (++l[1] == g.ablk[1]) {
                        l[1] = 0;
                        ++l[0];
                    }
                }
            }
        }
    }
    for (int d = 0; d < DIM; ++d) a[d] = 0;
    run_stage<2, PtrNum, Expand, Opt, POST_I, Umi::N_POST_IN, POST_O, Umi::N_POST_OUT, NI>(g, lds, lo, pt, a, valid,
                                                                                          ctx, uc);
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
}  // namespace UMI

#include "UMI_route.h"

namespace UMI {

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
    // a call a specialised kernel computes exactly goes to it (UMI_route.h)
    if (route::enabled() && route::try_stencil<Umi, PtrNum, EType, CType>(stream, p_dims, a_dims, imems, omems, c, st))
        return st;

    using X = engine::Ex<Expand>;
    engine::Args<EType, NI, NO> g{};
    int64_t nwg = 1, nthreads = 1;
    int32_t eb[DIM];
    bool empty = false;
    for (int d = 0; d < DIM; ++d) {
        g.ptot[d] = p_dims[d].total_size;
        g.pblk[d] = p_dims[d].block_size;
        eb[d] = X::e[d] * g.pblk[d];
        g.pgrid[d] = CeilDiv(g.ptot[d], eb[d]);
        g.atot[d] = a_dims[d].total_size;
        g.ablk[d] = a_dims[d].block_size;
        g.agrid[d] = g.atot[d] > 0 ? CeilDiv(g.atot[d], g.ablk[d]) : 0;
        nwg *= g.pgrid[d];
        nthreads *= g.pblk[d];
        empty |= g.ptot[d] <= 0;
    }
    if (nthreads > engine::MAX_THREADS || nwg > 0x7fffffff) return UMI_ERR_UNSUPPORTED;
    for (int i = 0; i < NI; ++i) {
        g.in[i].m = engine::make_map(imems[i].p_view_dims, imems[i].a_view_dims, imems[i].physical_dims);
        for (int j = 0; j < MAX_PTR; ++j) {
            g.in[i].ptr[j] = imems[i].ptr[j];
            g.in[i].pad[j] = imems[i].padding[j];
        }
    }
    // LDS staging of Loop inputs: in order, while the footprint is smaller
    // than the reads it replaces (operands are reused) and the budget holds
    {
        int64_t reads = 1, used = 0;
        for (int d = 0; d < DIM; ++d) reads *= (int64_t)eb[d] * g.ablk[d];
        for (int i = Umi::N_PRE_IN; i < Umi::N_PRE_IN + Umi::N_LOOP_IN; ++i) {
            const int npl = engine::planes<PtrNum>(i);
            if (npl == 0) continue;
            engine::Staging st = engine::make_staging(g.in[i].m, eb, g.ablk);
            const int64_t bytes = (int64_t)st.box * npl * (int64_t)sizeof(EType);
            if (st.box < reads && used + bytes <= (int64_t)engine::LDS_BUDGET) {
                st.on = 1;
                st.lds_ofs = (int32_t)(used / (int64_t)sizeof(EType));
                used += (bytes + 15) / 16 * 16;
                g.in[i].st = st;
            }
        }
        for (int o = 0; o < NO; ++o) {
            g.out[o].m = engine::make_map(omems[o].p_view_dims, omems[o].a_view_dims, omems[o].physical_dims);
            for (int j = 0; j < MAX_PTR; ++j) g.out[o].ptr[j] = omems[o].ptr[j];
        }
        if (empty || nwg == 0) return UMI_OK;
        void *args[] = {(void *)&g, (void *)&c};
        const void *fn = nthreads <= 256
                             ? (const void *)&engine::execute_kernel<Umi, PtrNum, Expand, ZeroStrideOpt, EType, CType, 256>
                             : (const void *)&engine::execute_kernel<Umi, PtrNum, Expand, ZeroStrideOpt, EType, CType, 1024>;
        return umi_launch_kernel(fn, (unsigned)nwg, 1, 1, (unsigned)nthreads, 1, 1, (size_t)used, args, stream);
    }
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

// ---------------------------------------------------------------- tuning
// Autotuner over the template parameters of a UMI program (no reference
// counterpart; the reference's mapping is whatever the caller wrote into
// Expand and the ParallelDim block / warp sizes, src/UMI_gpu.h:409-423,
// src/UMI.cpp:39-66).  Candidates: every Expand of the list, crossed with
// workgroup shapes for the two innermost parallel dims (the caller's own
// first, then 256 threads as 16 x 16, 8 x 32, 32 x 8, 4 x 64, 1 x 256 and 128
// threads as 8 x 16, 4 x 32), each with warp sizes that keep the reference's
// rules (product of warp sizes 32, block a multiple of warp x shuf); outer
// dims keep the caller's values.  A candidate the validator refuses (for
// instance a ZeroStrideOpt hint that is unsound under a larger Expand) is
// skipped.  Each is timed with events over `reps` launches on `stream`
// after one warm-up; results never depend on the choice (the engine's
// mapping does not change results), so ExecuteTuned may run any of them.
struct TuneResult {
    int expand_index = -1;             // position in the Expand list
    uint64_t expand = 0;
    NDTuple<ParallelDim> p_dims{};     // block / warp sizes of the fastest candidate
    float us = 0;                      // its time per launch
    int candidates = 0;                // candidates timed
};

namespace engine {
inline bool shape_pd(const NDTuple<ParallelDim> &base, int b2, int b3, NDTuple<ParallelDim> &out) {
    out = base;
    const int w3 = b3 < 32 ? b3 : 32, w2 = 32 / w3;
    if (32 % w3 != 0 || b2 % w2 != 0) return false;
    for (int d = 0; d < 2; ++d) out[d] = ParallelDim{base[d].total_size, 1, 1, 1};
    out[2] = ParallelDim{base[2].total_size, b2, w2, 1};
    out[3] = ParallelDim{base[3].total_size, b3, w3, 1};
    return true;
}
template <class F>
inline float time_launches(F &&launch, umi_stream stream, int reps) {
    if (launch() != UMI_OK) return -1.f;  // warm-up; a refused candidate is skipped
    umi_timer t;
    if (umi_timer_create(&t) != UMI_OK) return -1.f;
    umi_timer_start(t, stream);
    for (int i = 0; i < reps; ++i) launch();
    umi_timer_stop(t, stream);
    float ms = -1.f;
    umi_timer_elapsed_ms(t, &ms);
    umi_timer_destroy(t);
    return ms * 1e3f / reps;
}
}  // namespace engine

template <class Umi, uint64_t PtrNum, uint64_t ZeroStrideOpt, uint64_t... Expands>
struct Tuner {
    using EType = typename remove_cv<typename Umi::ElementType>::type;
    using CType = typename remove_cv<typename Umi::UserConstantType>::type;
    static constexpr uint64_t list[sizeof...(Expands)] = {Expands...};

    template <size_t... I>
    static umi_status run_index(int idx, umi_stream s, const NDTuple<ParallelDim> &pd, const NDTuple<AccumulationDim> &ad,
                                const array<UnrolledI<EType>, Umi::TOTAL_IN> &in,
                                const array<UnrolledO<EType>, Umi::TOTAL_OUT> &out, const CType &c,
                                std::index_sequence<I...>) {
        umi_status st = UMI_ERR_INVALID;
        ((idx == (int)I ? (st = ExecuteAsync<Umi, PtrNum, list[I], ZeroStrideOpt>(s, pd, ad, in, out, c), 0) : 0), ...);
        return st;
    }
    static umi_status run(int idx, umi_stream s, const NDTuple<ParallelDim> &pd, const NDTuple<AccumulationDim> &ad,
                          const array<UnrolledI<EType>, Umi::TOTAL_IN> &in,
                          const array<UnrolledO<EType>, Umi::TOTAL_OUT> &out, const CType &c) {
        return run_index(idx, s, pd, ad, in, out, c, std::make_index_sequence<sizeof...(Expands)>{});
    }
};

// Times the candidates and returns the fastest; the outputs hold the results
// of the last launch (identical for every candidate).
template <class Umi, uint64_t PtrNum, uint64_t ZeroStrideOpt, uint64_t... Expands>
TuneResult Autotune(umi_stream stream, const NDTuple<ParallelDim> &p_dims, const NDTuple<AccumulationDim> &a_dims,
                    const array<UnrolledI<typename remove_cv<typename Umi::ElementType>::type>, Umi::TOTAL_IN> &imems,
                    const array<UnrolledO<typename remove_cv<typename Umi::ElementType>::type>, Umi::TOTAL_OUT> &omems,
                    const typename remove_cv<typename Umi::UserConstantType>::type &c =
                        typename remove_cv<typename Umi::UserConstantType>::type(),
                    int reps = 10) {
    using T = Tuner<Umi, PtrNum, ZeroStrideOpt, Expands...>;
    static const int shapes[][2] = {{0, 0}, {16, 16}, {8, 32}, {32, 8}, {4, 64}, {1, 256}, {8, 16}, {4, 32}};
    TuneResult best;
    for (int e = 0; e < (int)sizeof...(Expands); ++e)
        for (const auto &sh : shapes) {
            NDTuple<ParallelDim> pd = p_dims;
            if (sh[0] && !engine::shape_pd(p_dims, sh[0], sh[1], pd)) continue;
            const float us = engine::time_launches([&] { return T::run(e, stream, pd, a_dims, imems, omems, c); },
                                                   stream, reps);
            if (us < 0) continue;
            ++best.candidates;
            if (best.expand_index < 0 || us < best.us) {
                best.expand_index = e;
                best.expand = T::list[e];
                best.p_dims = pd;
                best.us = us;
            }
        }
    return best;
}

// Runs the program with the mapping Autotune picked (same Expand list).
template <class Umi, uint64_t PtrNum, uint64_t ZeroStrideOpt, uint64_t... Expands>
umi_status ExecuteTuned(const TuneResult &t, umi_stream stream, const NDTuple<AccumulationDim> &a_dims,
                        const array<UnrolledI<typename remove_cv<typename Umi::ElementType>::type>, Umi::TOTAL_IN> &imems,
                        const array<UnrolledO<typename remove_cv<typename Umi::ElementType>::type>, Umi::TOTAL_OUT> &omems,
                        const typename remove_cv<typename Umi::UserConstantType>::type &c =
                            typename remove_cv<typename Umi::UserConstantType>::type()) {
    if (t.expand_index < 0) return UMI_ERR_INVALID;
    return Tuner<Umi, PtrNum, ZeroStrideOpt, Expands...>::run(t.expand_index, stream, t.p_dims, a_dims, imems, omems,
                                                               c);
}

}  // namespace UMI
