// umi_interp.h — CPU interpreter of UMI::Execute semantics.
//
// TEST INFRASTRUCTURE ONLY (the checker of the generic GPU engine).  Runs a
// strategy struct on the host, point by point, exactly as the reference
// kernel orders the work (src/UMI_gpu.h:236-357, johnjohnlin/UMI):
//   for each parallel point p (raster, dim 3 fastest):
//     Umi ctx; Pre(c, in_pre(p)) -> write pre outputs at p
//     for ab in raster(ceil(U/V)): for as in raster(V):      (UMI.cpp:158-159)
//       a = ab*V + as; if all(a < U): Loop(c, in_loop(p, a)) -> write at (p, a)
//     Post(c, in_post(p)) -> write post outputs at p
// Coordinates of array X (views P, A): c_k = sum_{P[d].id=k}(P.ofs + p_d P.stride)
//   + sum_{A[d].id=k} A.ofs + [loop] sum_{A[d].id=k} a_d A.stride
//   (gbase + grid/eblock/warp/thread offsets, UMI_gpu.h:492-500,535-543).
// Reads clamp each coordinate to [0, size-1] and return padding[j] if a PAD
// dim was clamped (Int4::ClampedSum, src/UMI.h:205-218); the address is the
// row-major unfold of the clamped coordinates.  Writes use the unclamped
// unfold (Int4::Unfold, src/UMI.h:197-204).  Pointers are host pointers.
#pragma once
#include <cstdint>

#include "UMI.h"

namespace umi_oracle {
using namespace UMI;

template <class V>
inline void coords(const NDTuple<ViewDim> &pv, const NDTuple<ViewDim> &av, bool loop, const int32_t *p,
                   const int32_t *a, int64_t *c) {
    for (int k = 0; k < DIM; ++k) c[k] = 0;
    for (int d = 0; d < DIM; ++d) {
        c[pv[d].physical_dim_id] += pv[d].ofs + (int64_t)p[d] * pv[d].stride;
        c[av[d].physical_dim_id] += av[d].ofs;
        if (loop) c[av[d].physical_dim_id] += (int64_t)a[d] * av[d].stride;
    }
}

template <class T>
inline T read(const UnrolledI<T> &X, int j, const int64_t *c) {
    bool outside = false;
    int64_t lin = 0;
    for (int k = 0; k < DIM; ++k) {
        const int64_t hi = X.physical_dims[k].size - 1;
        const int64_t cl = c[k] < 0 ? 0 : (c[k] > hi ? hi : c[k]);
        if (cl != c[k] && X.physical_dims[k].mode == PhysicalDim::PAD) outside = true;
        lin = lin * X.physical_dims[k].size + cl;
    }
    return outside ? X.padding[j] : X.ptr[j][lin];
}

template <class T>
inline void write(const UnrolledO<T> &X, int j, const int64_t *c, const T &v) {
    int64_t lin = 0;
    for (int k = 0; k < DIM; ++k) lin = lin * X.physical_dims[k].size + c[k];
    X.ptr[j][lin] = v;
}

template <class Umi, uint64_t PtrNum = 0x1111111111111111llu,
          class EType = typename std::remove_cv<typename Umi::ElementType>::type,
          class CType = typename std::remove_cv<typename Umi::UserConstantType>::type>
void Execute(const NDTuple<ParallelDim> &p_dims, const NDTuple<AccumulationDim> &a_dims,
             const std::array<UnrolledI<EType>, Umi::TOTAL_IN> &imems,
             const std::array<UnrolledO<EType>, Umi::TOTAL_OUT> &omems, const CType &uc = CType()) {
    const int PRE_I = 0, LOOP_I = Umi::N_PRE_IN, POST_I = LOOP_I + Umi::N_LOOP_IN;
    const int PRE_O = 0, LOOP_O = Umi::N_PRE_OUT, POST_O = LOOP_O + Umi::N_LOOP_OUT;
    auto nptr = [](int array_index) { return (int)((PtrNum >> (4 * array_index)) & 0xf); };
    int32_t P[DIM], U[DIM], V[DIM], G[DIM];
    int64_t np = 1;
    for (int d = 0; d < DIM; ++d) {
        P[d] = p_dims[d].total_size;
        U[d] = a_dims[d].total_size;
        V[d] = a_dims[d].block_size;
        G[d] = (U[d] + V[d] - 1) / V[d];
        np *= P[d];
    }
    const int64_t ng = (int64_t)G[0] * G[1] * G[2] * G[3], nv = (int64_t)V[0] * V[1] * V[2] * V[3];
    for (int64_t f = 0; f < np; ++f) {
        int32_t p[DIM], a[DIM] = {0, 0, 0, 0};
        int64_t t = f;
        for (int d = DIM - 1; d >= 0; --d) {
            p[d] = (int32_t)(t % P[d]);
            t /= P[d];
        }
        auto gather = [&](int i0, int n, bool loop, EType *in) {
            for (int i = 0; i < n; ++i) {
                int64_t c[DIM];
                coords<void>(imems[i0 + i].p_view_dims, imems[i0 + i].a_view_dims, loop, p, a, c);
                for (int j = 0; j < nptr(i0 + i); ++j) in[i * MAX_PTR + j] = read(imems[i0 + i], j, c);
            }
        };
        auto scatter = [&](int o0, int n, bool loop, const EType *out) {
            for (int o = 0; o < n; ++o) {
                int64_t c[DIM];
                coords<void>(omems[o0 + o].p_view_dims, omems[o0 + o].a_view_dims, loop, p, a, c);
                for (int j = 0; j < nptr(Umi::TOTAL_IN + o0 + o); ++j) write(omems[o0 + o], j, c, out[o * MAX_PTR + j]);
            }
        };
        Umi ctx;
        {
            simple_array<EType, Umi::N_PRE_IN * MAX_PTR> in;
            simple_array<EType, Umi::N_PRE_OUT * MAX_PTR> out;
            gather(PRE_I, Umi::N_PRE_IN, false, in.data);
            ctx.Pre(uc, in, out);
            scatter(PRE_O, Umi::N_PRE_OUT, false, out.data);
        }
        for (int64_t ab = 0; ab < ng; ++ab) {
            int32_t abi[DIM];
            int64_t s = ab;
            for (int d = DIM - 1; d >= 0; --d) {
                abi[d] = (int32_t)(s % G[d]);
                s /= G[d];
            }
            for (int64_t as = 0; as < nv; ++as) {
                int64_t r = as;
                bool ok = true;
                for (int d = DIM - 1; d >= 0; --d) {
                    a[d] = abi[d] * V[d] + (int32_t)(r % V[d]);
                    r /= V[d];
                    ok = ok && a[d] < U[d];
                }
                if (!ok) continue;
                simple_array<EType, Umi::N_LOOP_IN * MAX_PTR> in;
                simple_array<EType, Umi::N_LOOP_OUT * MAX_PTR> out;
                gather(LOOP_I, Umi::N_LOOP_IN, true, in.data);
                ctx.Loop(uc, in, out);
                scatter(LOOP_O, Umi::N_LOOP_OUT, true, out.data);
            }
        }
        {
            simple_array<EType, Umi::N_POST_IN * MAX_PTR> in;
            simple_array<EType, Umi::N_POST_OUT * MAX_PTR> out;
            gather(POST_I, Umi::N_POST_IN, false, in.data);
            ctx.Post(uc, in, out);
            scatter(POST_O, Umi::N_POST_OUT, false, out.data);
        }
    }
}

}  // namespace umi_oracle
