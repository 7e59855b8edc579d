// Deterministic reductions shared by the kernels: the block tree, and the
// second stage over the first-stage partials with what it does with the sum
// (the PCG scalars of IterState).
#pragma once

#include "common.hpp"

namespace mf3d {

// sum over a workgroup of BS threads (wave64 shuffle tree, then the waves'
// sums in LDS); the result is valid in thread 0
template <int BS>
__device__ inline double block_sum(double v) {
    __shared__ double part[BS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (w == 0) {
        v = lane < BS / 64 ? part[lane] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    }
    return v;
}

// sum over a workgroup of any size that is a multiple of 64 (at most 1024);
// the result is valid in thread 0
__device__ inline double block_sum_dyn(double v) {
    __shared__ double part[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (w == 0) {
        v = lane < (int)(blockDim.x >> 6) ? part[lane] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    }
    return v;
}

// sum of partials[0 .. np) by one workgroup of kFinalBlock threads in a fixed
// order (16 independent loads per thread and round; the partials sit in L2);
// valid in thread 0
__device__ inline double finalize_sum(const double* __restrict__ partials, unsigned np) {
    constexpr int U = 16;
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = 0.0;
    unsigned i = threadIdx.x;
    for (; i + (U - 1) * kFinalBlock < np; i += U * kFinalBlock) {
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] += partials[i + u * kFinalBlock];
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u) {  // the remaining < U rounds
        const unsigned t = i + u * kFinalBlock;
        if (t < np) v[u] += partials[t];
    }
#pragma unroll
    for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
        for (int u = 0; u < w; ++u) v[u] += v[u + w];
    return block_sum<kFinalBlock>(v[0]);
}

// what the second stage does with the sum S (thread 0 only)
__device__ inline void finalize_apply(int op, double S, IterState* __restrict__ st,
                                      double* __restrict__ out) {
    switch (op) {
    case FIN_STORE: *out = S; break;
    case FIN_SQRT: *out = sqrt(S); break;
    case FIN_BNORM: st->bnorm = sqrt(S); break;
    case FIN_R0: // pcg.cuh:113-127
        st->r0 = sqrt(S);
        st->rr = S;
        st->den = st->stop_norm == MF3D_STOP_RHS_NORM ? st->bnorm : st->r0;
        if (st->r0 < 1e-15) st->stop = 1;
        break;
    case FIN_RZ0: st->rz = S; break;
    case FIN_PAP: // pcg.cuh:146-147
        if (st->stop) break;
        st->pAp = S;
        st->alpha = st->rz / S;
        if (!isfinite(st->alpha)) st->stop = 2;
        break;
    case FIN_RR: // pcg.cuh:156-164
        if (st->stop) {
            st->xa = 0.0;
            break;
        }
        st->xpend = 1;  // r is updated; x follows with the p update (or pcg_x_pending)
        st->xa = st->alpha;  // ... or with pcg_x_step on a second stream
        st->rr = S;
        st->iter += 1;
        if (st->iter % st->check_every == 0 || st->iter >= st->max_iter) {
            if (sqrt(S) / st->den < st->rtol)
                st->stop = 1;
            else if (st->iter >= st->max_iter)
                st->stop = 3;
        }
        break;
    case FIN_RZ: // pcg.cuh:170-177 (flexible: Polak-Ribiere numerator)
        if (st->stop) break;
        st->xpend = 0;  // the p update that follows also updates x
        st->beta = (st->flexible ? S - st->rzo : S) / st->rz;
        st->rz = S;
        break;
    case FIN_RZO:
        if (st->stop) break;
        st->rzo = S;
        break;
    case FIN_RESUME: {  // true residual after a converged pass; vanishing as in FIN_R0
        st->xpend = 0;
        st->rr = S;
        const double rn = sqrt(S);
        st->stop = (rn < 1e-15 || rn / st->den < st->rtol) ? 1 : 0;
        if (!st->stop && st->iter >= st->max_iter) st->stop = 3;
        break;
    }
    case FIN_CG_RR0: // cg.cuh:47-77: rr = (r,r), tol = r0 * rtol
        st->rz = S;
        st->rr = S;
        st->r0 = sqrt(S);
        st->den = st->stop_norm == MF3D_STOP_RHS_NORM ? st->bnorm : st->r0;
        if (st->r0 <= st->den * st->rtol) st->stop = 1;
        break;
    case FIN_CG_PAP: { // cg.cuh:84-124, operator negated: (p, -Ap) = -S exactly
        if (st->stop) break;
        const double pAp = -S;
        st->pAp = pAp;
        if ((st->iter + 1) % st->check_every == 0 && (!isfinite(pAp) || fabs(pAp) < 1e-30)) {
            st->iter += 1;
            st->stop = 2;
            break;
        }
        st->alpha = st->rz / pAp;
        break;
    }
    case FIN_CG_RR: // cg.cuh:126-156
        if (st->stop) break;
        st->iter += 1;
        st->rr = S;
        if (st->iter % st->check_every == 0 && sqrt(S) <= st->den * st->rtol) {
            st->stop = 1;
            break;
        }
        if (st->iter >= st->max_iter) st->stop = 3;
        st->beta = S / st->rz;
        st->rz = S;
        break;
    default: break;
    }
}

__device__ inline void finalize_body(const double* __restrict__ partials, unsigned np, int op,
                                     IterState* __restrict__ st, double* __restrict__ out) {
    const double S = finalize_sum(partials, np);
    if (threadIdx.x == 0) finalize_apply(op, S, st, out);
}

}  // namespace mf3d
