// jfem_device.h -- device helpers shared by the kernel files.
//
// Reductions are deterministic: a fixed xor-butterfly inside each wave, the
// waves of a block in index order, and the per-block partials summed by the
// last block to finish (found with an integer atomic counter) in a fixed
// strided order.  Which block happens to finish last does not change the
// order of any floating-point addition.
#pragma once

#include "jfem_internal.h"

namespace jfem {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in thread 0 only.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[16]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sh[k][w];
            v[k] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool finite(double x) { return x - x == 0.0; }

__device__ inline void finalize(PcgScalars *S, int fin, const double *t) {
    switch (fin) {
    case FIN_ALPHA:
        S->pq = t[0];
        if (t[0] > 0.0 && finite(t[0])) {
            S->alpha = S->rz / t[0];
        } else {
            S->alpha = 0.0;
            S->done = DONE_BREAKDOWN;
        }
        break;
    case FIN_UPDATE: {
        const double rzn = t[0];
        S->rr = t[1];
        S->iter += 1;
        S->beta = rzn / S->rz;
        S->rz = rzn;
        if (S->rr <= S->tol2) S->done = DONE_CONVERGED;
        else if (!(rzn > 0.0) || !finite(rzn)) S->done = DONE_BREAKDOWN;
        else if (S->iter >= S->maxit) S->done = DONE_MAXIT;
        break;
    }
    case FIN_INIT:
        S->rz = t[0];
        S->rr = t[1];
        S->iter = 0;
        S->beta = 0.0;
        S->done = (S->rr <= S->tol2) ? DONE_CONVERGED : DONE_RUNNING;
        break;
    case FIN_DOT: S->tmp0 = t[0]; break;
    case FIN_DOT2: S->tmp0 = t[0]; S->tmp1 = t[1]; break;
    case FIN_BB: S->bb = t[0]; S->tol2 = S->rtol2 * t[0]; break;
    default: break;
    }
}

// Block sums of v[] go to partials[k*gridDim.x + blockIdx.x]; the last block
// to arrive sums them and applies `fin` to the scalars.
template <int NV>
__device__ __forceinline__ void reduce_finalize(double (&v)[NV], double *partials, unsigned *counter,
                                                PcgScalars *S, int fin) {
    __shared__ double sh[NV][16];
    __shared__ int s_last;
    block_sum<NV>(v, sh);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * nb + blockIdx.x] = v[k];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        s_last = (prev == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            s += __hip_atomic_load(partials + k * nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t[k] = s;
    }
    block_sum<NV>(t, sh);
    if (threadIdx.x == 0) {
        *counter = 0u;
        finalize(S, fin, t);
    }
}

template <typename T> struct Vec3 { T x, y, z; };

}  // namespace jfem
