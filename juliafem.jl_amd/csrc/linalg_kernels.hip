// linalg_kernels.hip -- sparse matrix-vector product and PCG vector kernels.
//
// The assembled operator is the reference's CSR (sparse(COO), src/sparse/
// sparse.jl:54) stored as dense 3x3 node blocks: the column indices of one
// block row are kept once per block (bcol), the values in the scalar CSR
// order, so the scalar CSR handed to the caller is the values array itself.
// y = K x reads 8 (or 4) bytes per value plus 4 bytes per 9 values of index.
//
// PCG (conjugate gradients with a 3x3 block-diagonal preconditioner) keeps
// every scalar on the device.  Each iteration is three kernels --
//   q = K p (+ partials of p.q)            [operator, slot s = i & 1]
//   x += a p, r -= a q, z = M r (+ r.z, r.r)  a = rz[s] / sum(p.q partials)
//   p = z + b p                            b = sum(r.z partials) / rz[s],
//                                          block 0 writes rz[s^1], done[s^1]
// -- with no host round trip and no atomics.  The reference's CG
// (ext/JuliaFEMCUDAExt.jl:531-577) has no preconditioner and copies two dot
// products to the host per iteration.
#include <algorithm>

#include "jfem_device.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// nodes per thread of the two-level direction kernel, all loaded before its
// reductions (its grid is sized for this many)
#ifndef JFEM_DIR_NODES
#define JFEM_DIR_NODES 3
#endif
constexpr int kDirNodes = JFEM_DIR_NODES;

// Non-temporal value loads measured slower on gfx950 (0.26 vs 0.18 ms at 87M
// values, 8 lanes per row): the default keeps ordinary loads.
#ifndef JFEM_SPMV_NT
#define JFEM_SPMV_NT 0
#endif
// block-row order: 0 grid-stride, 1 one contiguous run per block (XCD-grouped),
// 2 grid-stride sweeps with one contiguous window per XCD in each sweep
#ifndef JFEM_SPMV_MAP
#define JFEM_SPMV_MAP 0
#endif
template <typename T> __device__ __forceinline__ T ldnt(const T *p) {
#if JFEM_SPMV_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// G lanes per block row; lane k handles blocks k, k+G, ... of the row.
template <typename T, int G>
__global__ __launch_bounds__(256) void k_bsr_spmv(int N, const int32_t *__restrict__ brow,
                                                 const int32_t *__restrict__ bcol, const T *__restrict__ vals,
                                                 const T *__restrict__ x, const uint8_t *__restrict__ fixed,
                                                 T *__restrict__ y, const T *__restrict__ dotv,
                                                 double *__restrict__ part, const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int RPB = kBlock / G;  // rows per block sweep
    const int lane = threadIdx.x % G;
    double dv = 0.0;
#if JFEM_SPMV_MAP == 1
    // one contiguous run of block rows per block (XCD-grouped)
    const int per = ((N + (int)gridDim.x - 1) / (int)gridDim.x + RPB - 1) / RPB * RPB;
    const int begin = xcd_block() * per, end = min(N, begin + per), step = RPB;
#elif JFEM_SPMV_MAP == 2
    // grid-stride sweeps; within a sweep each XCD takes one contiguous window
    const int begin = xcd_block() * RPB, end = N, step = (int)gridDim.x * RPB;
#else
    const int begin = (int)blockIdx.x * RPB, end = N, step = (int)gridDim.x * RPB;
#endif
    for (int a0 = begin; a0 < end; a0 += step) {
        const int a = a0 + threadIdx.x / G;
        T acc0 = T(0), acc1 = T(0), acc2 = T(0);
        if (a < end) {
            const int r0 = brow[a], nb = brow[a + 1] - r0;
            const T *v = vals + 9 * (size_t)r0;
            const size_t rs = 3 * (size_t)nb;
            for (int k = lane; k < nb; k += G) {
                const size_t col = 3 * (size_t)bcol[r0 + k];
                const T x0 = x[col], x1 = x[col + 1], x2 = x[col + 2];
                const T *v0 = v + 3 * (size_t)k;
                const T *v1 = v0 + rs;
                const T *v2 = v1 + rs;
                acc0 += ldnt(v0) * x0 + ldnt(v0 + 1) * x1 + ldnt(v0 + 2) * x2;
                acc1 += ldnt(v1) * x0 + ldnt(v1 + 1) * x1 + ldnt(v1 + 2) * x2;
                acc2 += ldnt(v2) * x0 + ldnt(v2 + 1) * x1 + ldnt(v2 + 2) * x2;
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            acc0 += __shfl_xor(acc0, off, G);
            acc1 += __shfl_xor(acc1, off, G);
            acc2 += __shfl_xor(acc2, off, G);
        }
        if (a < end && lane < 3) {
            const size_t row = 3 * (size_t)a + lane;
            T val = lane == 0 ? acc0 : (lane == 1 ? acc1 : acc2);
            if (fixed && fixed[row]) val = T(0);
            y[row] = val;
            if (dotv) dv += (double)val * (double)dotv[row];
        }
    }
    if (part) {
        double vv[1] = {dv};
        write_partials<1>(vv, part);
    }
}

// z = Minv r, p = z; partials (r.z, r.r) -> misc
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_init(int N, const T *__restrict__ r, T *__restrict__ z,
                                                 T *__restrict__ p, const T *__restrict__ Minv,
                                                 double *__restrict__ part) {
    double v[2] = {0.0, 0.0};
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += gridDim.x * blockDim.x) {
        const size_t o = 3 * (size_t)a;
        const T r0 = r[o], r1 = r[o + 1], r2 = r[o + 2];
        const T *M = Minv + 9 * (size_t)a;
        const T z0 = M[0] * r0 + M[1] * r1 + M[2] * r2;
        const T z1 = M[3] * r0 + M[4] * r1 + M[5] * r2;
        const T z2 = M[6] * r0 + M[7] * r1 + M[8] * r2;
        z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
        p[o] = z0; p[o + 1] = z1; p[o + 2] = z2;
        v[0] += (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
        v[1] += (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    }
    write_partials<2>(v, part);
}

// x += alpha p, r -= alpha q, z = Minv r; partials (r.z, r.r).  A thread's
// first kU nodes (4 in FP32, 2 in FP64: every node at the usual grid) are
// all loaded before the alpha reduction, so that their latencies overlap
// each other and the reduction.
// With rA, the masked residual also goes to rA in aggregate order (node a at
// position apos[a]) for the coarse restriction, which then reads it
// contiguously.
template <typename T>
__global__ __launch_bounds__(512) void k_pcg_update(int N, int slot, int npart_pq, T *__restrict__ x,
                                                   T *__restrict__ r, T *__restrict__ z, const T *__restrict__ p,
                                                   const T *__restrict__ q, const T *__restrict__ Minv,
                                                   double *__restrict__ partials, PcgScalars *S,
                                                   T *__restrict__ rA = nullptr, const int32_t *__restrict__ apos = nullptr,
                                                   const uint8_t *__restrict__ fixed = nullptr) {
    const int b = (int)blockIdx.x, nb = (int)gridDim.x;
    constexpr int kU = sizeof(T) == 4 ? 4 : 2;
    const int stride = nb * blockDim.x;
    const int a0 = b * blockDim.x + threadIdx.x;
    Trip<T> xv[kU], pv[kU], rv[kU], qv[kU];
    Blk9<T> M[kU];
    auto load = [&](int u, int aa) {
        xv[u] = ld3(x, aa);
        pv[u] = ld3(p, aa);
        rv[u] = ld3(r, aa);
        qv[u] = ld3(q, aa);
        M[u] = ld9(Minv, aa);
    };
#pragma unroll
    for (int u = 0; u < kU; ++u)
        if (a0 + u * stride < N) load(u, a0 + u * stride);  // in flight while the flag is read
    if (S->done[slot]) return;
    double pq[1];
    sum_partials<1>(partials + part_pq(slot), npart_pq, pq);
    const double alpha_d = (pq[0] > 0.0 && finite(pq[0])) ? S->rz[slot] / pq[0] : 0.0;
    const T alpha = (T)alpha_d;
    double v[2] = {0.0, 0.0};
    auto update = [&](int u, int aa) {
        const size_t o = 3 * (size_t)aa;
        const T rn[3] = {rv[u].a - alpha * qv[u].a, rv[u].b - alpha * qv[u].b, rv[u].c - alpha * qv[u].c};
        st3(x, aa, xv[u].a + alpha * pv[u].a, xv[u].b + alpha * pv[u].b, xv[u].c + alpha * pv[u].c);
        st3(r, aa, rn[0], rn[1], rn[2]);
        const T *Mu = M[u].v;
        const T z0 = Mu[0] * rn[0] + Mu[1] * rn[1] + Mu[2] * rn[2];
        const T z1 = Mu[3] * rn[0] + Mu[4] * rn[1] + Mu[5] * rn[2];
        const T z2 = Mu[6] * rn[0] + Mu[7] * rn[1] + Mu[8] * rn[2];
        st3(z, aa, z0, z1, z2);
        if (rA) {
            const size_t oa = 3 * (size_t)apos[aa];
#pragma unroll
            for (int c = 0; c < 3; ++c) rA[oa + c] = fixed[o + c] ? T(0) : rn[c];
        }
        v[0] += (double)rn[0] * z0 + (double)rn[1] * z1 + (double)rn[2] * z2;
        v[1] += (double)rn[0] * rn[0] + (double)rn[1] * rn[1] + (double)rn[2] * rn[2];
    };
#pragma unroll
    for (int u = 0; u < kU; ++u)
        if (a0 + u * stride < N) update(u, a0 + u * stride);
    for (int a = a0 + kU * stride; a < N; a += stride) {  // grids too small for kU nodes per thread
        load(0, a);
        update(0, a);
    }
    write_partials<2>(v, partials + part_rz(slot), b);
    if (b == 0 && threadIdx.x == 0) {
        S->alpha = alpha_d;
        S->pq = pq[0];
    }
}

// Two-level form of the update, with the aggregates' nodes as the blocks'
// work: block b takes aggregates b per .. (b + 1) per - 1, each one's nodes
// agg_nodes[agg_ptr[g] .. agg_ptr[g + 1]) (ascending node id).  Besides r, z
// and the partials of (r.Br, r.r) (slot b), it forms the coarse right-hand
// side of the new residual from the values in registers, w_g = sum_{a in g}
// [m_a r_a ; d_a x (m_a r_a)] (cdA: the offsets d in agg_nodes order), in
// FP64 (w_new) and FP32 (wc, the GEMV's operand).  The first aggregate's
// first kU nodes per thread are in flight before the alpha reduction.  x is
// updated by the direction kernel (k_pcg_dir_coarse), which reads p anyway:
// this kernel then reads neither x nor p.
// With qsrc, q holds only the nodes the patch kernel finished, and a shared
// node's q is the sum of its partial slots bpart[s .. s + c) in ascending
// patch order, masked, as the shared-node kernel of the patch operator forms
// it (the same sums in the same order, so the same values): the first kQS
// slots are loaded with the node, further ones kQS at a time.
// 512 threads with four FP32 (three FP64) nodes each cover a 1,536-node
// aggregate in one pass of loads; 1,024 threads with two nodes each measured
// 4% slower solves.
constexpr int kAggThreads = 512;
template <typename T>
__global__ __launch_bounds__(kAggThreads) void k_pcg_update_agg(
    int n_agg, int nblk, int per, int slot, int npart_pq, T *__restrict__ r, T *__restrict__ z,
    const T *__restrict__ q, const int32_t *__restrict__ qsrc, const T *__restrict__ bpart,
    const T *__restrict__ Minv, const uint8_t *__restrict__ fixed,
    const int32_t *__restrict__ agg_ptr, const int32_t *__restrict__ agg_nodes, const T *__restrict__ cdA,
    double *__restrict__ partials, PcgScalars *S, double *__restrict__ w_new, float *__restrict__ wc) {
    const int b = xcd_block();
    if (b >= nblk) return;  // whole block (grid padded to a multiple of 8)
    const int g0 = b * per, g1 = min(g0 + per, n_agg);
    constexpr int kU = sizeof(T) == 4 ? 4 : 3;
    constexpr int kQS = sizeof(T) == 4 ? 4 : 2;  // slots loaded with the node (counts are mostly 2 and 4)
    const int t = (int)threadIdx.x, nt = (int)blockDim.x;
    Trip<T> rv[kU], qv[kU][kQS], dv[kU];
    Blk9<T> M[kU];
    Trip<uint8_t> fv[kU];
    int av[kU], qc[kU];
    const T *qp[kU];
    auto load = [&](int u, int k) {
        const int a = agg_nodes[k];
        av[u] = a;
        rv[u] = ld3(r, a);
        if (qsrc) {
            // a finished node reads q as its single "slot"
            const int e = qsrc[k];
            qp[u] = e < 0 ? q + 3 * (size_t)a : bpart + 3 * (size_t)(e >> 5);
            qc[u] = e < 0 ? 1 : (e & 31);
#pragma unroll
            for (int j = 0; j < kQS; ++j) qv[u][j] = j < qc[u] ? ld3(qp[u], j) : Trip<T>{T(0), T(0), T(0)};
        } else {
            qv[u][0] = ld3(q, a);
        }
        M[u] = ld9(Minv, a);
        fv[u] = ld3(fixed, a);
        dv[u] = ld3(cdA, k);
    };
    // q of node u: its slots summed in order from zero, masked (with qsrc)
    auto qsum = [&](int u) -> Trip<T> {
        if (!qsrc) return qv[u][0];
        T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
        for (int j = 0; j < kQS; ++j)
            if (j < qc[u]) {
                s0 += qv[u][j].a;
                s1 += qv[u][j].b;
                s2 += qv[u][j].c;
            }
        for (int j0 = kQS; j0 < qc[u]; j0 += kQS) {  // nodes of more than kQS slots
            Trip<T> w[kQS];
#pragma unroll
            for (int j = 0; j < kQS; ++j) w[j] = j0 + j < qc[u] ? ld3(qp[u], j0 + j) : Trip<T>{T(0), T(0), T(0)};
#pragma unroll
            for (int j = 0; j < kQS; ++j)
                if (j0 + j < qc[u]) {
                    s0 += w[j].a;
                    s1 += w[j].b;
                    s2 += w[j].c;
                }
        }
        if (fv[u].a) s0 = T(0);
        if (fv[u].b) s1 = T(0);
        if (fv[u].c) s2 = T(0);
        return Trip<T>{s0, s1, s2};
    };
    int kb = 0, ke = 0;
    auto prefetch = [&](int g) {
        kb = agg_ptr[g];
        ke = agg_ptr[g + 1];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (kb + t + u * nt < ke) load(u, kb + t + u * nt);
    };
    // the p.q partials are loaded first and summed by each wave alone, so
    // that alpha is known while the node loads are still in flight and the
    // first nodes' stores overlap the later nodes' loads (0.8 us faster than
    // the block reduction after the loads)
    WavePartials<1> wp;
    wp.load(partials + part_pq(slot), npart_pq);
    if (g0 < g1) prefetch(g0);
    if (S->done[slot]) return;  // uniform
    double pq[1];
    wp.sum(pq);
    const double alpha_d = (pq[0] > 0.0 && finite(pq[0])) ? S->rz[slot] / pq[0] : 0.0;
    const T alpha = (T)alpha_d;
    double v[2] = {0.0, 0.0};
    auto update = [&](int u, double (&w)[6]) {
        const int a = av[u];
        const Trip<T> qn = qsum(u);
        const T rn[3] = {rv[u].a - alpha * qn.a, rv[u].b - alpha * qn.b, rv[u].c - alpha * qn.c};
        st3(r, a, rn[0], rn[1], rn[2]);
        const T *Mu = M[u].v;
        const T z0 = Mu[0] * rn[0] + Mu[1] * rn[1] + Mu[2] * rn[2];
        const T z1 = Mu[3] * rn[0] + Mu[4] * rn[1] + Mu[5] * rn[2];
        const T z2 = Mu[6] * rn[0] + Mu[7] * rn[1] + Mu[8] * rn[2];
        st3(z, a, z0, z1, z2);
        v[0] += (double)rn[0] * z0 + (double)rn[1] * z1 + (double)rn[2] * z2;
        v[1] += (double)rn[0] * rn[0] + (double)rn[1] * rn[1] + (double)rn[2] * rn[2];
        const double y0 = fv[u].a ? 0.0 : (double)rn[0], y1 = fv[u].b ? 0.0 : (double)rn[1],
                     y2 = fv[u].c ? 0.0 : (double)rn[2];
        const double d0 = dv[u].a, d1 = dv[u].b, d2 = dv[u].c;
        w[0] += y0;
        w[1] += y1;
        w[2] += y2;
        w[3] += d1 * y2 - d2 * y1;
        w[4] += d2 * y0 - d0 * y2;
        w[5] += d0 * y1 - d1 * y0;
    };
    for (int g = g0; g < g1; ++g) {
        if (g > g0) prefetch(g);
        double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (kb + t + u * nt < ke) update(u, w);
        for (int k = kb + t + kU * nt; k < ke; k += nt) {  // aggregates of more than kU x nt nodes
            load(0, k);
            update(0, w);
        }
        block_sum_all<6>(w);
        if (t == 0) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                w_new[6 * (size_t)g + j] = w[j];
                wc[6 * (size_t)g + j] = (float)w[j];
            }
        }
    }
    write_partials<2>(v, partials + part_rz(slot), b);
    if (b == 0 && t == 0) {
        S->alpha = alpha_d;
        S->pq = pq[0];
    }
}

// the next iteration's scalars and status, written by one thread at the end
// of the direction kernel
__device__ __forceinline__ void pcg_record(PcgScalars *S, int slot, double rz, double rr, double beta) {
    const int it = S->iter + 1;
    S->iter = it;
    S->rz[slot ^ 1] = rz;
    S->rr = rr;
    S->beta = beta;
    int nd = DONE_RUNNING;
    if (rr <= S->tol2) nd = DONE_CONVERGED;
    else if (!(S->pq > 0.0) || !(rz > 0.0) || !finite(rz)) nd = DONE_BREAKDOWN;
    else if (it >= S->maxit) nd = DONE_MAXIT;
    S->done[slot ^ 1] = nd;
}

// p = z + beta p; block 0 records the next iteration's scalars and status
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_dir(int n3, int slot, int npart_rz, T *__restrict__ p,
                                                const T *__restrict__ z, const double *__restrict__ partials,
                                                PcgScalars *S) {
    const int stride = gridDim.x * blockDim.x;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    // first operands in flight before the flag test and the beta reduction
    T z0 = i0 < n3 ? z[i0] : T(0), p0 = i0 < n3 ? p[i0] : T(0);
    const int d = S->done[slot];
    if (d) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->done[slot ^ 1] = d;
        return;
    }
    double t[2];
    sum_partials<2>(partials + part_rz(slot), npart_rz, t);
    const double beta_d = t[0] / S->rz[slot];
    const T beta = (T)beta_d;
    if (i0 < n3) p[i0] = z0 + beta * p0;
    for (int i = i0 + stride; i < n3; i += stride) p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) pcg_record(S, slot, t[0], t[1], beta_d);
}

// Two-level form: p_a = z_a + m_a o (t_g + omega_g x d_a) + beta p_a with
// (t_g, omega_g) = zc of the node's aggregate, one thread per node; r.z is
// r.Br (update partials) + w.zc (coarse product partials).
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_dir_coarse(int N, int slot, int npart_rz, int npart_wz, T *__restrict__ p,
                                                       T *__restrict__ x, const T *__restrict__ z,
                                                       const int32_t *__restrict__ node_agg,
                                                       const T *__restrict__ cd, const uint8_t *__restrict__ fixed,
                                                       const float *__restrict__ zc, const double *__restrict__ partials,
                                                       PcgScalars *S) {
    // the first kDN nodes' operands in flight before the flag test and the
    // reductions
    constexpr int kDN = kDirNodes;
    const int a0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    Trip<T> zv[kDN], pv[kDN], dv[kDN], xv[kDN];
    Trip<uint8_t> fx[kDN];
    Trip<float> yc[kDN][2];  // the node's aggregate's (t, omega), read as soon as its id arrives
    auto load = [&](int u, int a) {
        zv[u] = ld3(z, a);
        pv[u] = ld3(p, a);
        if (x) xv[u] = ld3(x, a);
        dv[u] = ld3(cd, a);
        fx[u] = ld3(fixed, a);
        const int g = node_agg[a];
        yc[u][0] = ld3(zc, 2 * (size_t)g);
        yc[u][1] = ld3(zc, 2 * (size_t)g + 1);
    };
#pragma unroll
    for (int u = 0; u < kDN; ++u)
        if (a0 + u * stride < N) load(u, a0 + u * stride);
    const int d = S->done[slot];
    if (d) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->done[slot ^ 1] = d;
        return;
    }
    // (r.Br, r.r, w.zc) in one block reduction: the same sums as three
    // separate ones, with one barrier round fewer
    double t[3] = {0.0, 0.0, 0.0};
    {
        const double *prz = partials + part_rz(slot), *pwz = partials + kPartCoarse;
        for (int i = threadIdx.x; i < npart_rz; i += blockDim.x) {
            t[0] += prz[i];
            t[1] += prz[kMaxPart + i];
        }
        for (int i = threadIdx.x; i < npart_wz; i += blockDim.x) t[2] += pwz[i];
        block_sum_all<3>(t);
    }
    const double rz = t[0] + t[2];
    const double beta_d = rz / S->rz[slot];
    const T beta = (T)beta_d;
    const T alpha = (T)S->alpha;  // this iteration's step, as the update kernel applied it to r
    auto update = [&](int u, int a) {
        if (x) st3(x, a, xv[u].a + alpha * pv[u].a, xv[u].b + alpha * pv[u].b, xv[u].c + alpha * pv[u].c);
        const Trip<float> y0 = yc[u][0], y1 = yc[u][1];
        const T y[6] = {(T)y0.a, (T)y0.b, (T)y0.c, (T)y1.a, (T)y1.b, (T)y1.c};
        const T d0[3] = {dv[u].a, dv[u].b, dv[u].c};
        const T c[3] = {y[0] + y[4] * d0[2] - y[5] * d0[1], y[1] + y[5] * d0[0] - y[3] * d0[2],
                        y[2] + y[3] * d0[1] - y[4] * d0[0]};
        st3(p, a, (zv[u].a + (fx[u].a ? T(0) : c[0])) + beta * pv[u].a,
            (zv[u].b + (fx[u].b ? T(0) : c[1])) + beta * pv[u].b, (zv[u].c + (fx[u].c ? T(0) : c[2])) + beta * pv[u].c);
    };
#pragma unroll
    for (int u = 0; u < kDN; ++u)
        if (a0 + u * stride < N) update(u, a0 + u * stride);
    for (int a = a0 + kDN * stride; a < N; a += stride) {  // grids too small for kDN nodes per thread
        load(0, a);
        update(0, a);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) pcg_record(S, slot, rz, t[1], beta_d);
}

// one block: sum partials and apply `fin`
__global__ __launch_bounds__(256) void k_finalize(const double *__restrict__ part, int npart, int nv, int fin,
                                                 PcgScalars *S) {
    double t[2];
    sum_partials<2>(part, npart, t);
    if (threadIdx.x != 0) return;
    if (nv < 2) t[1] = 0.0;
    switch (fin) {
    case FIN_INIT: {
        S->rz[0] = t[0];
        S->rr = t[1];
        S->rr0 = t[1];
        S->iter = 0;
        const int d = (t[1] <= S->tol2) ? DONE_CONVERGED : DONE_RUNNING;
        S->done[0] = d;
        S->done[1] = d;
        break;
    }
    case FIN_DOT: S->tmp0 = t[0]; break;
    case FIN_DOT2: S->tmp0 = t[0]; S->tmp1 = t[1]; break;
    case FIN_BB: S->bb = t[0]; S->tol2 = S->rtol2 * t[0]; break;
    default: break;
    }
}

__global__ void k_zero_u32(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_convert(const TI *__restrict__ in, TO *__restrict__ out, int64_t n,
                                                double scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (TO)((double)in[i] * scale);
}

// r = b - Ku on free DOFs, 0 on fixed; partials of r.r
__global__ __launch_bounds__(256) void k_residual(const double *__restrict__ b, const double *__restrict__ Ku,
                                                 const uint8_t *__restrict__ fixed, double *__restrict__ r,
                                                 int n, double *__restrict__ part) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ri = (fixed && fixed[i]) ? 0.0 : b[i] - Ku[i];
        r[i] = ri;
        v[0] += ri * ri;
    }
    write_partials<1>(v, part);
}

__global__ __launch_bounds__(256) void k_axpy_scaled(double *__restrict__ u, const float *__restrict__ d,
                                                    double scale, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += scale * (double)d[i];
}

__global__ __launch_bounds__(256) void k_add(const double *__restrict__ a, const double *__restrict__ b,
                                            double *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T *__restrict__ a, const T *__restrict__ b, int n,
                                            double *__restrict__ part) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] += (double)a[i] * (double)b[i];
    write_partials<1>(v, part);
}

// ------------------------------------------------------------- launchers ----
// 16 lanes per block row measured fastest at 87M values (0.164 ms FP64
// against 0.182 with 8 and 0.218 with 4); the grid-stride row order beat the
// XCD-grouped orders (0.187, 0.194).
#ifndef JFEM_SPMV_LANES
#define JFEM_SPMV_LANES 16
#endif
constexpr int kSpmvLanes = JFEM_SPMV_LANES;

template <typename T>
hipError_t launch_bsr_spmv(jfm_ctx *ctx, const T *vals, const T *x, T *y, const T *dotv, double *part,
                           const int *done, int *npart) {
    const int N = (int)ctx->N;
    const unsigned g = xcd_grid(capped_grid((int64_t)N * kSpmvLanes, kBlock));
    if (npart) *npart = (int)g;
    k_bsr_spmv<T, kSpmvLanes><<<g, kBlock, 0, ctx->stream>>>(N, ctx->brow, ctx->bcol, vals, x, ctx->fixed, y,
                                                            dotv, part, done);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_init(jfm_ctx *ctx, T *r, T *z, T *p, const T *Minv, int *npart) {
    const unsigned g = capped_grid(ctx->N, kBlock);
    *npart = (int)g;
    k_pcg_init<T><<<g, kBlock, 0, ctx->stream>>>((int)ctx->N, r, z, p, Minv, ctx->partials + kPartMisc);
    return hipGetLastError();
}

#ifndef JFEM_VEC_BLOCKS
#define JFEM_VEC_BLOCKS 384
#endif
#ifndef JFEM_DIR_BLOCKS
#define JFEM_DIR_BLOCKS 1024
#endif
constexpr int kVecBlocks = JFEM_VEC_BLOCKS;  // grid of the PCG update kernel (384: measured 0.4-1.3% faster solves than 512)
constexpr int kDirBlocks = JFEM_DIR_BLOCKS;  // grid of the PCG direction kernels

template <typename T>
hipError_t launch_pcg_update(jfm_ctx *ctx, int slot, int npart_pq, T *x, T *r, T *z, const T *p, const T *q,
                             const T *Minv, int *npart, T *rA, bool restrict_q, bool shared_q) {
    const unsigned g = std::min<unsigned>(capped_grid(ctx->N, kBlock), kVecBlocks);
    *npart = (int)g;
    if (restrict_q) {  // one block per aggregate (or per run of them), w to wc64 by parity
        const T *cdA = sizeof(T) == 8 ? (const T *)ctx->cdA64 : (const T *)ctx->cdA32;
        const int n_agg = (int)ctx->n_agg;
        const int per = (n_agg + kMaxPart - 1) / kMaxPart;
        const int nb = (n_agg + per - 1) / per;
        *npart = nb;
        if (shared_q && !ctx->qsrc) return hipErrorInvalidValue;
        k_pcg_update_agg<T><<<xcd_grid((unsigned)nb), kAggThreads, 0, ctx->stream>>>(
            n_agg, nb, per, slot, npart_pq, r, z, q, shared_q ? ctx->qsrc : nullptr,
            shared_q ? (const T *)ctx->bpart : nullptr, Minv, ctx->fixed, ctx->agg_ptr, ctx->agg_nodes, cdA,
            ctx->partials, ctx->S, ctx->wc64 + (size_t)(slot ^ 1) * ctx->nc, ctx->wc32);
        return hipGetLastError();
    }
    k_pcg_update<T><<<g, kBlock, 0, ctx->stream>>>((int)ctx->N, slot, npart_pq, x, r, z, p, q, Minv,
                                                  ctx->partials, ctx->S, rA, rA ? ctx->apos : nullptr, ctx->fixed);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_direction(jfm_ctx *ctx, int slot, int npart_rz, T *p, const T *z) {
    const int64_t n3 = 3 * ctx->N;
    k_pcg_dir<T><<<std::min<unsigned>(capped_grid(n3, kBlock), kDirBlocks), kBlock, 0, ctx->stream>>>((int)n3, slot, npart_rz, p, z, ctx->partials,
                                                                      ctx->S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_direction_coarse(jfm_ctx *ctx, int slot, int npart_rz, int npart_wz, T *p, const T *z, T *x) {
    const T *cd = sizeof(T) == 8 ? (const T *)ctx->cd64 : (const T *)ctx->cd32;
    // kDirNodes nodes per thread, at most kDirBlocks blocks (no partials are
    // written, so the grid is not capped at kMaxPart): 456 blocks of three
    // nodes per thread at 349,569 nodes measured 0.12 ms faster per solve
    // than 683 of two, and those 0.06-0.15 ms faster than 1024
    const int64_t per = (int64_t)kBlock * kDirNodes;
    const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (ctx->N + per - 1) / per), kDirBlocks);
    k_pcg_dir_coarse<T><<<g, kBlock, 0, ctx->stream>>>(
        (int)ctx->N, slot, npart_rz, npart_wz, p, x, z, ctx->node_agg, cd, ctx->fixed, ctx->zc32, ctx->partials, ctx->S);
    return hipGetLastError();
}
template hipError_t launch_pcg_direction_coarse<double>(jfm_ctx *, int, int, int, double *, const double *, double *);
template hipError_t launch_pcg_direction_coarse<float>(jfm_ctx *, int, int, int, float *, const float *, float *);

hipError_t launch_finalize(jfm_ctx *ctx, const double *part, int npart, int nv, int fin) {
    k_finalize<<<1, kBlock, 0, ctx->stream>>>(part, npart, nv, fin, ctx->S);
    return hipGetLastError();
}

hipError_t launch_zero_u32(jfm_ctx *ctx, unsigned *p, int64_t n) {
    k_zero_u32<<<nblk(n), kBlock, 0, ctx->stream>>>(p, (int)n);
    return hipGetLastError();
}

template <typename TI, typename TO>
hipError_t launch_convert(jfm_ctx *ctx, const TI *in, TO *out, int64_t n, double scale) {
    if (n == 0) return hipSuccess;
    k_convert<TI, TO><<<nblk(n), kBlock, 0, ctx->stream>>>(in, out, n, scale);
    return hipGetLastError();
}

hipError_t launch_residual(jfm_ctx *ctx, const double *b, const double *Ku, const uint8_t *fixed, double *r,
                           int64_t n, int *npart) {
    const unsigned g = capped_grid(n, kBlock);
    *npart = (int)g;
    k_residual<<<g, kBlock, 0, ctx->stream>>>(b, Ku, fixed, r, (int)n, ctx->partials + kPartMisc);
    return hipGetLastError();
}

hipError_t launch_axpy_scaled(jfm_ctx *ctx, double *u, const float *d, double scale, int64_t n) {
    k_axpy_scaled<<<nblk(n), kBlock, 0, ctx->stream>>>(u, d, scale, (int)n);
    return hipGetLastError();
}

hipError_t launch_add(jfm_ctx *ctx, const double *a, const double *b, double *out, int64_t n) {
    k_add<<<nblk(n), kBlock, 0, ctx->stream>>>(a, b, out, (int)n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dot(jfm_ctx *ctx, const T *a, const T *b, int64_t n, int *npart) {
    const unsigned g = capped_grid(n, kBlock);
    *npart = (int)g;
    k_dot<T><<<g, kBlock, 0, ctx->stream>>>(a, b, (int)n, ctx->partials + kPartMisc);
    return hipGetLastError();
}

template hipError_t launch_bsr_spmv<double>(jfm_ctx *, const double *, const double *, double *, const double *, double *, const int *, int *);
template hipError_t launch_bsr_spmv<float>(jfm_ctx *, const float *, const float *, float *, const float *, double *, const int *, int *);
template hipError_t launch_pcg_init<double>(jfm_ctx *, double *, double *, double *, const double *, int *);
template hipError_t launch_pcg_init<float>(jfm_ctx *, float *, float *, float *, const float *, int *);
template hipError_t launch_pcg_update<double>(jfm_ctx *, int, int, double *, double *, double *, const double *, const double *, const double *, int *, double *, bool, bool);
template hipError_t launch_pcg_update<float>(jfm_ctx *, int, int, float *, float *, float *, const float *, const float *, const float *, int *, float *, bool, bool);
template hipError_t launch_pcg_direction<double>(jfm_ctx *, int, int, double *, const double *);
template hipError_t launch_pcg_direction<float>(jfm_ctx *, int, int, float *, const float *);
template hipError_t launch_convert<double, float>(jfm_ctx *, const double *, float *, int64_t, double);
template hipError_t launch_convert<float, double>(jfm_ctx *, const float *, double *, int64_t, double);
template hipError_t launch_convert<double, double>(jfm_ctx *, const double *, double *, int64_t, double);
template hipError_t launch_dot<double>(jfm_ctx *, const double *, const double *, int64_t, int *);
template hipError_t launch_dot<float>(jfm_ctx *, const float *, const float *, int64_t, int *);

}  // namespace jfem
