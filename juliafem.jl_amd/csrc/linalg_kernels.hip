// linalg_kernels.hip -- sparse matrix-vector product and PCG vector kernels.
//
// The assembled operator is the reference's CSR (sparse(COO), src/sparse/
// sparse.jl:54) stored as dense 3x3 node blocks: the column indices of one
// block row are kept once per block (bcol), the values in the scalar CSR
// order, so the scalar CSR handed to the caller is the values array itself.
// y = K x reads 8 (or 4) bytes per value plus 4 bytes per 9 values of index.
//
// PCG (preconditioned conjugate gradients with a 3x3 block-diagonal
// preconditioner) keeps every scalar on the device: the dot products are
// fused into the producing kernels and finished by the last block (see
// jfem_device.h), so an iteration needs no host round trip.  The
// reference's CG (ext/JuliaFEMCUDAExt.jl:531-577) copies two dot products to
// the host per iteration and has no preconditioner.
#include "jfem_device.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// G lanes per block row; lane k handles blocks k, k+G, ... of the row.
template <typename T, int G>
__global__ __launch_bounds__(256) void k_bsr_spmv(int N, const int32_t *__restrict__ brow,
                                                 const int32_t *__restrict__ bcol, const T *__restrict__ vals,
                                                 const T *__restrict__ x, const uint8_t *__restrict__ fixed,
                                                 T *__restrict__ y, const T *__restrict__ dotv, double *partials,
                                                 unsigned *counter, PcgScalars *S, int fin,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = gid / G, lane = gid % G;
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    if (a < N) {
        const int r0 = brow[a], nb = brow[a + 1] - r0;
        const T *v = vals + 9 * (size_t)r0;
        const size_t rs = 3 * (size_t)nb;
        for (int k = lane; k < nb; k += G) {
            const size_t col = 3 * (size_t)bcol[r0 + k];
            const T x0 = x[col], x1 = x[col + 1], x2 = x[col + 2];
            const T *v0 = v + 3 * (size_t)k;
            const T *v1 = v0 + rs;
            const T *v2 = v1 + rs;
            acc0 += v0[0] * x0 + v0[1] * x1 + v0[2] * x2;
            acc1 += v1[0] * x0 + v1[1] * x1 + v1[2] * x2;
            acc2 += v2[0] * x0 + v2[1] * x1 + v2[2] * x2;
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, G);
        acc1 += __shfl_xor(acc1, off, G);
        acc2 += __shfl_xor(acc2, off, G);
    }
    double dv = 0.0;
    if (a < N && lane < 3) {
        const size_t row = 3 * (size_t)a + lane;
        T val = lane == 0 ? acc0 : (lane == 1 ? acc1 : acc2);
        if (fixed && fixed[row]) val = T(0);
        y[row] = val;
        if (dotv) dv = (double)val * (double)dotv[row];
    }
    if (fin != FIN_NONE) {
        double vv[1] = {dv};
        reduce_finalize<1>(vv, partials, counter, S, fin);
    }
}

// z = Minv r, p = z, (r.z, r.r) -> FIN_INIT
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_init(int N, const T *__restrict__ r, T *__restrict__ z,
                                                 T *__restrict__ p, const T *__restrict__ Minv,
                                                 double *partials, unsigned *counter, PcgScalars *S) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (a < N) {
        const T r0 = r[3 * (size_t)a], r1 = r[3 * (size_t)a + 1], r2 = r[3 * (size_t)a + 2];
        const T *M = Minv + 9 * (size_t)a;
        const T z0 = M[0] * r0 + M[1] * r1 + M[2] * r2;
        const T z1 = M[3] * r0 + M[4] * r1 + M[5] * r2;
        const T z2 = M[6] * r0 + M[7] * r1 + M[8] * r2;
        z[3 * (size_t)a] = z0; z[3 * (size_t)a + 1] = z1; z[3 * (size_t)a + 2] = z2;
        p[3 * (size_t)a] = z0; p[3 * (size_t)a + 1] = z1; p[3 * (size_t)a + 2] = z2;
        v[0] = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
        v[1] = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    }
    reduce_finalize<2>(v, partials, counter, S, FIN_INIT);
}

// x += alpha p, r -= alpha q, z = Minv r, (r.z, r.r) -> FIN_UPDATE
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_update(int N, T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ z, const T *__restrict__ p,
                                                   const T *__restrict__ q, const T *__restrict__ Minv,
                                                   double *partials, unsigned *counter, PcgScalars *S) {
    if (S->done) return;
    const T alpha = (T)S->alpha;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (a < N) {
        const size_t o = 3 * (size_t)a;
        T rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[o + c] += alpha * p[o + c];
            rv[c] = r[o + c] - alpha * q[o + c];
            r[o + c] = rv[c];
        }
        const T *M = Minv + 9 * (size_t)a;
        const T z0 = M[0] * rv[0] + M[1] * rv[1] + M[2] * rv[2];
        const T z1 = M[3] * rv[0] + M[4] * rv[1] + M[5] * rv[2];
        const T z2 = M[6] * rv[0] + M[7] * rv[1] + M[8] * rv[2];
        z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
        v[0] = (double)rv[0] * z0 + (double)rv[1] * z1 + (double)rv[2] * z2;
        v[1] = (double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2];
    }
    reduce_finalize<2>(v, partials, counter, S, FIN_UPDATE);
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_dir(int n, T *__restrict__ p, const T *__restrict__ z,
                                                const PcgScalars *__restrict__ S) {
    if (S->done) return;
    const T beta = (T)S->beta;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
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

// r = b - Ku on free DOFs, 0 on fixed; (r.r) -> FIN_DOT
__global__ __launch_bounds__(256) void k_residual(const double *__restrict__ b, const double *__restrict__ Ku,
                                                 const uint8_t *__restrict__ fixed, double *__restrict__ r,
                                                 int n, double *partials, unsigned *counter, PcgScalars *S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double ri = (fixed && fixed[i]) ? 0.0 : b[i] - Ku[i];
        r[i] = ri;
        v[0] = ri * ri;
    }
    reduce_finalize<1>(v, partials, counter, S, FIN_DOT);
}

__global__ __launch_bounds__(256) void k_axpy_scaled(double *__restrict__ u, const float *__restrict__ d,
                                                    double scale, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += scale * (double)d[i];
}

__global__ __launch_bounds__(256) void k_dot(const double *__restrict__ a, const double *__restrict__ b, int n,
                                            double *partials, unsigned *counter, PcgScalars *S, int fin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {i < n ? a[i] * b[i] : 0.0};
    reduce_finalize<1>(v, partials, counter, S, fin);
}

// ------------------------------------------------------------- launchers ----
constexpr int kSpmvLanes = 8;

template <typename T>
hipError_t launch_bsr_spmv(jfm_ctx *ctx, const T *vals, const T *x, T *y, const T *dotv, int fin,
                           const int *done) {
    const int N = (int)ctx->N;
    if (N == 0) return hipSuccess;
    k_bsr_spmv<T, kSpmvLanes><<<nblk((int64_t)N * kSpmvLanes), kBlock, 0, ctx->stream>>>(
        N, ctx->brow, ctx->bcol, vals, x, ctx->fixed, y, dotv, ctx->partials, ctx->counter, ctx->S, fin, done);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_init(jfm_ctx *ctx, int64_t N, T *r, T *z, T *p, const T *Minv) {
    k_pcg_init<T><<<nblk(N), kBlock, 0, ctx->stream>>>((int)N, r, z, p, Minv, ctx->partials, ctx->counter, ctx->S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update(jfm_ctx *ctx, int64_t N, T *x, T *r, T *z, const T *p, const T *q, const T *Minv) {
    k_pcg_update<T><<<nblk(N), kBlock, 0, ctx->stream>>>((int)N, x, r, z, p, q, Minv, ctx->partials, ctx->counter, ctx->S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_direction(jfm_ctx *ctx, int64_t N, T *p, const T *z) {
    k_pcg_dir<T><<<nblk(3 * N), kBlock, 0, ctx->stream>>>((int)(3 * N), p, z, ctx->S);
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
                           int64_t n) {
    k_residual<<<nblk(n), kBlock, 0, ctx->stream>>>(b, Ku, fixed, r, (int)n, ctx->partials, ctx->counter, ctx->S);
    return hipGetLastError();
}

hipError_t launch_axpy_scaled(jfm_ctx *ctx, double *u, const float *d, double scale, int64_t n) {
    k_axpy_scaled<<<nblk(n), kBlock, 0, ctx->stream>>>(u, d, scale, (int)n);
    return hipGetLastError();
}

hipError_t launch_dot(jfm_ctx *ctx, const double *a, const double *b, int64_t n, int fin) {
    k_dot<<<nblk(n), kBlock, 0, ctx->stream>>>(a, b, (int)n, ctx->partials, ctx->counter, ctx->S, fin);
    return hipGetLastError();
}

template hipError_t launch_bsr_spmv<double>(jfm_ctx *, const double *, const double *, double *, const double *, int, const int *);
template hipError_t launch_bsr_spmv<float>(jfm_ctx *, const float *, const float *, float *, const float *, int, const int *);
template hipError_t launch_pcg_init<double>(jfm_ctx *, int64_t, double *, double *, double *, const double *);
template hipError_t launch_pcg_init<float>(jfm_ctx *, int64_t, float *, float *, float *, const float *);
template hipError_t launch_pcg_update<double>(jfm_ctx *, int64_t, double *, double *, double *, const double *, const double *, const double *);
template hipError_t launch_pcg_update<float>(jfm_ctx *, int64_t, float *, float *, float *, const float *, const float *, const float *);
template hipError_t launch_pcg_direction<double>(jfm_ctx *, int64_t, double *, const double *);
template hipError_t launch_pcg_direction<float>(jfm_ctx *, int64_t, float *, const float *);
template hipError_t launch_convert<double, float>(jfm_ctx *, const double *, float *, int64_t, double);
template hipError_t launch_convert<float, double>(jfm_ctx *, const float *, double *, int64_t, double);
template hipError_t launch_convert<double, double>(jfm_ctx *, const double *, double *, int64_t, double);

}  // namespace jfem
