// learned.hip -- learned 3x3 block preconditioner (gfx950).
//
// No counterpart in the reference (its CG has no preconditioner,
// ext/JuliaFEMCUDAExt.jl:531-577).  Pieces:
//   k_features   24 local features per node (diagonal block, stiffness shares
//                of the incident elements, material field, boundary flags)
//   k_mlp_infer  24 -> 64 -> 64 -> 6 network on 16-node tiles with
//                v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation)
//   k_mlp        the training pass: the same forward pass, then the backward
//                pass, weight gradients again on MFMA
//   k_mblocks    M_a = C_a (L L' + eps I) C_a' from the outputs
//   k_block_apply / k_accum_dm   the loss pieces of a training step
// The loss of one (residual r, error e = K^-1 r) pair is the energy-norm
// error left after the best step along z = M r:
//     loss = 1 - (z.r)^2 / ((z.Kz)(e.r)),
// a scale-free measure of how well M approximates K^-1 on r.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "jfem_device.h"
#include "jfem_learned.h"

using namespace jfem;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace jfem {

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ------------------------------------------------------------------ features --
template <int NPE> constexpr int corners() { return NPE == 10 ? 4 : NPE; }

// Per-element inputs of the node features, once per element instead of once
// per incidence: ef[e] = log E, ef[n_el + e] = shape quality |invJ|_F detJ^(1/3)
// (first Gauss point), ef[2 n_el + e] = 1 if the element touches a fixed DOF.
template <int NPE>
__global__ __launch_bounds__(256) void k_elem_features(int n_el, const int32_t *__restrict__ conn,
                                                      const double *__restrict__ gf, const double *__restrict__ E,
                                                      const double *__restrict__ nu, const uint8_t *__restrict__ fixed,
                                                      double *__restrict__ ef) {
    using EL = Elem<NPE>;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    const double *G = gf + e;
    double fro = 0.0;
    for (int m = 0; m < 9; ++m) {
        const double v = G[(size_t)m * n_el];
        fro += v * v;
    }
    const double cl = G[(size_t)9 * n_el];
    const double Ee = E[e], ne = nu[e];
    const double la = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double det = cl / (la * EL::w(0));
    bool fx = false;
    for (int j = 0; j < NPE; ++j) {
        const size_t nd = 3 * (size_t)conn[(size_t)j * n_el + e];
        fx = fx || fixed[nd] || fixed[nd + 1] || fixed[nd + 2];
    }
    ef[e] = log(Ee);
    ef[(size_t)n_el + e] = sqrt(fro) * cbrt(det);
    ef[2 * (size_t)n_el + e] = fx ? 1.0 : 0.0;
}

template <int NPE>
__global__ __launch_bounds__(256) void k_features(int N, int n_el, int ngf, const int32_t *__restrict__ ptr,
                                                 const int32_t *__restrict__ inc, const double *__restrict__ ef,
                                                 const double *__restrict__ gf, const double *__restrict__ E,
                                                 const uint8_t *__restrict__ fixed, const int32_t *__restrict__ fptr,
                                                 const double *__restrict__ Dover, float *__restrict__ feat) {
    using EL = Elem<NPE>;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double tmin = 1e300, tmax = 0, tsum = 0, tsq = 0;
    double emin = 1e300, emax = 0, esum = 0, lesum = 0, tle = 0;
    double qsum = 0, qmin = 1e300, qmax = 0;
    int ncorner = 0, nfixel = 0;
    const int k0 = ptr[a], k1 = ptr[a + 1];
    const int n = k1 - k0;
    for (int k = k0; k < k1; ++k) {
        const int idx = inc[k];
        const int e = idx / NPE, i = idx - e * NPE;
        double te = 0.0;
        double Ji[9], cl = 0.0, cm = 0.0;
        for (int qp = 0; qp < EL::NGP; ++qp) {
            if (qp == 0 || ngf > 1) {   // affine elements: one set of factors
                const double *G = gf + (size_t)(ngf == 1 ? 0 : qp) * GF * n_el + e;
                for (int m = 0; m < 9; ++m) Ji[m] = G[(size_t)m * n_el];
                cl = G[(size_t)9 * n_el];
                cm = G[(size_t)10 * n_el];
            }
            double g[3];
            const double d0 = dn_rt<NPE>(qp, i, 0), d1 = dn_rt<NPE>(qp, i, 1), d2 = dn_rt<NPE>(qp, i, 2);
            for (int b = 0; b < 3; ++b) g[b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
            const double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) D[r * 3 + c] += (cl + cm) * g[r] * g[c] + (r == c ? cm * gg : 0.0);
            te += (cl + 4.0 * cm) * gg;
        }
        const double Ee = E[e], lE = ef[e], q = ef[(size_t)n_el + e];
        tmin = fmin(tmin, te); tmax = fmax(tmax, te); tsum += te; tsq += te * te;
        emin = fmin(emin, Ee); emax = fmax(emax, Ee); esum += Ee; lesum += lE; tle += te * lE;
        qsum += q; qmin = fmin(qmin, q); qmax = fmax(qmax, q);
        if (i < corners<NPE>()) ++ncorner;
        if (ef[2 * (size_t)n_el + e] != 0.0) ++nfixel;
    }
    if (Dover)  // diagonal block of the tangent in place of the linear one
        for (int m = 0; m < 9; ++m) D[m] = Dover[9 * (size_t)a + m];
    const double emean = esum / n;
    int nabove = 0;
    for (int k = k0; k < k1; ++k) nabove += (E[inc[k] / NPE] > emean) ? 1 : 0;
    const double T = D[0] + D[4] + D[8];
    const double s = 3.0 / T;
    const double d00 = D[0] * s, d11 = D[4] * s, d22 = D[8] * s, d01 = D[1] * s, d12 = D[5] * s, d02 = D[2] * s;
    const double tmean = tsum / n;
    float *f = feat + (size_t)a * kFeat;
    f[0] = (float)d00; f[1] = (float)d11; f[2] = (float)d22;
    f[3] = (float)d01; f[4] = (float)d12; f[5] = (float)d02;
    f[6] = (float)log((double)n);
    f[7] = (float)(tmax / tsum);
    f[8] = (float)(n * tmin / tsum);
    f[9] = (float)(sqrt(fmax(tsq / n - tmean * tmean, 0.0)) / tmean);
    f[10] = (float)log(emax / emin);
    f[11] = (float)(log(emean) - lesum / n);
    f[12] = (float)((double)ncorner / n);
    const size_t o = 3 * (size_t)a;
    f[13] = (float)((fixed[o] + fixed[o + 1] + fixed[o + 2]) / 3.0);
    f[14] = (float)((double)nfixel / n);
    f[15] = (fptr[a + 1] > fptr[a]) ? 1.0f : 0.0f;  // node on a loaded surface element
    const double det = d00 * (d11 * d22 - d12 * d12) - d01 * (d01 * d22 - d12 * d02) + d02 * (d01 * d12 - d11 * d02);
    f[16] = (float)det;
    f[17] = (float)sqrt(2.0 * (d01 * d01 + d12 * d12 + d02 * d02));
    f[18] = (float)(qsum / n);
    f[19] = (float)qmin;
    f[20] = (float)qmax;
    f[21] = (float)(tle / tsum - log(emax));
    f[22] = (float)((tmax - tmin) / tsum);
    f[23] = (float)((double)nabove / n);
}

// ------------------------------------------------------------- network -----
constexpr int kTile = 16;      // nodes per wave tile (the MFMA M dimension)
constexpr int kWaves = 2;      // waves per block (keeps the LDS image under 64 KiB)
constexpr int kXPad = 32;      // input width padded to two MFMA K steps

struct MlpLds {
    uint16_t W1[kHid][kXPad];  // [out][in], bf16, zero-padded
    uint16_t W2[kHid][kHid];
    uint16_t W3[16][kHid];     // rows >= kOut zero
    float b1[kHid], b2[kHid], b3[16];
    float shift[kFeat], scale[kFeat];
    struct Wave {
        uint16_t X[kTile][kXPad];
        uint16_t H1[kTile][kHid], H2[kTile][kHid];
        float Z1[kTile][kHid], Z2[kTile][kHid];  // pre-activations, then FP32 dZ
        float O[kTile][16];                      // outputs, then FP32 dO
        uint16_t dO[kTile][16], dZ2[kTile][kHid], dZ1[kTile][kHid];
    } w[kWaves];
};

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *(const bf16x4 *)p; }
__device__ __forceinline__ bf16x4 ld4_strided(const uint16_t *p, int stride) {
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[stride];
    v[2] = (short)p[2 * stride];
    v[3] = (short)p[3 * stride];
    return v;
}

// v_mfma_f32_16x16x16_bf16 operand layout (lane l, r = l % 16, q = 4 * (l / 16)):
//   A: A[r][q..q+3]   B: B[q..q+3][r]   C: C[q + j][r], j = 0..3
// C(16 x 16*NT) = A(16 x K) * W'(K x 16*NT), W row-major [out][in] (stride ws)
template <int K, int NT>
__device__ __forceinline__ void dense_fwd(const uint16_t *A, int as, const uint16_t *W, int ws, f32x4 (&acc)[NT]) {
    const int l = threadIdx.x & 63, r = l & 15, q = (l >> 4) * 4;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x4 a = ld4(A + r * as + k0 + q);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = mfma(a, ld4(W + (t * 16 + r) * ws + k0 + q), acc[t]);
    }
}

// C(16 x 16*NT) = dY(16 x K) * W(K x 16*NT), W row-major [out][in]
template <int K, int NT>
__device__ __forceinline__ void dense_bwd(const uint16_t *dY, int ys, const uint16_t *W, int ws, f32x4 (&acc)[NT]) {
    const int l = threadIdx.x & 63, r = l & 15, q = (l >> 4) * 4;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x4 a = ld4(dY + r * ys + k0 + q);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = mfma(a, ld4_strided(W + (k0 + q) * ws + t * 16 + r, ws), acc[t]);
    }
}

// acc[mt][nt] += dY'(16*MT x 16 nodes) * X(16 nodes x 16*NT)
template <int MT, int NT>
__device__ __forceinline__ void dense_wgrad(const uint16_t *dY, int ys, const uint16_t *X, int xs,
                                            f32x4 (&acc)[MT][NT]) {
    const int l = threadIdx.x & 63, r = l & 15, q = (l >> 4) * 4;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const bf16x4 a = ld4_strided(dY + q * ys + mt * 16 + r, ys);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(a, ld4_strided(X + q * xs + nt * 16 + r, xs), acc[mt][nt]);
    }
}

__device__ __forceinline__ double clampo(double o) { return fmin(fmax(o, -5.0), 5.0); }

// Cholesky factor C (lower) of a 3x3 SPD matrix S (row-major): C C' = S.
__device__ __forceinline__ void chol3(const double *S, double *C) {
    const double c00 = sqrt(S[0]);
    const double c10 = S[3] / c00, c20 = S[6] / c00;
    const double c11 = sqrt(S[4] - c10 * c10);
    const double c21 = (S[7] - c20 * c10) / c11;
    const double c22 = sqrt(S[8] - c20 * c20 - c21 * c21);
    C[0] = c00; C[1] = 0; C[2] = 0;
    C[3] = c10; C[4] = c11; C[5] = 0;
    C[6] = c20; C[7] = c21; C[8] = c22;
}

__device__ __forceinline__ bool node_fixed(const uint8_t *fixed, int a) {
    const size_t o = 3 * (size_t)a;
    return fixed[o] || fixed[o + 1] || fixed[o + 2];
}

// Training pass over all nodes: recomputes the forward pass, forms dLoss/do
// from G = dLoss/dM (N x 9) and writes one set of weight-gradient partials per
// wave to gpart.  (TRAIN = true is the only instantiation; inference is
// k_mlp_infer below.)
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp(int N, const float *__restrict__ feat, const uint8_t *__restrict__ blob,
                                            float *__restrict__ out, const double *__restrict__ G,
                                            const double *__restrict__ Minv, const uint8_t *__restrict__ fixed,
                                            float *__restrict__ gpart) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MlpLds &S = *reinterpret_cast<MlpLds *>(lds_raw);
    const LearnedBlob *B = reinterpret_cast<const LearnedBlob *>(blob);
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, r = l & 15, q = (l >> 4) * 4;
    // weights to LDS (zero padding)
    for (int i = tid; i < kHid * kXPad; i += blockDim.x) {
        const int o = i / kXPad, c = i % kXPad;
        S.W1[o][c] = c < kFeat ? B->W1[o * kFeat + c] : (uint16_t)0;
    }
    for (int i = tid; i < kHid * kHid; i += blockDim.x) S.W2[i / kHid][i % kHid] = B->W2[i];
    for (int i = tid; i < 16 * kHid; i += blockDim.x) {
        const int o = i / kHid, c = i % kHid;
        S.W3[o][c] = o < kOut ? B->W3[o * kHid + c] : (uint16_t)0;
    }
    for (int i = tid; i < kHid; i += blockDim.x) { S.b1[i] = B->b1[i]; S.b2[i] = B->b2[i]; }
    for (int i = tid; i < 16; i += blockDim.x) S.b3[i] = i < kOut ? B->b3[i] : 0.f;
    for (int i = tid; i < kFeat; i += blockDim.x) { S.shift[i] = B->feat_shift[i]; S.scale[i] = B->feat_scale[i]; }
    __syncthreads();
    MlpLds::Wave &W = S.w[wv];
    f32x4 gW1[4][2], gW2[4][4], gW3[1][4];
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f;
    if (TRAIN) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            gW1[i][0] = gW1[i][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 4; ++j) gW2[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            gW3[0][i] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
    // A lane reads feature column fc of nodes (l + 64 u) / kXPad of the tile,
    // all kFL loads at once, and the next tile's while this one is computed.
    constexpr int kFL = kTile * kXPad / 64;
    const int fc = l % kXPad;
    float fv[kFL];
    auto fetch = [&](int tile_base) {
        const int t0 = tile_base + wv * kTile;
#pragma unroll
        for (int u = 0; u < kFL; ++u) {
            const int node = t0 + (l + 64 * u) / kXPad;
            fv[u] = (fc < kFeat && node < N) ? feat[(size_t)node * kFeat + fc] : 0.f;
        }
    };
    const int stride = gridDim.x * kWaves * kTile;
    if (blockIdx.x * kWaves * kTile < N) fetch(blockIdx.x * kWaves * kTile);
    for (int base = blockIdx.x * kWaves * kTile; base < N; base += stride) {
        const int n0 = base + wv * kTile;
#pragma unroll
        for (int u = 0; u < kFL; ++u) {
            const int m = (l + 64 * u) / kXPad, node = n0 + m;
            W.X[m][fc] = f2bf((fc < kFeat && node < N) ? (fv[u] - S.shift[fc]) * S.scale[fc] : 0.f);
        }
        if (base + stride < N) fetch(base + stride);
        __syncthreads();
        {
            f32x4 acc[4];
            dense_fwd<kXPad, 4>(&W.X[0][0], kXPad, &S.W1[0][0], kXPad, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    const float z = acc[t][j] + S.b1[n];
                    if (TRAIN) W.Z1[q + j][n] = z;  // pre-activations for the backward pass
                    W.H1[q + j][n] = f2bf(gelu(z));
                }
        }
        __syncthreads();
        {
            f32x4 acc[4];
            dense_fwd<kHid, 4>(&W.H1[0][0], kHid, &S.W2[0][0], kHid, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    const float z = acc[t][j] + S.b2[n];
                    if (TRAIN) W.Z2[q + j][n] = z;
                    W.H2[q + j][n] = f2bf(gelu(z));
                }
        }
        __syncthreads();
        {
            f32x4 acc[1];
            dense_fwd<kHid, 1>(&W.H2[0][0], kHid, &S.W3[0][0], kHid, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) W.O[q + j][r] = acc[0][j] + S.b3[r];
        }
        __syncthreads();
        // dLoss/do per node (lane m < 16)
        if (l < kTile) {
            const int node = n0 + l;
            double dO[6] = {0, 0, 0, 0, 0, 0};
            if (node < N && !node_fixed(fixed, node) && Minv[9 * (size_t)node] > 0.0) {
                const double *Mi = Minv + 9 * (size_t)node;
                double C[9];
                chol3(Mi, C);
                const double o0 = W.O[l][0], o2 = W.O[l][2], o5 = W.O[l][5];
                const double L[9] = {exp(clampo(o0)), 0, 0, W.O[l][1], exp(clampo(o2)), 0, W.O[l][3], W.O[l][4],
                                     exp(clampo(o5))};
                const double *Gn = G + 9 * (size_t)node;
                double Gs[9], T1[9], Q[9], dL[9];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Gs[i * 3 + j] = Gn[i * 3 + j] + Gn[j * 3 + i];
                for (int i = 0; i < 3; ++i)  // T1 = Gs C
                    for (int j = 0; j < 3; ++j) T1[i * 3 + j] = Gs[i * 3] * C[j] + Gs[i * 3 + 1] * C[3 + j] + Gs[i * 3 + 2] * C[6 + j];
                for (int i = 0; i < 3; ++i)  // Q = C' T1
                    for (int j = 0; j < 3; ++j) Q[i * 3 + j] = C[i] * T1[j] + C[3 + i] * T1[3 + j] + C[6 + i] * T1[6 + j];
                for (int i = 0; i < 3; ++i)  // dL = Q L
                    for (int j = 0; j < 3; ++j) dL[i * 3 + j] = Q[i * 3] * L[j] + Q[i * 3 + 1] * L[3 + j] + Q[i * 3 + 2] * L[6 + j];
                dO[0] = (o0 > -5.0 && o0 < 5.0) ? dL[0] * L[0] : 0.0;
                dO[1] = dL[3];
                dO[2] = (o2 > -5.0 && o2 < 5.0) ? dL[4] * L[4] : 0.0;
                dO[3] = dL[6];
                dO[4] = dL[7];
                dO[5] = (o5 > -5.0 && o5 < 5.0) ? dL[8] * L[8] : 0.0;
            }
            for (int c = 0; c < 16; ++c) {
                const float v = c < kOut ? (float)dO[c] : 0.f;
                W.O[l][c] = v;
                W.dO[l][c] = f2bf(v);
            }
        }
        __syncthreads();
        {   // dZ2 = (dO W3) * gelu'(Z2)
            f32x4 acc[4];
            dense_bwd<16, 4>(&W.dO[0][0], 16, &S.W3[0][0], kHid, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    const float d = acc[t][j] * gelu_grad(W.Z2[q + j][n]);
                    W.Z2[q + j][n] = d;
                    W.dZ2[q + j][n] = f2bf(d);
                }
        }
        __syncthreads();
        {   // dZ1 = (dZ2 W2) * gelu'(Z1)
            f32x4 acc[4];
            dense_bwd<kHid, 4>(&W.dZ2[0][0], kHid, &S.W2[0][0], kHid, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    const float d = acc[t][j] * gelu_grad(W.Z1[q + j][n]);
                    W.Z1[q + j][n] = d;
                    W.dZ1[q + j][n] = f2bf(d);
                }
        }
        __syncthreads();
        dense_wgrad<1, 4>(&W.dO[0][0], 16, &W.H2[0][0], kHid, gW3);
        dense_wgrad<4, 4>(&W.dZ2[0][0], kHid, &W.H1[0][0], kHid, gW2);
        dense_wgrad<4, 2>(&W.dZ1[0][0], kHid, &W.X[0][0], kXPad, gW1);
        for (int m = 0; m < kTile; ++m) {
            gb1 += W.Z1[m][l];
            gb2 += W.Z2[m][l];
            if (l < 16) gb3 += W.O[m][l];
        }
        __syncthreads();
    }
    if (!TRAIN) return;
    float *gp = gpart + (size_t)(blockIdx.x * kWaves + wv) * kNParam;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = mt * 16 + q + j;
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int in = nt * 16 + r;
                if (in < kFeat) gp[kOffW1 + o * kFeat + in] = gW1[mt][nt][j];
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gp[kOffW2 + o * kHid + nt * 16 + r] = gW2[mt][nt][j];
        }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int o = q + j;
        if (o < kOut)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gp[kOffW3 + o * kHid + nt * 16 + r] = gW3[0][nt][j];
    }
    gp[kOffB1 + l] = gb1;
    gp[kOffB2 + l] = gb2;
    if (l < kOut) gp[kOffB3 + l] = gb3;
}

// M_a = C (L L' + eps I) C'; nodes with a fixed DOF (or no element) keep Minv
__global__ __launch_bounds__(256) void k_mblocks(int N, const float *__restrict__ o, const double *__restrict__ Minv,
                                                const uint8_t *__restrict__ fixed, float eps, double *__restrict__ M64,
                                                float *__restrict__ M32) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const double *Mi = Minv + 9 * (size_t)a;
    double M[9];
    if (node_fixed(fixed, a) || !(Mi[0] > 0.0)) {
        for (int m = 0; m < 9; ++m) M[m] = Mi[m];
    } else {
        double C[9];
        chol3(Mi, C);
        const float *oa = o + (size_t)a * kOut;
        const double L[9] = {exp(clampo(oa[0])), 0, 0, oa[1], exp(clampo(oa[2])), 0, oa[3], oa[4], exp(clampo(oa[5]))};
        double A[9], T1[9];
        for (int i = 0; i < 3; ++i)  // A = L L' + eps I
            for (int j = 0; j < 3; ++j)
                A[i * 3 + j] = L[i * 3] * L[j * 3] + L[i * 3 + 1] * L[j * 3 + 1] + L[i * 3 + 2] * L[j * 3 + 2] + (i == j ? eps : 0.0);
        for (int i = 0; i < 3; ++i)  // T1 = C A
            for (int j = 0; j < 3; ++j) T1[i * 3 + j] = C[i * 3] * A[j] + C[i * 3 + 1] * A[3 + j] + C[i * 3 + 2] * A[6 + j];
        for (int i = 0; i < 3; ++i)  // M = T1 C'
            for (int j = 0; j < 3; ++j) M[i * 3 + j] = T1[i * 3] * C[j * 3] + T1[i * 3 + 1] * C[j * 3 + 1] + T1[i * 3 + 2] * C[j * 3 + 2];
        for (int i = 0; i < 3; ++i)  // exact symmetry
            for (int j = 0; j < i; ++j) M[i * 3 + j] = M[j * 3 + i];
    }
    for (int m = 0; m < 9; ++m) {
        M64[9 * (size_t)a + m] = M[m];
        M32[9 * (size_t)a + m] = (float)M[m];
    }
}

// z = M r (node blocks)
__global__ __launch_bounds__(256) void k_block_apply(int N, const double *__restrict__ M, const double *__restrict__ r,
                                                    double *__restrict__ z) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const double *Ma = M + 9 * (size_t)a;
    const size_t o = 3 * (size_t)a;
    for (int i = 0; i < 3; ++i) z[o + i] = Ma[i * 3] * r[o] + Ma[i * 3 + 1] * r[o + 1] + Ma[i * 3 + 2] * r[o + 2];
}

// G_a += coef (r - c w)_a r_a'
__global__ __launch_bounds__(256) void k_accum_dm(int N, double coef, double c, const double *__restrict__ r,
                                                 const double *__restrict__ w, double *__restrict__ G) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const size_t o = 3 * (size_t)a;
    for (int i = 0; i < 3; ++i) {
        const double gi = coef * (r[o + i] - c * w[o + i]);
        for (int j = 0; j < 3; ++j) G[9 * (size_t)a + i * 3 + j] += gi * r[o + j];
    }
}

// grad[p] = sum over wave partials, fixed order
__global__ __launch_bounds__(256) void k_grad_reduce(int nparts, const float *__restrict__ gpart, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam) return;
    double s = 0.0;
    for (int w = 0; w < nparts; ++w) s += gpart[(size_t)w * kNParam + p];
    grad[p] = (float)s;
}

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }
constexpr int kTrainBlocks = 256;

// ef: 3 n_el doubles of per-element inputs (k_elem_features)
hipError_t launch_features(jfm_ctx *ctx, double *ef, float *feat, hipStream_t s) {
    const int N = (int)ctx->N, ne = (int)ctx->n_el;
    const double *Dover = ctx->minv_tangent ? ctx->Dtan : nullptr;
    switch (ctx->npe) {
    case 10:
        if (ne) k_elem_features<10><<<nblk(ne), kBlock, 0, s>>>(ne, ctx->conn, ctx->gf64, ctx->E, ctx->nu, ctx->fixed, ef);
        k_features<10><<<nblk(N), kBlock, 0, s>>>(N, ne, ctx->ngf, ctx->inc_ptr, ctx->inc, ef, ctx->gf64, ctx->E, ctx->fixed, ctx->finc_ptr, Dover, feat);
        break;
    case 4:
        if (ne) k_elem_features<4><<<nblk(ne), kBlock, 0, s>>>(ne, ctx->conn, ctx->gf64, ctx->E, ctx->nu, ctx->fixed, ef);
        k_features<4><<<nblk(N), kBlock, 0, s>>>(N, ne, ctx->ngf, ctx->inc_ptr, ctx->inc, ef, ctx->gf64, ctx->E, ctx->fixed, ctx->finc_ptr, Dover, feat);
        break;
    case 8:
        if (ne) k_elem_features<8><<<nblk(ne), kBlock, 0, s>>>(ne, ctx->conn, ctx->gf64, ctx->E, ctx->nu, ctx->fixed, ef);
        k_features<8><<<nblk(N), kBlock, 0, s>>>(N, ne, ctx->ngf, ctx->inc_ptr, ctx->inc, ef, ctx->gf64, ctx->E, ctx->fixed, ctx->finc_ptr, Dover, feat);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Inference: the same forward pass as k_mlp (same MFMA tiles, roundings and
// order), in persistent 4-wave blocks that stage the weights once; each wave
// walks its own 16-node tiles with wave-local barriers only, and its LDS holds
// no training buffers (about 39 KiB per block, four blocks per CU).
constexpr int kInfWaves = 4;
struct MlpInfLds {
    uint16_t W1[kHid][kXPad];
    uint16_t W2[kHid][kHid];
    uint16_t W3[16][kHid];
    float b1[kHid], b2[kHid], b3[16];
    float shift[kFeat], scale[kFeat];
    struct Wave {
        uint16_t X[kTile][kXPad];
        uint16_t H1[kTile][kHid], H2[kTile][kHid];
        float O[kTile][16];
    } w[kInfWaves];
};

__global__ __launch_bounds__(kInfWaves * 64) void k_mlp_infer(int N, const float *__restrict__ feat,
                                                             const uint8_t *__restrict__ blob,
                                                             float *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MlpInfLds &S = *reinterpret_cast<MlpInfLds *>(lds_raw);
    const LearnedBlob *B = reinterpret_cast<const LearnedBlob *>(blob);
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, r = l & 15, q = (l >> 4) * 4;
    for (int i = tid; i < kHid * kXPad; i += blockDim.x) {
        const int o = i / kXPad, c = i % kXPad;
        S.W1[o][c] = c < kFeat ? B->W1[o * kFeat + c] : (uint16_t)0;
    }
    for (int i = tid; i < kHid * kHid; i += blockDim.x) S.W2[i / kHid][i % kHid] = B->W2[i];
    for (int i = tid; i < 16 * kHid; i += blockDim.x) {
        const int o = i / kHid, c = i % kHid;
        S.W3[o][c] = o < kOut ? B->W3[o * kHid + c] : (uint16_t)0;
    }
    for (int i = tid; i < kHid; i += blockDim.x) { S.b1[i] = B->b1[i]; S.b2[i] = B->b2[i]; }
    for (int i = tid; i < 16; i += blockDim.x) S.b3[i] = i < kOut ? B->b3[i] : 0.f;
    for (int i = tid; i < kFeat; i += blockDim.x) { S.shift[i] = B->feat_shift[i]; S.scale[i] = B->feat_scale[i]; }
    __syncthreads();   // the only block-wide barrier
    MlpInfLds::Wave &W = S.w[wv];
    constexpr int kFL = kTile * kXPad / 64;
    const int fc = l % kXPad;
    float fv[kFL];
    auto fetch = [&](int n0) {
#pragma unroll
        for (int u = 0; u < kFL; ++u) {
            const int node = n0 + (l + 64 * u) / kXPad;
            fv[u] = (fc < kFeat && node < N) ? feat[(size_t)node * kFeat + fc] : 0.f;
        }
    };
    const int waves = gridDim.x * kInfWaves;
    const int first = (blockIdx.x * kInfWaves + wv) * kTile;
    if (first < N) fetch(first);
    for (int n0 = first; n0 < N; n0 += waves * kTile) {
#pragma unroll
        for (int u = 0; u < kFL; ++u) {
            const int m = (l + 64 * u) / kXPad, node = n0 + m;
            W.X[m][fc] = f2bf((fc < kFeat && node < N) ? (fv[u] - S.shift[fc]) * S.scale[fc] : 0.f);
        }
        if (n0 + waves * kTile < N) fetch(n0 + waves * kTile);   // next tile's features in flight
        wave_sync();
        {
            f32x4 acc[4];
            dense_fwd<kXPad, 4>(&W.X[0][0], kXPad, &S.W1[0][0], kXPad, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    W.H1[q + j][n] = f2bf(gelu(acc[t][j] + S.b1[n]));
                }
        }
        wave_sync();
        {
            f32x4 acc[4];
            dense_fwd<kHid, 4>(&W.H1[0][0], kHid, &S.W2[0][0], kHid, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int n = t * 16 + r;
                    W.H2[q + j][n] = f2bf(gelu(acc[t][j] + S.b2[n]));
                }
        }
        wave_sync();
        {
            f32x4 acc[1];
            dense_fwd<kHid, 1>(&W.H2[0][0], kHid, &S.W3[0][0], kHid, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) W.O[q + j][r] = acc[0][j] + S.b3[r];
        }
        wave_sync();
        for (int i = l; i < kTile * kOut; i += 64) {
            const int m = i / kOut, c = i % kOut, node = n0 + m;
            if (node < N) out[(size_t)node * kOut + c] = W.O[m][c];
        }
        wave_sync();
    }
}

hipError_t launch_mlp(jfm_ctx *ctx, const float *feat, const void *blob, float *out, const double *G,
                      const double *Minv, float *gpart, int *nwaves, hipStream_t s) {
    const int N = (int)ctx->N;
    const size_t lds = sizeof(MlpLds);
    const unsigned tiles = (unsigned)((N + kWaves * kTile - 1) / (kWaves * kTile));
    if (G) {
        const unsigned g = std::min<unsigned>(tiles, kTrainBlocks);
        *nwaves = (int)g * kWaves;
        k_mlp<true><<<g, kWaves * 64, lds, s>>>(N, feat, (const uint8_t *)blob, out, G, Minv, ctx->fixed, gpart);
    } else {
        // four resident blocks per CU on 256 CUs; each wave takes every 4096th tile
        const unsigned tiles16 = (unsigned)((N + kTile - 1) / kTile);
        const unsigned g = std::max(1u, std::min<unsigned>((tiles16 + kInfWaves - 1) / kInfWaves, 1024));
        k_mlp_infer<<<g, kInfWaves * 64, sizeof(MlpInfLds), s>>>(N, feat, (const uint8_t *)blob, out);
    }
    return hipGetLastError();
}

hipError_t launch_mblocks(jfm_ctx *ctx, const float *out, const double *Minv, float eps, double *M64, float *M32,
                          hipStream_t s) {
    const int N = (int)ctx->N;
    k_mblocks<<<nblk(N), kBlock, 0, s>>>(N, out, Minv, ctx->fixed, eps, M64, M32);
    return hipGetLastError();
}

}  // namespace jfem

// ==================================================================== API ===
static_assert(sizeof(LearnedBlob) == 24 + 2 * (kHid * kFeat + kHid * kHid + kOut * kHid) + 4 * (2 * kHid + kOut + 2 * kFeat + 1),
              "weight blob must be packed");
static_assert(sizeof(MlpLds) <= 64 * 1024, "network LDS image must fit 64 KiB");

extern "C" size_t jfm_learned_blob_bytes(void) { return sizeof(LearnedBlob); }
extern "C" int32_t jfm_learned_n_params(void) { return kNParam; }

#define LALLOC(ctx, ptr, T, n)                                                             \
    do {                                                                                   \
        if (!(ptr)) {                                                                      \
            (ptr) = jfm_alloc<T>((ctx), (n));                                              \
            if (!(ptr)) return jfm_fail((ctx), JFM_ENOMEM, "device allocation failed: " #ptr); \
        }                                                                                  \
    } while (0)

static int ensure_buffers(jfm_ctx *ctx) {
    const int64_t N = ctx->N;
    LALLOC(ctx, ctx->feat, float, N * kFeat);
    LALLOC(ctx, ctx->efeat, double, 3 * std::max<int64_t>(ctx->n_el, 1));
    LALLOC(ctx, ctx->lout, float, N * kOut);
    LALLOC(ctx, ctx->Mlearn64, double, N * 9);
    LALLOC(ctx, ctx->Mlearn32, float, N * 9);
    LALLOC(ctx, ctx->mlp_w, unsigned char, (int64_t)sizeof(LearnedBlob));
    return JFM_OK;
}

extern "C" int jfm_learned_set_weights(jfm_ctx *ctx, const void *blob, size_t bytes) {
    if (!ctx || !blob) return JFM_EINVAL;
    if (bytes != sizeof(LearnedBlob)) return jfm_fail(ctx, JFM_EINVAL, "weight blob has the wrong size");
    const LearnedBlob *b = (const LearnedBlob *)blob;
    if (b->magic != kBlobMagic || b->n_in != kFeat || b->n_h1 != kHid || b->n_h2 != kHid || b->n_out != kOut)
        return jfm_fail(ctx, JFM_EINVAL, "weight blob header does not match the 24-64-64-6 network");
    if (!(b->eps >= 0.0f)) return jfm_fail(ctx, JFM_EINVAL, "eps must be >= 0");
    jfem_learned_drop(ctx);
    ctx->mlp_host.assign((const unsigned char *)blob, (const unsigned char *)blob + bytes);
    ctx->has_mlp = true;
    ctx->learned_ready = false;
    // the device copy now, so that no solve waits on a host-to-device copy
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    LALLOC(ctx, ctx->mlp_w, unsigned char, (int64_t)sizeof(LearnedBlob));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->mlp_w, ctx->mlp_host.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    JFM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mlp_dev = true;
    return JFM_OK;
}

extern "C" int jfm_learned_get_weights(jfm_ctx *ctx, void *blob, size_t bytes) {
    if (!ctx || !blob) return JFM_EINVAL;
    if (!ctx->has_mlp) return jfm_fail(ctx, JFM_ESTATE, "no weights loaded");
    if (bytes != sizeof(LearnedBlob)) return jfm_fail(ctx, JFM_EINVAL, "weight blob has the wrong size");
    std::memcpy(blob, ctx->mlp_host.data(), bytes);
    return JFM_OK;
}

// features for the current mesh and boundary conditions
static int compute_features(jfm_ctx *ctx, hipStream_t s) {
    if (int rc = jfem_ensure_workspaces(ctx)) return rc;
    if (int rc = ensure_buffers(ctx)) return rc;
    hipError_t e = launch_features(ctx, ctx->efeat, ctx->feat, s);
    if (e != hipSuccess) return jfm_hip_fail(ctx, e, "launch_features");
    return JFM_OK;
}

// features, network and blocks, enqueued on stream s
static int learned_enqueue(jfm_ctx *ctx, hipStream_t s) {
    if (int rc = compute_features(ctx, s)) return rc;
    if (!ctx->mlp_dev) {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->mlp_w, ctx->mlp_host.data(), sizeof(LearnedBlob), hipMemcpyHostToDevice, s));
        ctx->mlp_dev = true;
    }
    int nw = 0;
    hipError_t e = launch_mlp(ctx, ctx->feat, ctx->mlp_w, ctx->lout, nullptr, (const double *)ctx->ws64.Minv, nullptr,
                              &nw, s);
    if (e != hipSuccess) return jfm_hip_fail(ctx, e, "launch_mlp");
    const float eps = ((const LearnedBlob *)ctx->mlp_host.data())->eps;
    e = launch_mblocks(ctx, ctx->lout, (const double *)ctx->ws64.Minv, eps, ctx->Mlearn64, ctx->Mlearn32, s);
    if (e != hipSuccess) return jfm_hip_fail(ctx, e, "launch_mblocks");
    return JFM_OK;
}

static int learned_events(jfm_ctx *ctx) {
    for (hipEvent_t *ev : {&ctx->ev_fork, &ctx->ev_l0, &ctx->ev_l1})
        if (!*ev) JFM_HIP(ctx, hipEventCreate(ev));
    return JFM_OK;
}

int jfem_learned_launch(jfm_ctx *ctx) {
    if (!ctx->has_mlp || !ctx->has_mesh || ctx->learned_ready || ctx->learned_pending) return JFM_OK;
    // buffers and block Jacobi first (they allocate, and the blocks are an input)
    if (int rc = jfem_ensure_workspaces(ctx)) return rc;
    if (int rc = ensure_buffers(ctx)) return rc;
    if (!ctx->stream2) JFM_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    if (int rc = learned_events(ctx)) return rc;
    JFM_HIP(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
    JFM_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
    JFM_HIP(ctx, hipEventRecord(ctx->ev_l0, ctx->stream2));
    int rc = learned_enqueue(ctx, ctx->stream2);
    if (rc == JFM_OK)
        if (hipError_t e = hipEventRecord(ctx->ev_l1, ctx->stream2)) rc = jfm_hip_fail(ctx, e, "learned launch");
    if (rc) {  // nothing half-launched is left to be joined later
        (void)hipStreamSynchronize(ctx->stream2);
        return rc;
    }
    ctx->learned_pending = true;
    return JFM_OK;
}

void jfem_learned_drop(jfm_ctx *ctx) {
    if (ctx->learned_pending && ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    ctx->learned_pending = false;
}

int jfem_learned_prepare(jfm_ctx *ctx, double *ms) {
    if (!ctx->has_mlp) return jfm_fail(ctx, JFM_ESTATE, "learned preconditioner requested but no weights are loaded");
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (ms) *ms = 0.0;
    if (ctx->learned_pending) {  // started beside earlier work: join it
        JFM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_l1, 0));
        JFM_HIP(ctx, hipEventSynchronize(ctx->ev_l1));
        float t = 0.f;
        JFM_HIP(ctx, hipEventElapsedTime(&t, ctx->ev_l0, ctx->ev_l1));
        if (ms) *ms = t;
        ctx->learned_pending = false;
        ctx->learned_ready = true;
        return JFM_OK;
    }
    if (ctx->learned_ready) return JFM_OK;
    if (int rc = learned_events(ctx)) return rc;
    JFM_HIP(ctx, hipEventRecord(ctx->ev_l0, ctx->stream));
    if (int rc = learned_enqueue(ctx, ctx->stream)) return rc;
    JFM_HIP(ctx, hipEventRecord(ctx->ev_l1, ctx->stream));
    JFM_HIP(ctx, hipEventSynchronize(ctx->ev_l1));
    float t = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&t, ctx->ev_l0, ctx->ev_l1));
    if (ms) *ms = t;
    ctx->learned_ready = true;
    return JFM_OK;
}

extern "C" int jfm_learned_features(jfm_ctx *ctx, float *feat) {
    if (!ctx) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = compute_features(ctx, ctx->stream)) return rc;
    if (feat) JFM_HIP(ctx, hipMemcpyAsync(feat, ctx->feat, sizeof(float) * ctx->N * kFeat, hipMemcpyDeviceToHost, ctx->stream));
    JFM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFM_OK;
}

extern "C" int jfm_learned_outputs(jfm_ctx *ctx, float *out, double *M) {
    if (!ctx) return JFM_EINVAL;
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    jfem_learned_drop(ctx);
    ctx->learned_ready = false;
    if (int rc = jfem_learned_prepare(ctx, nullptr)) return rc;
    if (out) JFM_HIP(ctx, hipMemcpyAsync(out, ctx->lout, sizeof(float) * ctx->N * kOut, hipMemcpyDeviceToHost, ctx->stream));
    if (M) JFM_HIP(ctx, hipMemcpyAsync(M, ctx->Mlearn64, sizeof(double) * ctx->N * 9, hipMemcpyDeviceToHost, ctx->stream));
    JFM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFM_OK;
}

extern "C" int jfm_learned_grad(jfm_ctx *ctx, int32_t n_pairs, const double *r, const double *e, int32_t coarse_nodes,
                                double *loss, float *grad) {
    if (!ctx || n_pairs <= 0 || !r || !e || !grad || coarse_nodes < 0) return JFM_EINVAL;
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    jfem_learned_drop(ctx);
    ctx->learned_ready = false;
    if (int rc = jfem_learned_prepare(ctx, nullptr)) return rc;
    // the blocks act inside the two-level preconditioner: z = M r + Q A_c^-1 Q' r
    const bool coarse = coarse_nodes > 0;
    if (coarse && !(ctx->coarse_ready && ctx->coarse_built_nodes == coarse_nodes && ctx->coarse_tag == 0))
        if (int rc = jfem_coarse_setup(ctx, JFM_OP_MATFREE, coarse_nodes, nullptr)) return rc;
    const int64_t n = ctx->ndof, N = ctx->N;
    hipStream_t s = ctx->stream;
    double *dr = jfm_alloc<double>(ctx, n), *de = jfm_alloc<double>(ctx, n), *dz = jfm_alloc<double>(ctx, n),
           *dw = jfm_alloc<double>(ctx, n), *G = jfm_alloc<double>(ctx, 9 * N);
    float *gpart = jfm_alloc<float>(ctx, (int64_t)kTrainBlocks * kWaves * kNParam), *dgrad = jfm_alloc<float>(ctx, kNParam);
    auto cleanup = [&] {
        for (void *p : {(void *)dr, (void *)de, (void *)dz, (void *)dw, (void *)G, (void *)gpart, (void *)dgrad}) jfm_free(ctx, p);
    };
    if (!dr || !de || !dz || !dw || !G || !gpart || !dgrad) {
        cleanup();
        return jfm_fail(ctx, JFM_ENOMEM, "device allocation failed in jfm_learned_grad");
    }
    int rc = JFM_OK;
    double lsum = 0.0;
    JFM_HIP(ctx, hipMemsetAsync(G, 0, sizeof(double) * 9 * N, s));
    for (int p = 0; p < n_pairs && rc == JFM_OK; ++p) {
        JFM_HIP(ctx, hipMemcpyAsync(dr, r + (size_t)p * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(de, e + (size_t)p * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
        k_block_apply<<<nblk(N), kBlock, 0, s>>>((int)N, ctx->Mlearn64, dr, dz);
        if (coarse) {
            int np = 0;
            hipError_t err = launch_coarse_correct<double>(ctx, dr, dz, nullptr, ctx->partials + kPartMisc, nullptr, &np);
            if (err != hipSuccess) { rc = jfm_hip_fail(ctx, err, "coarse correction"); break; }
        }
        if ((rc = jfem_apply(ctx, JFM_OP_MATFREE, dz, dw))) break;
        double zr = 0, zw = 0, er = 0;
        if ((rc = jfem_dot(ctx, dz, dr, n, &zr)) || (rc = jfem_dot(ctx, dz, dw, n, &zw)) || (rc = jfem_dot(ctx, de, dr, n, &er)))
            break;
        if (!(zw > 0.0) || !(er > 0.0)) {
            rc = jfm_fail(ctx, JFM_EINVAL, "training pair with non-positive energy (is e = K^-1 r on the free DOFs?)");
            break;
        }
        const double c = zr / zw;
        lsum += 1.0 - zr * c / er;
        k_accum_dm<<<nblk(N), kBlock, 0, s>>>((int)N, -2.0 * c / (er * n_pairs), c, dr, dw, G);
    }
    if (rc == JFM_OK) {
        int nw = 0;
        hipError_t err = launch_mlp(ctx, ctx->feat, ctx->mlp_w, ctx->lout, G, (const double *)ctx->ws64.Minv, gpart, &nw,
                                    ctx->stream);
        if (err == hipSuccess) {
            k_grad_reduce<<<nblk(kNParam), kBlock, 0, s>>>(nw, gpart, dgrad);
            err = hipGetLastError();
        }
        if (err != hipSuccess) rc = jfm_hip_fail(ctx, err, "launch_mlp(train)");
    }
    if (rc == JFM_OK) {
        hipError_t err = hipMemcpyAsync(grad, dgrad, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s);
        if (err == hipSuccess) err = hipStreamSynchronize(s);
        if (err != hipSuccess) rc = jfm_hip_fail(ctx, err, "gradient copy");
        if (loss) *loss = lsum / n_pairs;
    }
    (void)hipStreamSynchronize(s);
    cleanup();
    return rc;
}
