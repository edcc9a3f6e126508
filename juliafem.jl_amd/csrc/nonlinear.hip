// nonlinear.hip -- total-Lagrangian compressible neo-Hookean kernels (gfx950).
//
// Reference: src/materials/neo_hookean.jl:129-137 (psi = mu/2 (I1 - 3) -
// mu ln J + lambda/2 ln^2 J), :205-231 (S = 2 dpsi/dC, D = 4 d2psi/dC2 by AD),
// src/problems_elasticity.jl:255-289 (BL(F)), :378-404 (geometric stiffness),
// src/solvers.jl:575-620 (Newton).  The closed forms used here are
//     S  = mu A (C - I) + lambda ln J A,               A = C^-1
//     dS = 2 (mu - lambda ln J) A dE A + lambda tr(A dE) A,
// with dE = sym(F' grad v) and dP = grad v S + F dS (P = F S).
// Weighted by w detJ0 these use the stored cl = lambda w detJ, cm = mu w detJ.
// Element results go to the incidence slots like the linear operators, so
// the nodal sums run in ascending element id without atomics.
#include "jfem_device.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

template <typename T> struct Mat3 { T m[9]; };

// det(I + M) - 1 = I1(M) + I2(M) + I3(M), exact at small M
template <typename T>
__device__ __forceinline__ T det_ipm_minus1(const T *M) {
    const T i1 = M[0] + M[4] + M[8];
    const T i2 = M[0] * M[4] - M[1] * M[3] + M[4] * M[8] - M[5] * M[7] + M[0] * M[8] - M[2] * M[6];
    const T i3 = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
    return i1 + i2 + i3;
}

// State at one Gauss point from the reference displacement gradient Hu
// (sum_i (u_i - u_0) (x) dN_i) and invJ: F = I + H, A = C^-1, ln J and the
// weighted S = cm A (C - I) + cl ln J A, with C - I = H + H' + H'H so that S
// keeps its relative precision at small strain.
template <typename T>
__device__ __forceinline__ void nh_state(const T *Hu, const T *Ji, T cl, T cm, T *F, T *A, T *S, T &lnJ) {
    T H[9], E2[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) H[a * 3 + b] = Hu[a * 3] * Ji[b * 3] + Hu[a * 3 + 1] * Ji[b * 3 + 1] + Hu[a * 3 + 2] * Ji[b * 3 + 2];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = H[k] + (k % 4 == 0 ? T(1) : T(0));
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) {
            const T v = H[a * 3 + b] + H[b * 3 + a] + H[a] * H[b] + H[3 + a] * H[3 + b] + H[6 + a] * H[6 + b];
            E2[a * 3 + b] = v;
            E2[b * 3 + a] = v;
        }
    // J = det F from H (NAN ln J for an inverted point, which C alone cannot show)
    const T jm1 = det_ipm_minus1<T>(H);
    lnJ = log1p(jm1);
    const T J = T(1) + jm1;
    const T id = T(1) / (J * J);            // det C = J^2
    const T C0 = T(1) + E2[0], C4 = T(1) + E2[4], C8 = T(1) + E2[8];
    A[0] = (C4 * C8 - E2[5] * E2[7]) * id;
    A[1] = (E2[2] * E2[7] - E2[1] * C8) * id;
    A[2] = (E2[1] * E2[5] - E2[2] * C4) * id;
    A[4] = (C0 * C8 - E2[2] * E2[6]) * id;
    A[5] = (E2[2] * E2[3] - C0 * E2[5]) * id;
    A[8] = (C0 * C4 - E2[1] * E2[3]) * id;
    A[3] = A[1]; A[6] = A[2]; A[7] = A[5];
    const T sl = cl * lnJ;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) {
            const T ae = A[a * 3] * E2[b] + A[a * 3 + 1] * E2[3 + b] + A[a * 3 + 2] * E2[6 + b];
            const T ea = A[b * 3] * E2[a] + A[b * 3 + 1] * E2[3 + a] + A[b * 3 + 2] * E2[6 + a];
            const T v = T(0.5) * cm * (ae + ea) + sl * A[a * 3 + b];
            S[a * 3 + b] = v;
            S[b * 3 + a] = v;
        }
}

// weighted dP for a direction gradient G (grad v): dP = G S + F dS,
// dS = 2 (cm - cl ln J) A dE A + cl tr(A dE) A, dE = sym(F' G)
template <typename T>
__device__ __forceinline__ void nh_dP(const T *F, const T *A, const T *S, T lnJ, T cl, T cm, const T *G, T *dP) {
    T dE[9], AdE[9], dS[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)      // F' G, then symmetrise
#pragma unroll
        for (int b = 0; b < 3; ++b) dE[a * 3 + b] = F[a] * G[b] + F[3 + a] * G[3 + b] + F[6 + a] * G[6 + b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a + 1; b < 3; ++b) {
            const T s = T(0.5) * (dE[a * 3 + b] + dE[b * 3 + a]);
            dE[a * 3 + b] = s;
            dE[b * 3 + a] = s;
        }
    const T trAdE = A[0] * dE[0] + A[1] * dE[1] + A[2] * dE[2] + A[3] * dE[3] + A[4] * dE[4] + A[5] * dE[5] +
                    A[6] * dE[6] + A[7] * dE[7] + A[8] * dE[8];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) AdE[a * 3 + b] = A[a * 3] * dE[b] + A[a * 3 + 1] * dE[3 + b] + A[a * 3 + 2] * dE[6 + b];
    const T c1 = T(2) * (cm - cl * lnJ);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            dS[a * 3 + b] = c1 * (AdE[a * 3] * A[b] + AdE[a * 3 + 1] * A[3 + b] + AdE[a * 3 + 2] * A[6 + b]) + cl * trAdE * A[a * 3 + b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            dP[a * 3 + b] = G[a * 3] * S[b] + G[a * 3 + 1] * S[3 + b] + G[a * 3 + 2] * S[6 + b] +
                            F[a * 3] * dS[b] + F[a * 3 + 1] * dS[3 + b] + F[a * 3 + 2] * dS[6 + b];
}

// Element phase of the residual (TANGENT = false: y_e = f_int,e(u)) or the
// tangent product (TANGENT = true: y_e = K_T,e(u) v_e).
template <int NPE, typename T, bool AFF, bool TANGENT>
__global__ __launch_bounds__(256) void k_nh_elem(int n_el, const int32_t *__restrict__ conn,
                                                const int32_t *__restrict__ slot, const T *__restrict__ gf,
                                                const T *__restrict__ u, const T *__restrict__ v, T *__restrict__ ye,
                                                const int *__restrict__ done, const int32_t *__restrict__ elist,
                                                int n_items, int64_t vs = 0, int64_t ys = 0) {
    using EL = Elem<NPE>;
    constexpr int NG = EL::NGP;
    if (done && *done) return;
    const int idx = xcd_block() * blockDim.x + threadIdx.x;
    if (idx >= n_items) return;
    const int e = elist ? elist[idx] : idx;   // optional element subset
    if (TANGENT) v += blockIdx.y * vs;        // vector blockIdx.y of a batch
    ye += blockIdx.y * ys;
    const int n0 = conn[e];
    const T u0[3] = {u[3 * (size_t)n0], u[3 * (size_t)n0 + 1], u[3 * (size_t)n0 + 2]};
    T v0[3] = {T(0), T(0), T(0)};
    if (TANGENT) {
        v0[0] = v[3 * (size_t)n0]; v0[1] = v[3 * (size_t)n0 + 1]; v0[2] = v[3 * (size_t)n0 + 2];
    }
    T f[NPE][3];
#pragma unroll
    for (int i = 0; i < NPE; ++i) f[i][0] = f[i][1] = f[i][2] = T(0);
    T Ji[9], cl, cm;
    auto load = [&](int q) {
        const T *g = gf + (size_t)q * GF * n_el + e;
#pragma unroll
        for (int k = 0; k < 9; ++k) Ji[k] = g[(size_t)k * n_el];
        cl = g[(size_t)9 * n_el];
        cm = g[(size_t)10 * n_el];
    };
    if (AFF) load(0);
#pragma unroll
    for (int q = 0; q < NG; ++q) {
        if (!AFF) load(q);
        T Hu[9], Hv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) { Hu[k] = T(0); Hv[k] = T(0); }
#pragma unroll
        for (int i = 1; i < NPE; ++i) {
            const size_t nd = 3 * (size_t)conn[(size_t)i * n_el + e];
            const T a0 = u[nd] - u0[0], a1 = u[nd + 1] - u0[1], a2 = u[nd + 2] - u0[2];
            T b0 = T(0), b1 = T(0), b2 = T(0);
            if (TANGENT) { b0 = v[nd] - v0[0]; b1 = v[nd + 1] - v0[1]; b2 = v[nd + 2] - v0[2]; }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double d = EL::dN(q, i, c);
                if (d != 0.0) {
                    Hu[c] += a0 * T(d); Hu[3 + c] += a1 * T(d); Hu[6 + c] += a2 * T(d);
                    if (TANGENT) { Hv[c] += b0 * T(d); Hv[3 + c] += b1 * T(d); Hv[6 + c] += b2 * T(d); }
                }
            }
        }
        T F[9], A[9], S[9], lnJ;
        nh_state<T>(Hu, Ji, cl, cm, F, A, S, lnJ);
        T P[9];
        if (TANGENT) {
            T G[9];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) G[a * 3 + b] = Hv[a * 3] * Ji[b * 3] + Hv[a * 3 + 1] * Ji[b * 3 + 1] + Hv[a * 3 + 2] * Ji[b * 3 + 2];
            nh_dP<T>(F, A, S, lnJ, cl, cm, G, P);
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) P[a * 3 + b] = F[a * 3] * S[b] + F[a * 3 + 1] * S[3 + b] + F[a * 3 + 2] * S[6 + b];
        }
        // Q = P invJ: f_i[c] += sum_d Q[c][d] dN(q, i, d)
        T Q[9];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) Q[c * 3 + d] = P[c * 3] * Ji[d] + P[c * 3 + 1] * Ji[3 + d] + P[c * 3 + 2] * Ji[6 + d];
#pragma unroll
        for (int i = 0; i < NPE; ++i)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dd = EL::dN(q, i, d);
                if (dd != 0.0) {
                    f[i][0] += T(dd) * Q[d];
                    f[i][1] += T(dd) * Q[3 + d];
                    f[i][2] += T(dd) * Q[6 + d];
                }
            }
    }
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        T *out = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        out[0] = f[i][0];
        out[1] = f[i][1];
        out[2] = f[i][2];
    }
}

// Diagonal 3x3 blocks of the tangent K_T(u), summed over the node's
// incidences in ascending element id, inverted like the linear block Jacobi;
// the raw blocks go to D (N x 9) for the learned preconditioner's features.
template <int NPE>
__global__ __launch_bounds__(256) void k_nh_block_jacobi(int N, const int32_t *__restrict__ ptr,
                                                        const int32_t *__restrict__ inc, const int32_t *__restrict__ conn,
                                                        int n_el, int ngf, const double *__restrict__ gf,
                                                        const double *__restrict__ u, const uint8_t *__restrict__ fixed,
                                                        double *__restrict__ M64, float *__restrict__ M32,
                                                        double *__restrict__ D) {
    using EL = Elem<NPE>;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
        const int idx = inc[k];
        const int e = idx / NPE, i = idx - e * NPE;
        const size_t n0 = 3 * (size_t)conn[e];
        for (int q = 0; q < EL::NGP; ++q) {
            const double *G0 = gf + (size_t)(ngf == 1 ? 0 : q) * GF * n_el + e;
            double Ji[9];
            for (int m = 0; m < 9; ++m) Ji[m] = G0[(size_t)m * n_el];
            const double cl = G0[(size_t)9 * n_el], cm = G0[(size_t)10 * n_el];
            double Hu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = 1; j < NPE; ++j) {
                const size_t nd = 3 * (size_t)conn[(size_t)j * n_el + e];
                const double d0 = dn_rt<NPE>(q, j, 0), d1 = dn_rt<NPE>(q, j, 1), d2 = dn_rt<NPE>(q, j, 2);
                for (int c = 0; c < 3; ++c) {
                    const double uc = u[nd + c] - u[n0 + c];
                    Hu[c * 3] += uc * d0;
                    Hu[c * 3 + 1] += uc * d1;
                    Hu[c * 3 + 2] += uc * d2;
                }
            }
            double F[9], A[9], S[9], lnJ;
            nh_state<double>(Hu, Ji, cl, cm, F, A, S, lnJ);
            const double d0 = dn_rt<NPE>(q, i, 0), d1 = dn_rt<NPE>(q, i, 1), d2 = dn_rt<NPE>(q, i, 2);
            double g[3];
            for (int b = 0; b < 3; ++b) g[b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
            for (int kc = 0; kc < 3; ++kc) {   // column kc: unit direction e_kc at node i
                double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dP[9];
                for (int b = 0; b < 3; ++b) G[kc * 3 + b] = g[b];
                nh_dP<double>(F, A, S, lnJ, cl, cm, G, dP);
                for (int c = 0; c < 3; ++c) B[c * 3 + kc] += dP[c * 3] * g[0] + dP[c * 3 + 1] * g[1] + dP[c * 3 + 2] * g[2];
            }
        }
    }
    for (int m = 0; m < 9; ++m) D[9 * (size_t)a + m] = B[m];
    const uint8_t fx[3] = {fixed[3 * (size_t)a], fixed[3 * (size_t)a + 1], fixed[3 * (size_t)a + 2]};
    double M[9];
    if (!(B[0] + B[4] + B[8] > 0.0)) {
        for (int m = 0; m < 9; ++m) M[m] = 0.0;
    } else {
        for (int c = 0; c < 3; ++c)
            if (fx[c])
                for (int d = 0; d < 3; ++d) { B[c * 3 + d] = (c == d); B[d * 3 + c] = (c == d); }
        const double c00 = B[4] * B[8] - B[5] * B[7], c01 = B[5] * B[6] - B[3] * B[8], c02 = B[3] * B[7] - B[4] * B[6];
        const double id = 1.0 / (B[0] * c00 + B[1] * c01 + B[2] * c02);
        M[0] = c00 * id;
        M[1] = (B[2] * B[7] - B[1] * B[8]) * id;
        M[2] = (B[1] * B[5] - B[2] * B[4]) * id;
        M[3] = c01 * id;
        M[4] = (B[0] * B[8] - B[2] * B[6]) * id;
        M[5] = (B[2] * B[3] - B[0] * B[5]) * id;
        M[6] = c02 * id;
        M[7] = (B[1] * B[6] - B[0] * B[7]) * id;
        M[8] = (B[0] * B[4] - B[1] * B[3]) * id;
        // the tangent is symmetric; symmetrise the inverse against rounding
        for (int r = 0; r < 3; ++r)
            for (int c = r + 1; c < 3; ++c) {
                const double s = 0.5 * (M[r * 3 + c] + M[c * 3 + r]);
                M[r * 3 + c] = s;
                M[c * 3 + r] = s;
            }
        for (int c = 0; c < 3; ++c)
            if (fx[c])
                for (int d = 0; d < 3; ++d) { M[c * 3 + d] = 0.0; M[d * 3 + c] = 0.0; }
    }
    for (int m = 0; m < 9; ++m) {
        M64[9 * (size_t)a + m] = M[m];
        M32[9 * (size_t)a + m] = (float)M[m];
    }
}

// --------------------------------------------------------------- launchers ---
template <int NPE, typename T, bool TANGENT>
static void nh_elem(jfm_ctx *ctx, const T *gf, const T *u, const T *v, T *ye, const int *done, const int32_t *elist,
                    int m, int nv, int64_t vs, int64_t ys) {
    const int n = (int)ctx->n_el;
    const dim3 grid(xcd_grid(nblk(m)), nv);
    if (ctx->affine)
        k_nh_elem<NPE, T, true, TANGENT><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, gf, u, v, ye, done,
                                                                           elist, m, vs, ys);
    else
        k_nh_elem<NPE, T, false, TANGENT><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, gf, u, v, ye, done,
                                                                            elist, m, vs, ys);
}

template <typename T>
hipError_t launch_nh_elem(jfm_ctx *ctx, bool tangent, const T *gf, const T *u, const T *v, T *ye, const int *done,
                          const int32_t *elist, int n_items, int nv, int64_t vs, int64_t ys) {
    if (ctx->n_el == 0) return hipSuccess;
    const int m = elist ? n_items : (int)ctx->n_el;
    if (m == 0) return hipSuccess;
    switch (ctx->npe) {
    case 10: tangent ? nh_elem<10, T, true>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys) : nh_elem<10, T, false>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys); break;
    case 4: tangent ? nh_elem<4, T, true>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys) : nh_elem<4, T, false>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys); break;
    case 8: tangent ? nh_elem<8, T, true>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys) : nh_elem<8, T, false>(ctx, gf, u, v, ye, done, elist, m, nv, vs, ys); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_nh_block_jacobi(jfm_ctx *ctx, const double *u, double *M64, float *M32, double *D) {
    const int N = (int)ctx->N;
    switch (ctx->npe) {
    case 10: k_nh_block_jacobi<10><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, ctx->conn, (int)ctx->n_el, ctx->ngf, ctx->gf64, u, ctx->fixed, M64, M32, D); break;
    case 4: k_nh_block_jacobi<4><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, ctx->conn, (int)ctx->n_el, ctx->ngf, ctx->gf64, u, ctx->fixed, M64, M32, D); break;
    case 8: k_nh_block_jacobi<8><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, ctx->conn, (int)ctx->n_el, ctx->ngf, ctx->gf64, u, ctx->fixed, M64, M32, D); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_nh_elem<double>(jfm_ctx *, bool, const double *, const double *, const double *, double *, const int *,
                                           const int32_t *, int, int, int64_t, int64_t);
template hipError_t launch_nh_elem<float>(jfm_ctx *, bool, const float *, const float *, const float *, float *, const int *,
                                          const int32_t *, int, int, int64_t, int64_t);

}  // namespace jfem
