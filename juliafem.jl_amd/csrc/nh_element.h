// nh_element.h -- the neo-Hookean total-Lagrangian element (gfx950): the
// Gauss-point state and its directional derivative, and the element's
// internal forces f_int,e(u) or tangent product K_T,e(u) v_e, for the
// two-phase kernels of nonlinear.hip.
//
// Reference: src/materials/neo_hookean.jl:129-137 (psi = mu/2 (I1 - 3) -
// mu ln J + lambda/2 ln^2 J), :205-231 (S = 2 dpsi/dC, D = 4 d2psi/dC2 by AD),
// src/problems_elasticity.jl:255-289 (BL(F)), :378-404 (geometric stiffness).
// The closed forms used here are
//     S  = mu A (C - I) + lambda ln J A,               A = C^-1
//     dS = 2 (mu - lambda ln J) A dE A + lambda tr(A dE) A,
// with dE = sym(F' grad v) and dP = grad v S + F dS (P = F S).
// Weighted by w detJ0 these use the stored cl = lambda w detJ, cm = mu w detJ.
#pragma once
#include "jfem_device.h"

namespace jfem {

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

// Element e (SoA connectivity conn[i * n_el + e], geometric factors
// gf[(q * GF + k) * n_el + e]) at the state u: f_int,e(u) (TANGENT = false)
// or K_T,e(u) v_e (TANGENT = true) into f (an output array rather than a
// callback, which cost the FP32 tangent kernel 38 VGPRs).
template <int NPE, typename T, bool AFF, bool TANGENT>
__device__ __forceinline__ void nh_element(int n_el, int e, const int32_t *__restrict__ conn, const T *__restrict__ gf,
                                           const T *__restrict__ u, const T *__restrict__ v, T (&f)[NPE][3]) {
    using EL = Elem<NPE>;
    constexpr int NG = EL::NGP;
    const int n0 = conn[e];
    const T u0[3] = {u[3 * (size_t)n0], u[3 * (size_t)n0 + 1], u[3 * (size_t)n0 + 2]};
    T v0[3] = {T(0), T(0), T(0)};
    if (TANGENT) {
        v0[0] = v[3 * (size_t)n0]; v0[1] = v[3 * (size_t)n0 + 1]; v0[2] = v[3 * (size_t)n0 + 2];
    }
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
}

}  // namespace jfem
