// mf_element.h -- the matrix-free element product y_e = K_e x_e (gfx950),
// shared by the two-phase operator (fem_kernels.hip) and the patch operator
// (patch.hip).
//
// Reference algorithm: src/problems_elasticity.jl:203-451 (Km += w BL' D BL,
// :369-375), applied to x_e without forming K_e.  H_q = sum_i (x_i - x_0) (x)
// dN_i(xi_q) (reference gradient; sum_i dN_i = 0, so subtracting node 0's
// value removes the element's translation, which in FP32 would otherwise
// cancel away the strain), grad x = H_q invJ', sigma_w = w detJ (lambda tr(eps)
// I + 2 mu eps), P_q = invJ' sigma_w, y_i = sum_q P_q' dN_i(xi_q).
#pragma once
#include "jfem_device.h"

namespace jfem {

// Element e with node values read through xl(i, v): v = x at local node i;
// geometric factors gf[(q * GF + k) * n_el + e]; out(i, f0, f1, f2) receives
// node i's terms.
template <int NPE, typename T, bool AFF, typename XL, typename Out>
__device__ __forceinline__ void mf_element_x(int n_el, int e, const T *__restrict__ gf, XL &&xl, Out &&out) {
    using EL = Elem<NPE>;
    constexpr int NG = EL::NGP;
    T H[NG][9];
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
        for (int k = 0; k < 9; ++k) H[q][k] = T(0);
    T v0[3];
    xl(0, v0);
#pragma unroll
    for (int i = 1; i < NPE; ++i) {
        T v[3];
        xl(i, v);
        const T u0 = v[0] - v0[0], u1 = v[1] - v0[1], u2 = v[2] - v0[2];
#pragma unroll
        for (int q = 0; q < NG; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double d = EL::dN(q, i, c);
                if (d != 0.0) {
                    H[q][c] += u0 * T(d);
                    H[q][3 + c] += u1 * T(d);
                    H[q][6 + c] += u2 * T(d);
                }
            }
    }
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
        T G[9];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                G[a * 3 + b] = H[q][a * 3] * Ji[b * 3] + H[q][a * 3 + 1] * Ji[b * 3 + 1] + H[q][a * 3 + 2] * Ji[b * 3 + 2];
        const T tr = cl * (G[0] + G[4] + G[8]);
        T s[9];
        s[0] = tr + T(2) * cm * G[0];
        s[4] = tr + T(2) * cm * G[4];
        s[8] = tr + T(2) * cm * G[8];
        s[1] = s[3] = cm * (G[1] + G[3]);
        s[2] = s[6] = cm * (G[2] + G[6]);
        s[5] = s[7] = cm * (G[5] + G[7]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                H[q][d * 3 + c] = Ji[d] * s[c] + Ji[3 + d] * s[3 + c] + Ji[6 + d] * s[6 + c];
    }
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        T f0 = T(0), f1 = T(0), f2 = T(0);
#pragma unroll
        for (int q = 0; q < NG; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dd = EL::dN(q, i, d);
                if (dd != 0.0) {
                    f0 += T(dd) * H[q][d * 3];
                    f1 += T(dd) * H[q][d * 3 + 1];
                    f2 += T(dd) * H[q][d * 3 + 2];
                }
            }
        out(i, f0, f1, f2);
    }
}

// Element e with SoA connectivity conn[i * n_el + e] and x gathered from HBM.
template <int NPE, typename T, bool AFF, typename Out>
__device__ __forceinline__ void mf_element(int n_el, int e, const int32_t *__restrict__ conn, const T *__restrict__ gf,
                                           const T *__restrict__ x, Out &&out) {
    mf_element_x<NPE, T, AFF>(n_el, e, gf, [&](int i, T (&v)[3]) {
        const size_t o = 3 * (size_t)conn[(size_t)i * n_el + e];
        v[0] = x[o];
        v[1] = x[o + 1];
        v[2] = x[o + 2];
    }, out);
}

}  // namespace jfem
