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

// Affine Tet10 (straight edges: constant invJ, linear strain and stress).
// The 4-point rule integrates the degree-2 integrand grad N_i : sigma exactly,
// so the product equals the exact integral, evaluated here from the linear
// fields' values at the 4 vertices with the P1 x P1 identity
//     int_T f g = V/20 (sum_k f_k g_k + (sum_k f_k)(sum_k g_k)).
// In barycentric coordinates L_0..L_3 (grad L_0 = -(1,1,1), grad L_j = e_j)
// the reference gradients at vertex k are sparse: corner i gives
// (4 delta_ik - 1) grad L_i, edge (a,b) gives 4 (delta_kb grad L_a +
// delta_ka grad L_b).  So the reference displacement gradient at vertex k is
// sum_j c_j (x) grad L_j with c_k = 3 u_k, c_j = 4 u_kj - u_j, and with
// P_k = invJ' sigma_k (w detJ folded in), r_k(j) = P_k grad L_j, S = sum_k P_k:
//     y_corner i = (4 r_i(i) - S grad L_i) / 5,
//     y_edge (a,b) = 4 (r_b(a) + r_a(b) + S grad L_a + S grad L_b) / 5
// (the 1/5 is 24 / 120: the rule's weight w = 1/24 sits in cl, cm).  About 40%
// fewer operations than the Gauss-point form below, same result up to rounding.
// FOLD: gf holds sqrt(cm) invJ (9 rows) and rho = cl / cm = lambda / mu is
// the same for every element; then G and sigma carry a factor sqrt(cm) and
// 1 / sqrt(cm), and P = invJ' sigma is unchanged (two fewer loads and
// multiplications per element).
template <typename T, bool FOLD = false, typename XL, typename Out>
__host__ __device__ __forceinline__ void mf_element_tet10_affine(int n_el, int e, const T *__restrict__ gf, XL &&xl,
                                                                 Out &&out, T rho = T(0)) {
    constexpr int EDG[4][4] = {{-1, 4, 6, 7}, {4, -1, 5, 8}, {6, 5, -1, 9}, {7, 8, 9, -1}};
    T u[10][3];
    T v0[3];
    xl(0, v0);
#pragma unroll
    for (int c = 0; c < 3; ++c) u[0][c] = T(0);
#pragma unroll
    for (int i = 1; i < 10; ++i) {
        T v[3];
        xl(i, v);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[i][c] = v[c] - v0[c];   // translation removed (FP32)
    }
    T Ji[9];
    auto gld = [&](int k) -> T {
        return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(gf) +
                                            (uint32_t)sizeof(T) * (uint32_t)(k * n_el + e));
    };
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji[k] = gld(k);
    const T cl = FOLD ? rho : T(0.2) * gld(9), cm = FOLD ? T(1) : T(0.2) * gld(10);
    T P[4][9];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T cv[4][3];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) cv[j][c] = (j == k) ? T(3) * u[k][c] : T(4) * u[EDG[k][j]][c] - u[j][c];
        T G[9];   // G[a][b] = sum_d H[a][d] Ji[b][d], H[a][d] = cv[d+1][a] - cv[0][a]
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T h0 = cv[1][a] - cv[0][a], h1 = cv[2][a] - cv[0][a], h2 = cv[3][a] - cv[0][a];
#pragma unroll
            for (int b = 0; b < 3; ++b) G[a * 3 + b] = h0 * Ji[b * 3] + h1 * Ji[b * 3 + 1] + h2 * Ji[b * 3 + 2];
        }
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
                P[k][d * 3 + c] = Ji[d] * s[c] + Ji[3 + d] * s[3 + c] + Ji[6 + d] * s[6 + c];
    }
    T S[9], r0[4][3], s0[3];
#pragma unroll
    for (int m = 0; m < 9; ++m) S[m] = P[0][m] + P[1][m] + P[2][m] + P[3][m];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s0[c] = -(S[c] + S[3 + c] + S[6 + c]);
#pragma unroll
        for (int k = 0; k < 4; ++k) r0[k][c] = -(P[k][c] + P[k][3 + c] + P[k][6 + c]);
    }
    // r_k(j): j = 0 -> r0[k], else row j-1 of P_k; likewise S grad L_j
    auto r = [&](int k, int j, int c) -> T { return j == 0 ? r0[k][c] : P[k][(j - 1) * 3 + c]; };
    auto sg = [&](int j, int c) -> T { return j == 0 ? s0[c] : S[(j - 1) * 3 + c]; };
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out(i, T(4) * r(i, i, 0) - sg(i, 0), T(4) * r(i, i, 1) - sg(i, 1), T(4) * r(i, i, 2) - sg(i, 2));
    constexpr int EA[6] = {0, 1, 2, 0, 1, 2}, EB[6] = {1, 2, 0, 3, 3, 3};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int a = EA[m], b = EB[m];
        T f[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] = T(4) * ((r(b, a, c) + r(a, b, c)) + (sg(a, c) + sg(b, c)));
        out(4 + m, f[0], f[1], f[2]);
    }
}

// Element e with node values read through xl(i, v): v = x at local node i;
// geometric factors gf[(q * GF + k) * n_el + e]; out(i, f0, f1, f2) receives
// node i's terms.
template <int NPE, typename T, bool AFF, typename XL, typename Out>
__host__ __device__ __forceinline__ void mf_element_gp(int n_el, int e, const T *__restrict__ gf, XL &&xl, Out &&out);

template <int NPE, typename T, bool AFF, bool FOLD = false, typename XL, typename Out>
__host__ __device__ __forceinline__ void mf_element_x(int n_el, int e, const T *__restrict__ gf, XL &&xl, Out &&out,
                                                      T rho = T(0)) {
    if constexpr (NPE == 10 && AFF)
        mf_element_tet10_affine<T, FOLD>(n_el, e, gf, xl, out, rho);
    else
        mf_element_gp<NPE, T, AFF>(n_el, e, gf, xl, out);
}

// The Gauss-point form: every element type, affine or not.
template <int NPE, typename T, bool AFF, typename XL, typename Out>
__host__ __device__ __forceinline__ void mf_element_gp(int n_el, int e, const T *__restrict__ gf, XL &&xl, Out &&out) {
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


// Element e with SoA connectivity conn[i * n_el + e] and x gathered from HBM
// (one three-component load per node).
template <int NPE, typename T, bool AFF, typename Out>
__device__ __forceinline__ void mf_element(int n_el, int e, const int32_t *__restrict__ conn, const T *__restrict__ gf,
                                           const T *__restrict__ x, Out &&out) {
    mf_element_x<NPE, T, AFF>(n_el, e, gf, [&](int i, T (&v)[3]) {
        const uint32_t a = (uint32_t)load_at<int32_t>(conn, 4u * (uint32_t)(i * n_el + e));
        const Trip<T> w = load_at<Trip<T>>(x, (uint32_t)(3 * sizeof(T)) * a);
        v[0] = w.a;
        v[1] = w.b;
        v[2] = w.c;
    }, out);
}

}  // namespace jfem
