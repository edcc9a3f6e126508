// nonlinear.hip -- total-Lagrangian compressible neo-Hookean kernels (gfx950).
//
// The element (state, tangent, internal forces) is in nh_element.h;
// reference: src/materials/neo_hookean.jl:129-231,
// src/problems_elasticity.jl:255-289, :378-404, src/solvers.jl:575-620
// (Newton).  Element results go to the incidence slots like the linear
// operators, so the nodal sums run in ascending element id without atomics.
#include "jfem_device.h"
#include "nh_element.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// Element phase of the residual (TANGENT = false: y_e = f_int,e(u)) or the
// tangent product (TANGENT = true: y_e = K_T,e(u) v_e).
template <int NPE, typename T, bool AFF, bool TANGENT>
__global__ __launch_bounds__(256) void k_nh_elem(int n_el, const int32_t *__restrict__ conn,
                                                const int32_t *__restrict__ slot, const T *__restrict__ gf,
                                                const T *__restrict__ u, const T *__restrict__ v, T *__restrict__ ye,
                                                const int *__restrict__ done, const int32_t *__restrict__ elist,
                                                int n_items, int64_t vs = 0, int64_t ys = 0) {
    if (done && *done) return;
    const int idx = xcd_block() * blockDim.x + threadIdx.x;
    if (idx >= n_items) return;
    const int e = elist ? elist[idx] : idx;   // optional element subset
    if (TANGENT) v += blockIdx.y * vs;        // vector blockIdx.y of a batch
    ye += blockIdx.y * ys;
    T f[NPE][3];
    nh_element<NPE, T, AFF, TANGENT>(n_el, e, conn, gf, u, v, f);
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        T *o = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        o[0] = f[i][0];
        o[1] = f[i][1];
        o[2] = f[i][2];
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
