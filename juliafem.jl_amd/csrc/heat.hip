// heat.hip -- steady heat conduction on the uploaded mesh (Problem{Heat}).
//
// Reference: src/problems_heat.jl:36-66 (volume: Ke += w detJ k dN' dN,
// fe += w detJ N' f), :81-113 (boundary: fe += w detJ N' g for a heat flux g;
// Ke += w detJ h N' N, fe += w detJ N' h T_ext for convection).  The
// temperature is component 0 of the library's three-component node vectors,
// so the operator runs through the same incidence slots, gather and PCG as
// elasticity; components 1 and 2 are held at zero by the row mask.  The
// conduction operator is matrix-free from the stored geometric factors
// (w detJ = w / det invJ); the convection faces add their contributions in
// the gather from face-incidence slots, in ascending face order.
#include <vector>

#include "jfem_device.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ double det3x3(const double *A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// MODE 0: y_i = sum_q w detJ k grad N_i . grad T (all three components);
// MODE 1: source load y_i[0] = sum_q w detJ q N_i; MODE 2: diagonal k |grad N_i|^2
template <int NPE, bool AFF, int MODE>
__global__ __launch_bounds__(256) void k_heat_elem(int n_el, const int32_t *__restrict__ conn,
                                                  const int32_t *__restrict__ slot, const double *__restrict__ gf,
                                                  const double *__restrict__ kc, const double *__restrict__ x,
                                                  double *__restrict__ ye, const int *__restrict__ done) {
    using EL = Elem<NPE>;
    if (done && *done) return;
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    const double ke = kc[e];
    double y[NPE][3];
#pragma unroll
    for (int i = 0; i < NPE; ++i) y[i][0] = y[i][1] = y[i][2] = 0.0;
    double xe[NPE][3];
    if (MODE == 0) {
        const size_t n0 = 3 * (size_t)conn[e];
        const double t0 = x[n0], t1 = x[n0 + 1], t2 = x[n0 + 2];
#pragma unroll
        for (int i = 0; i < NPE; ++i) {
            const size_t nd = 3 * (size_t)conn[(size_t)i * n_el + e];
            xe[i][0] = x[nd] - t0; xe[i][1] = x[nd + 1] - t1; xe[i][2] = x[nd + 2] - t2;  // sum grad N = 0
        }
    }
    double Ji[9];
    for (int q = 0; q < EL::NGP; ++q) {
        if (AFF ? q == 0 : true) {
            const double *g = gf + (size_t)(AFF ? 0 : q) * GF * n_el + e;
            for (int k = 0; k < 9; ++k) Ji[k] = g[(size_t)k * n_el];
        }
        const double wd = EL::w(q) / det3x3(Ji);
        double gr[NPE][3];
#pragma unroll
        for (int i = 0; i < NPE; ++i) {
            const double d0 = EL::dN(q, i, 0), d1 = EL::dN(q, i, 1), d2 = EL::dN(q, i, 2);
            for (int b = 0; b < 3; ++b) gr[i][b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
        }
        if (MODE == 0) {
            double gT[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // gT[c][b] = d x_c / dX_b
#pragma unroll
            for (int j = 1; j < NPE; ++j)
                for (int c = 0; c < 3; ++c)
                    for (int b = 0; b < 3; ++b) gT[c][b] += xe[j][c] * gr[j][b];
            const double s = wd * ke;
#pragma unroll
            for (int i = 0; i < NPE; ++i)
                for (int c = 0; c < 3; ++c) y[i][c] += s * (gr[i][0] * gT[c][0] + gr[i][1] * gT[c][1] + gr[i][2] * gT[c][2]);
        } else if (MODE == 1) {
#pragma unroll
            for (int i = 0; i < NPE; ++i) y[i][0] += wd * ke * EL::N(q, i);  // ke holds the source here
        } else {
#pragma unroll
            for (int i = 0; i < NPE; ++i) {
                const double v = wd * ke * (gr[i][0] * gr[i][0] + gr[i][1] * gr[i][1] + gr[i][2] * gr[i][2]);
                y[i][0] += v; y[i][1] += v; y[i][2] += v;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        double *o = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        o[0] = y[i][0]; o[1] = y[i][1]; o[2] = y[i][2];
    }
}

// Boundary faces.  MODE 0: convection y_i = sum_q w detJ h N_i sum_j N_j x_j
// (three components); MODE 1: load y_i[0] = sum_q w detJ (g + h T_ext) N_i;
// MODE 2: diagonal h N_i^2.  Output to the face slots fe[(f * nfn + i) * 3].
template <int MODE>
__global__ __launch_bounds__(256) void k_heat_face(int nf, int nfn, const int32_t *__restrict__ fconn,
                                                  const double *__restrict__ X, const double *__restrict__ h,
                                                  const double *__restrict__ g, const double *__restrict__ text,
                                                  const double *__restrict__ x, double *__restrict__ fe,
                                                  const int *__restrict__ done) {
    if (done && *done) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double Xf[6][3], xf[6][3], out[6][3];
    for (int i = 0; i < nfn; ++i) {
        const size_t nd = 3 * (size_t)fconn[(size_t)i * nf + f];
        for (int c = 0; c < 3; ++c) {
            Xf[i][c] = X[nd + c];
            xf[i][c] = MODE == 0 ? x[nd + c] : 0.0;
            out[i][c] = 0.0;
        }
    }
    const double hf = h ? h[f] : 0.0;
    const double load = (g ? g[f] : 0.0) + hf * (text ? text[f] : 0.0);
    double w[4], pu[4], pv[4];
    const int ng = surf_rule(nfn, w, pu, pv);
    for (int q = 0; q < ng; ++q) {
        double N[6], dN[6][2];
        surf_basis(nfn, pu[q], pv[q], N, dN);
        double J[2][3] = {{0, 0, 0}, {0, 0, 0}};
        for (int i = 0; i < nfn; ++i)
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 3; ++b) J[a][b] += dN[i][a] * Xf[i][b];
        const double cx = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        const double cy = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        const double cz = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double wq = w[q] * sqrt(cx * cx + cy * cy + cz * cz);
        if (MODE == 0) {
            double s[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < nfn; ++j)
                for (int c = 0; c < 3; ++c) s[c] += N[j] * xf[j][c];
            for (int i = 0; i < nfn; ++i)
                for (int c = 0; c < 3; ++c) out[i][c] += wq * hf * N[i] * s[c];
        } else if (MODE == 1) {
            for (int i = 0; i < nfn; ++i) out[i][0] += wq * load * N[i];
        } else {
            for (int i = 0; i < nfn; ++i)
                for (int c = 0; c < 3; ++c) out[i][c] += wq * hf * N[i] * N[i];
        }
    }
    for (int i = 0; i < nfn; ++i)
        for (int c = 0; c < 3; ++c) fe[((size_t)f * nfn + i) * 3 + c] = out[i][c];
}

// y_a = sum (element slots of a) + sum (face slots of a), rows masked;
// optional partials of y . dotv.  The same fixed order as k_gather.
__global__ __launch_bounds__(256) void k_heat_gather(int N, const int32_t *__restrict__ ptr, const double *__restrict__ yinc,
                                                    const int32_t *__restrict__ fptr, const int32_t *__restrict__ finc,
                                                    const double *__restrict__ fe, const uint8_t *__restrict__ fixed,
                                                    double *__restrict__ y, const double *__restrict__ dotv,
                                                    double *__restrict__ part, const int *__restrict__ done) {
    if (done && *done) return;
    double dv = 0.0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += gridDim.x * blockDim.x) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
            s0 += yinc[3 * (size_t)k]; s1 += yinc[3 * (size_t)k + 1]; s2 += yinc[3 * (size_t)k + 2];
        }
        if (fptr)
            for (int k = fptr[a]; k < fptr[a + 1]; ++k) {
                const double *v = fe + 3 * (size_t)finc[k];
                s0 += v[0]; s1 += v[1]; s2 += v[2];
            }
        const size_t o = 3 * (size_t)a;
        if (fixed) {
            if (fixed[o]) s0 = 0.0;
            if (fixed[o + 1]) s1 = 0.0;
            if (fixed[o + 2]) s2 = 0.0;
        }
        y[o] = s0; y[o + 1] = s1; y[o + 2] = s2;
        if (dotv) dv += s0 * dotv[o] + s1 * dotv[o + 1] + s2 * dotv[o + 2];
    }
    if (part) {
        double v1[1] = {dv};
        write_partials<1>(v1, part);
    }
}

// Jacobi blocks diag(1/d) of the temperature component (0 where fixed)
__global__ __launch_bounds__(256) void k_heat_jacobi(int N, const double *__restrict__ d, const uint8_t *__restrict__ fixed,
                                                    double *__restrict__ Minv) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double *M = Minv + 9 * (size_t)a;
    for (int k = 0; k < 9; ++k) M[k] = 0.0;
    if (!fixed[3 * (size_t)a] && d[3 * (size_t)a] > 0.0) M[0] = 1.0 / d[3 * (size_t)a];
}

template <int MODE>
static hipError_t heat_elem(jfm_ctx *ctx, const double *kc, const double *x, const int *done) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    const unsigned g = xcd_grid(nblk(n));
#define HEAT(NPE)                                                                                                   \
    if (ctx->affine)                                                                                                \
        k_heat_elem<NPE, true, MODE><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->gf64, kc, x, ctx->ye64, done); \
    else                                                                                                            \
        k_heat_elem<NPE, false, MODE><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->gf64, kc, x, ctx->ye64, done);
    switch (ctx->npe) {
    case 10: HEAT(10) break;
    case 4: HEAT(4) break;
    case 8: HEAT(8) break;
    default: return hipErrorInvalidValue;
    }
#undef HEAT
    return hipGetLastError();
}

template <int MODE>
static hipError_t heat_face(jfm_ctx *ctx, const double *x, const int *done) {
    const int nf = (int)ctx->hf_n;
    if (nf == 0) return hipSuccess;
    k_heat_face<MODE><<<nblk(nf), kBlock, 0, ctx->stream>>>(nf, ctx->hf_nfn, ctx->hf_conn, ctx->X, ctx->hf_h, ctx->hf_g,
                                                            ctx->hf_text, x, ctx->hf_fe, done);
    return hipGetLastError();
}

// faces: add the face slots (loads always; operator and diagonal only with convection)
static hipError_t heat_gather(jfm_ctx *ctx, bool faces, const uint8_t *fixed, double *y, const double *dotv,
                              double *part, const int *done, int *npart) {
    const unsigned g = capped_grid(ctx->N, kBlock);
    if (npart) *npart = (int)g;
    const bool f = faces && ctx->hf_n > 0;
    k_heat_gather<<<g, kBlock, 0, ctx->stream>>>((int)ctx->N, ctx->inc_ptr, ctx->ye64, f ? ctx->hf_ptr : nullptr,
                                                 ctx->hf_inc, ctx->hf_fe, fixed, y, dotv, part, done);
    return hipGetLastError();
}

// y = K_heat x (masked by ctx->fixed), the conduction operator of kOpHeat
hipError_t launch_heat_apply(jfm_ctx *ctx, const double *x, double *y, const double *dotv, double *part, const int *done,
                             int *npart) {
    const bool conv = ctx->hf_n > 0 && ctx->hf_h;
    hipError_t e = heat_elem<0>(ctx, ctx->heat_k, x, done);
    if (e == hipSuccess && conv) e = heat_face<0>(ctx, x, done);
    if (e != hipSuccess) return e;
    return heat_gather(ctx, conv, ctx->fixed, y, dotv, part, done, npart);
}

// load vector (component 0): sources and boundary flux/convection terms
hipError_t launch_heat_load(jfm_ctx *ctx, const double *src, double *f) {
    hipError_t e = hipSuccess;
    if (src) e = heat_elem<1>(ctx, src, nullptr, nullptr);
    else e = hipMemsetAsync(ctx->ye64, 0, sizeof(double) * 3 * (size_t)ctx->npe * ctx->n_el, ctx->stream);
    if (e == hipSuccess) e = heat_face<1>(ctx, nullptr, nullptr);
    if (e != hipSuccess) return e;
    return heat_gather(ctx, true, nullptr, f, nullptr, nullptr, nullptr, nullptr);
}

// Jacobi blocks of the heat operator (ctx->fixed is the heat mask)
hipError_t launch_heat_jacobi(jfm_ctx *ctx, double *Minv) {
    const bool conv = ctx->hf_n > 0 && ctx->hf_h;
    hipError_t e = heat_elem<2>(ctx, ctx->heat_k, nullptr, nullptr);
    if (e == hipSuccess && conv) e = heat_face<2>(ctx, nullptr, nullptr);
    if (e == hipSuccess) e = heat_gather(ctx, conv, nullptr, ctx->vy64, nullptr, nullptr, nullptr, nullptr);
    if (e != hipSuccess) return e;
    k_heat_jacobi<<<nblk(ctx->N), kBlock, 0, ctx->stream>>>((int)ctx->N, ctx->vy64, ctx->fixed, Minv);
    return hipGetLastError();
}

}  // namespace jfem
