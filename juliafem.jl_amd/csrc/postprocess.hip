// postprocess.hip -- Gauss-point strain and stress, nodal least-squares fit.
//
// Reference: src/problems_elasticity.jl:520-548 (get_strain_vector /
// get_stress_vector: small strain eps = (grad u + grad u')/2, linear
// isotropic stress, ABAQUS order 11, 22, 33, 12, 23, 13 with tensor shear),
// :550-594 (lsq_fit: M x = b with M = sum w N N', b = sum w f N, solved for
// every node).  Here M x = b is solved by PCG with the consistent mass
// operator applied matrix-free through the same incidence slots and gather
// as K, three components at a time, diagonal preconditioner, on the fit rule
// described below.
#include "jfem_device.h"

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

template <int NPE>
__device__ __forceinline__ void load_invj(const double *__restrict__ gf, int n_el, int e, int set, double (&Ji)[9]) {
    const double *g = gf + (size_t)set * GF * n_el + e;
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji[k] = g[(size_t)k * n_el];
}

__device__ __forceinline__ double det9(const double (&A)[9]) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// strain and stress at every Gauss point: out[(e * NGP + q) * 6 + k]
template <int NPE, bool AFF>
__global__ __launch_bounds__(256) void k_post_gp(int n_el, const int32_t *__restrict__ conn, const double *__restrict__ gf,
                                                const double *__restrict__ E, const double *__restrict__ nu,
                                                const double *__restrict__ u, double *__restrict__ strain,
                                                double *__restrict__ stress) {
    using EL = Elem<NPE>;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    const double Ee = E[e], ne = nu[e];
    const double la = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne)), mu = Ee / (2.0 * (1.0 + ne));
    const size_t n0 = 3 * (size_t)conn[e];
    double Ji[9];
    if (AFF) load_invj<NPE>(gf, n_el, e, 0, Ji);
    for (int q = 0; q < EL::NGP; ++q) {
        if (!AFF) load_invj<NPE>(gf, n_el, e, q, Ji);
        double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // H[a][k] = sum_i (u_i - u_0)[a] dN_i/dxi_k
        for (int i = 1; i < NPE; ++i) {
            const size_t nd = 3 * (size_t)conn[(size_t)i * n_el + e];
            const double d0 = dn_rt<NPE>(q, i, 0), d1 = dn_rt<NPE>(q, i, 1), d2 = dn_rt<NPE>(q, i, 2);
            for (int a = 0; a < 3; ++a) {
                const double v = u[nd + a] - u[n0 + a];
                H[a * 3] += v * d0;
                H[a * 3 + 1] += v * d1;
                H[a * 3 + 2] += v * d2;
            }
        }
        double G[9];  // grad u: G[a][b] = du_a/dX_b
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) G[a * 3 + b] = H[a * 3] * Ji[b * 3] + H[a * 3 + 1] * Ji[b * 3 + 1] + H[a * 3 + 2] * Ji[b * 3 + 2];
        const double ev[6] = {G[0], G[4], G[8], 0.5 * (G[1] + G[3]), 0.5 * (G[5] + G[7]), 0.5 * (G[2] + G[6])};
        const double tr = ev[0] + ev[1] + ev[2];
        double *so = strain + ((size_t)e * EL::NGP + q) * 6, *to = stress + ((size_t)e * EL::NGP + q) * 6;
        for (int k = 0; k < 6; ++k) {
            so[k] = ev[k];
            to[k] = 2.0 * mu * ev[k] + (k < 3 ? la * tr : 0.0);
        }
    }
}

// ---- quadrature of the least-squares fit ---------------------------------
// The reference integrates the fit with the element's own rule
// (src/problems_elasticity.jl:555-560); for Tet10 (4 points) and Tet4 (1)
// that leaves the fit matrix singular, so a constant field is not recovered.
// Tetrahedra use a collapsed (Duffy) Gauss product rule instead: 3 x 3 x 4
// Gauss-Legendre points mapped by xi = a(1-b)(1-c), eta = b(1-c), zeta = c
// with weight factor (1-b)(1-c)^2, exact to degree 4 (the Tet10 mass entries);
// Hex8 keeps its 2 x 2 x 2 rule, which is exact for its mass entries.
__device__ __forceinline__ void gauss3(int k, double &x, double &w) {   // on [0, 1]
    const double g = 0.77459666924148337704, w0 = 8.0 / 9.0, w1 = 5.0 / 9.0;
    x = k == 0 ? 0.5 * (1.0 - g) : (k == 1 ? 0.5 : 0.5 * (1.0 + g));
    w = 0.5 * (k == 1 ? w0 : w1);
}
__device__ __forceinline__ void gauss4(int k, double &x, double &w) {
    const double g0 = 0.33998104358485626480, g1 = 0.86113631159405257522;
    const double w0 = 0.65214515486254614263, w1 = 0.34785484513745385737;
    const double t[4] = {-g1, -g0, g0, g1}, ww[4] = {w1, w0, w0, w1};
    x = 0.5 * (1.0 + t[k]);
    w = 0.5 * ww[k];
}

template <int NPE> struct LsqRule;
template <> struct LsqRule<10> { static constexpr int NP = 36; };
template <> struct LsqRule<4> { static constexpr int NP = 36; };
template <> struct LsqRule<8> { static constexpr int NP = 8; };

template <int NPE>
__device__ __forceinline__ void lsq_point(int p, double (&xi)[3], double &w) {
    if constexpr (NPE == 8) {
        for (int c = 0; c < 3; ++c) xi[c] = Elem<8>::xi(p, c);
        w = Elem<8>::w(p);
        return;
    }
    const int ia = p / 12, ib = (p / 4) % 3, ic = p % 4;
    double a, b, c, wa, wb, wc;
    gauss3(ia, a, wa);
    gauss3(ib, b, wb);
    gauss4(ic, c, wc);
    xi[0] = a * (1.0 - b) * (1.0 - c);
    xi[1] = b * (1.0 - c);
    xi[2] = c;
    w = wa * wb * wc * (1.0 - b) * (1.0 - c) * (1.0 - c);
}

// N and dN/dxi at an arbitrary reference point (the formulas of Elem<NPE>)
template <int NPE>
__device__ __forceinline__ void shape_at(const double (&xi)[3], double (&N)[NPE], double (&dN)[NPE][3]) {
    const double u = xi[0], v = xi[1], s = xi[2];
    if constexpr (NPE == 10) {
        const double l = 1.0 - u - v - s;
        N[0] = l * (2.0 * l - 1.0); N[1] = u * (2.0 * u - 1.0); N[2] = v * (2.0 * v - 1.0); N[3] = s * (2.0 * s - 1.0);
        N[4] = 4.0 * u * l; N[5] = 4.0 * u * v; N[6] = 4.0 * v * l; N[7] = 4.0 * s * l; N[8] = 4.0 * u * s;
        N[9] = 4.0 * v * s;
        const double d0 = 1.0 - 4.0 * l;
        const double t[10][3] = {{d0, d0, d0}, {4.0 * u - 1.0, 0.0, 0.0}, {0.0, 4.0 * v - 1.0, 0.0}, {0.0, 0.0, 4.0 * s - 1.0},
                                 {4.0 * (l - u), -4.0 * u, -4.0 * u}, {4.0 * v, 4.0 * u, 0.0}, {-4.0 * v, 4.0 * (l - v), -4.0 * v},
                                 {-4.0 * s, -4.0 * s, 4.0 * (l - s)}, {4.0 * s, 0.0, 4.0 * u}, {0.0, 4.0 * s, 4.0 * v}};
        for (int i = 0; i < NPE; ++i)
            for (int c = 0; c < 3; ++c) dN[i][c] = t[i][c];
    } else if constexpr (NPE == 4) {
        N[0] = 1.0 - u - v - s; N[1] = u; N[2] = v; N[3] = s;
        for (int i = 0; i < NPE; ++i)
            for (int c = 0; c < 3; ++c) dN[i][c] = i == 0 ? -1.0 : ((i - 1) == c ? 1.0 : 0.0);
    } else {
        for (int i = 0; i < NPE; ++i) {
            const double r0 = Elem<8>::ref(i, 0), r1 = Elem<8>::ref(i, 1), r2 = Elem<8>::ref(i, 2);
            const double a = 1.0 + r0 * u, b = 1.0 + r1 * v, d = 1.0 + r2 * s;
            N[i] = 0.125 * a * b * d;
            dN[i][0] = 0.125 * r0 * b * d;
            dN[i][1] = 0.125 * a * r1 * d;
            dN[i][2] = 0.125 * a * b * r2;
        }
    }
}

// mode 0: consistent mass M x (three components), mode 1: fit right-hand side
// sum_p w detJ f_p[comp0 + c] N_i with f the strain (field 0) or stress of u,
// mode 2: diagonal of M; into the incidence slots.  Geometry from X at every
// point of the fit rule, so curved elements need no stored factors.
template <int NPE, int MODE>
__global__ __launch_bounds__(256) void k_lsq_elem(int n_el, const int32_t *__restrict__ conn,
                                                 const int32_t *__restrict__ slot, const double *__restrict__ X,
                                                 const double *__restrict__ E, const double *__restrict__ nu,
                                                 const double *__restrict__ u, int field, int comp0,
                                                 const double *__restrict__ x, double *__restrict__ ye,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double Xe[NPE][3], ve[NPE][3], y[NPE][3];
    for (int i = 0; i < NPE; ++i) {
        const size_t nd = 3 * (size_t)conn[(size_t)i * n_el + e];
        for (int c = 0; c < 3; ++c) {
            Xe[i][c] = X[nd + c];
            ve[i][c] = MODE == 0 ? x[nd + c] : (MODE == 1 ? u[nd + c] : 0.0);
            y[i][c] = 0.0;
        }
    }
    double la = 0.0, mu = 0.0;
    if (MODE == 1 && field == 1) {
        const double Ee = E[e], ne = nu[e];
        la = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
        mu = Ee / (2.0 * (1.0 + ne));
    }
    for (int p = 0; p < LsqRule<NPE>::NP; ++p) {
        double xi[3], wp, N[NPE], dN[NPE][3];
        lsq_point<NPE>(p, xi, wp);
        shape_at<NPE>(xi, N, dN);
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < NPE; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) J[a * 3 + b] += dN[i][a] * Xe[i][b];
        const double det = det9(J);
        const double w = wp * det;
        if (MODE == 0) {
            double s3[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < NPE; ++j)
                for (int c = 0; c < 3; ++c) s3[c] += N[j] * ve[j][c];
            for (int i = 0; i < NPE; ++i)
                for (int c = 0; c < 3; ++c) y[i][c] += w * N[i] * s3[c];
        } else if (MODE == 2) {
            for (int i = 0; i < NPE; ++i)
                for (int c = 0; c < 3; ++c) y[i][c] += w * N[i] * N[i];
        } else {
            // inverse Jacobian, grad u = sum_i u_i (x) invJ dN_i (translation removed)
            double Ji[9];
            const double id = 1.0 / det;
            Ji[0] = (J[4] * J[8] - J[5] * J[7]) * id; Ji[1] = (J[2] * J[7] - J[1] * J[8]) * id; Ji[2] = (J[1] * J[5] - J[2] * J[4]) * id;
            Ji[3] = (J[5] * J[6] - J[3] * J[8]) * id; Ji[4] = (J[0] * J[8] - J[2] * J[6]) * id; Ji[5] = (J[2] * J[3] - J[0] * J[5]) * id;
            Ji[6] = (J[3] * J[7] - J[4] * J[6]) * id; Ji[7] = (J[1] * J[6] - J[0] * J[7]) * id; Ji[8] = (J[0] * J[4] - J[1] * J[3]) * id;
            double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 1; i < NPE; ++i) {
                double g[3];
                for (int b = 0; b < 3; ++b) g[b] = Ji[b * 3] * dN[i][0] + Ji[b * 3 + 1] * dN[i][1] + Ji[b * 3 + 2] * dN[i][2];
                for (int c = 0; c < 3; ++c) {
                    const double d = ve[i][c] - ve[0][c];
                    for (int b = 0; b < 3; ++b) G[c * 3 + b] += d * g[b];
                }
            }
            double f[6] = {G[0], G[4], G[8], 0.5 * (G[1] + G[3]), 0.5 * (G[5] + G[7]), 0.5 * (G[2] + G[6])};
            if (field == 1) {
                const double tr = f[0] + f[1] + f[2];
                for (int k = 0; k < 6; ++k) f[k] = 2.0 * mu * f[k] + (k < 3 ? la * tr : 0.0);
            }
            for (int i = 0; i < NPE; ++i)
                for (int c = 0; c < 3; ++c) y[i][c] += w * N[i] * f[comp0 + c];
        }
    }
    for (int i = 0; i < NPE; ++i) {
        double *o = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        o[0] = y[i][0]; o[1] = y[i][1]; o[2] = y[i][2];
    }
}

// Minv blocks diag(1 / d) from the gathered diagonal d (N x 3)
__global__ __launch_bounds__(256) void k_diag_blocks(int N, const double *__restrict__ d, double *__restrict__ Minv) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double *M = Minv + 9 * (size_t)a;
    for (int k = 0; k < 9; ++k) M[k] = 0.0;
    for (int c = 0; c < 3; ++c) M[4 * c] = 1.0 / d[3 * (size_t)a + c];
}

template <int MODE>
static hipError_t lsq_dispatch(jfm_ctx *ctx, const double *u, int field, int comp0, const double *x, const int *done) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    const unsigned g = xcd_grid(nblk(n));
    switch (ctx->npe) {
    case 10: k_lsq_elem<10, MODE><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->X, ctx->E, ctx->nu, u, field, comp0, x, ctx->ye64, done); break;
    case 4: k_lsq_elem<4, MODE><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->X, ctx->E, ctx->nu, u, field, comp0, x, ctx->ye64, done); break;
    case 8: k_lsq_elem<8, MODE><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->X, ctx->E, ctx->nu, u, field, comp0, x, ctx->ye64, done); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mass_elem(jfm_ctx *ctx, const double *x, const int *done) {
    return lsq_dispatch<0>(ctx, nullptr, 0, 0, x, done);
}

hipError_t launch_post_rhs(jfm_ctx *ctx, const double *u, int field, int comp0, double *b) {
    hipError_t e = lsq_dispatch<1>(ctx, u, field, comp0, nullptr, nullptr);
    if (e != hipSuccess) return e;
    return launch_gather<double>(ctx, ctx->ye64, nullptr, b, nullptr, nullptr, nullptr, nullptr);
}

hipError_t launch_mass_diag(jfm_ctx *ctx, double *Minv) {
    hipError_t e = lsq_dispatch<2>(ctx, nullptr, 0, 0, nullptr, nullptr);
    if (e != hipSuccess) return e;
    e = launch_gather<double>(ctx, ctx->ye64, nullptr, ctx->vy64, nullptr, nullptr, nullptr, nullptr);
    if (e != hipSuccess) return e;
    k_diag_blocks<<<nblk(ctx->N), kBlock, 0, ctx->stream>>>((int)ctx->N, ctx->vy64, Minv);
    return hipGetLastError();
}

hipError_t launch_post_gp(jfm_ctx *ctx, const double *u, double *strain, double *stress) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    const unsigned g = nblk(n);
#define POST(NPE)                                                                                                    \
    if (ctx->affine)                                                                                                 \
        k_post_gp<NPE, true><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->gf64, ctx->E, ctx->nu, u, strain, stress); \
    else                                                                                                             \
        k_post_gp<NPE, false><<<g, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->gf64, ctx->E, ctx->nu, u, strain, stress);
    switch (ctx->npe) {
    case 10: POST(10) break;
    case 4: POST(4) break;
    case 8: POST(8) break;
    default: return hipErrorInvalidValue;
    }
#undef POST
    return hipGetLastError();
}

}  // namespace jfem
