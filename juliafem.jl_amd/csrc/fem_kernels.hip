// fem_kernels.hip -- element-level kernels of the elasticity path (gfx950).
//
// Reference algorithm (paths relative to the reference repository root):
//   element stiffness      src/problems_elasticity.jl:203-451
//                          (Km += w BL' D BL, :369-375; w = ip.weight*detJ, :248;
//                          finite strain BL(F), Kg, f_int, :255-289, 378-409)
//   Jacobian and gradients src/basis/math.jl:185-212 (J = sum dN_i (x) X_i,
//                          grad_i = invJ . dN_i, signed detJ)
//   surface traction       src/problems_elasticity.jl:454-502
//   nodal-assembly order   ext/JuliaFEMCUDAExt.jl:179-193, 293-361
//
// For the isotropic D of src/problems_elasticity.jl:310-317 the 3x3 block
// (i,j) of BL' D BL is  lambda g_i (x) g_j + mu (g_j (x) g_i + (g_i . g_j) I)
// (g = grad N), which is what the kernels evaluate instead of the 6x30
// products; the oracle keeps the reference's BL form and the tests compare.
//
// Every nodal sum runs over the node's incidences in ascending element id;
// there are no floating-point atomics.
#include <algorithm>

#include "jfem_device.h"
#include "ebe_element.h"
#include "mf_element.h"

namespace jfem {

// ------------------------------------------------------- geometric factors --
// invJ, lambda*w*detJ, mu*w*detJ per element (affine) or per Gauss point.
template <int NPE>
__global__ __launch_bounds__(256) void k_geom(int n_el, int ngf, const int32_t *__restrict__ conn,
                                             const double *__restrict__ X, const double *__restrict__ E,
                                             const double *__restrict__ nu, double *__restrict__ gf64,
                                             float *__restrict__ gf32, int *__restrict__ bad) {
    using EL = Elem<NPE>;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double Xe[NPE][3];
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        const int nd = conn[(size_t)i * n_el + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) Xe[i][c] = X[3 * (size_t)nd + c];
    }
    const double Ev = E[e], nv = nu[e];
    const double la = Ev * nv / ((1.0 + nv) * (1.0 - 2.0 * nv));
    const double mu = Ev / (2.0 * (1.0 + nv));
#pragma unroll
    for (int q = 0; q < EL::NGP; ++q) {
        if (q >= ngf) break;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int i = 0; i < NPE; ++i)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double d = EL::dN(q, i, a);
                if (d != 0.0)
#pragma unroll
                    for (int b = 0; b < 3; ++b) J[a][b] += d * Xe[i][b];
            }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (!(det > 0.0)) atomicMax(bad, 1);
        const double id = 1.0 / det;
        double inv[9];
        inv[0] = c00 * id;
        inv[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        inv[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        inv[3] = c01 * id;
        inv[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        inv[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        inv[6] = c02 * id;
        inv[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        inv[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        const double wd = EL::w(q) * det;
        double *g = gf64 + (size_t)q * GF * n_el + e;
        float *h = gf32 + (size_t)q * GF * n_el + e;
#pragma unroll
        for (int k = 0; k < 9; ++k) { g[(size_t)k * n_el] = inv[k]; h[(size_t)k * n_el] = (float)inv[k]; }
        g[(size_t)9 * n_el] = la * wd;
        g[(size_t)10 * n_el] = mu * wd;
        h[(size_t)9 * n_el] = (float)(la * wd);
        h[(size_t)10 * n_el] = (float)(mu * wd);
    }
}

// ---------------------------------------------------- matrix-free element --
// Phase 1 of the two-phase operator: y_e = K_e x_e from the geometric factors
// (mf_element.h), written to the element's incidence slots.
template <int NPE, typename T, bool AFF>
__global__ __launch_bounds__(256) void k_mf_elem(int n_el, const int32_t *__restrict__ conn,
                                                const int32_t *__restrict__ slot, const T *__restrict__ gf,
                                                const T *__restrict__ x, T *__restrict__ ye,
                                                const int *__restrict__ done, const int32_t *__restrict__ elist,
                                                int n_items, int64_t xs = 0, int64_t ys = 0) {
    if (done && *done) return;
    const int idx = xcd_block() * blockDim.x + threadIdx.x;
    if (idx >= n_items) return;
    const int e = elist ? elist[idx] : idx;   // optional element subset
    x += blockIdx.y * xs;                     // vector blockIdx.y of a batch
    ye += blockIdx.y * ys;
    mf_element<NPE, T, AFF>(n_el, e, conn, gf, x, [&](int i, T f0, T f1, T f2) {
        T *o = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        o[0] = f0;
        o[1] = f1;
        o[2] = f2;
    });
}

// ------------------------------------------------------------- EBE element --
// Phase 1 with stored element matrices: y_e = K_e x_e (ebe_element.h), one
// thread per element in the order the matrices are stored.  With eperm the
// matrices are in the patch operator's order: position k holds element
// eperm[k], conn is the patch-ordered connectivity (indexed by k) and the
// results go to element eperm[k]'s incidence slots.
template <int NPE, typename T>
__global__ __launch_bounds__(256) void k_ebe_elem(int n_el, const int32_t *__restrict__ conn,
                                                 const int32_t *__restrict__ eperm, const int32_t *__restrict__ slot,
                                                 const T *__restrict__ Ke, const T *__restrict__ x, T *__restrict__ ye,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    if (k >= n_el) return;
    const int e = eperm ? eperm[k] : k;
    T f[NPE][3];
    ebe_element<NPE, T>(n_el, k, conn, Ke, x, f);
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        T *out = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
        out[0] = f[i][0];
        out[1] = f[i][1];
        out[2] = f[i][2];
    }
}

// ------------------------------------------------------------ nodal gather --
// Phase 2: y[a] = sum over the incidences (e, i) of node a, in ascending
// element id, of the element results.  Phase 1 wrote the result of incidence
// (e, i) into its slot k of the node-sorted incidence list (ptr/inc order),
// so a node's terms are contiguous and the gather is a streaming segmented
// sum.  Fixed DOFs are zeroed; optional partials of dot(y, dotv).
template <typename T>
__global__ __launch_bounds__(256) void k_gather(int N, const int32_t *__restrict__ ptr, const T *__restrict__ yinc,
                                               const uint8_t *__restrict__ fixed, T *__restrict__ y,
                                               const T *__restrict__ dotv, double *__restrict__ part,
                                               const int *__restrict__ done, const int32_t *__restrict__ nlist,
                                               int64_t ys = 0, int64_t ns = 0) {
    if (done && *done) return;
    yinc += blockIdx.y * ys;   // vector blockIdx.y of a batch
    y += blockIdx.y * ns;
    double dv = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < N; k += gridDim.x * blockDim.x) {
        const int a = nlist ? nlist[k] : k;   // optional node subset (N items)
        const int k0 = ptr[a], k1 = ptr[a + 1];
        T s0 = T(0), s1 = T(0), s2 = T(0);
        const T *v = yinc + 3 * (size_t)k0;
#pragma unroll 4
        for (int k = 0; k < k1 - k0; ++k) {
            s0 += v[3 * k];
            s1 += v[3 * k + 1];
            s2 += v[3 * k + 2];
        }
        const size_t o = 3 * (size_t)a;
        if (fixed) {
            if (fixed[o]) s0 = T(0);
            if (fixed[o + 1]) s1 = T(0);
            if (fixed[o + 2]) s2 = T(0);
        }
        y[o] = s0;
        y[o + 1] = s1;
        y[o + 2] = s2;
        if (dotv) dv += (double)s0 * (double)dotv[o] + (double)s1 * (double)dotv[o + 1] + (double)s2 * (double)dotv[o + 2];
    }
    if (part) {
        double v1[1] = {dv};
        write_partials<1>(v1, part);
    }
}

// Thread-per-node gather (face loads; not on the iteration path).
__global__ __launch_bounds__(256) void k_gather_simple(int nrows, const int32_t *__restrict__ ptr,
                                                      const int32_t *__restrict__ inc, const double *__restrict__ ye,
                                                      double *__restrict__ y) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nrows) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
        const size_t idx = 3 * (size_t)inc[k];
        s0 += ye[idx];
        s1 += ye[idx + 1];
        s2 += ye[idx + 2];
    }
    y[3 * (size_t)a] = s0;
    y[3 * (size_t)a + 1] = s1;
    y[3 * (size_t)a + 2] = s2;
}

// --------------------------------------------------------- element matrices --
// Gradients of shape function i at Gauss point q of element e, from its
// geometric factors: grad N_i = invJ dN_i(xi_q).
template <int NPE>
__device__ __forceinline__ void elem_grad(int n_el, int e, const double (&Ji)[9], int q, int i, double (&g)[3]) {
    const double d0 = dn_rt<NPE>(q, i, 0), d1 = dn_rt<NPE>(q, i, 1), d2 = dn_rt<NPE>(q, i, 2);
#pragma unroll
    for (int b = 0; b < 3; ++b) g[b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
}

// Block (i, j) of K_e: sum_q cl g_i (x) g_j + cm (g_j (x) g_i + (g_i . g_j) I),
// cl = lambda w detJ, cm = mu w detJ (the isotropic BL' D BL, Voigt D of
// src/problems_elasticity.jl:310-317).
template <int NPE>
__device__ __forceinline__ void elem_block(int n_el, int ngf, const double *__restrict__ gf, int e, int i, int j,
                                           double (&K)[9]) {
    constexpr int NG = Elem<NPE>::NGP;
#pragma unroll
    for (int m = 0; m < 9; ++m) K[m] = 0.0;
    double Ji[9], cl = 0.0, cm = 0.0;
#pragma unroll
    for (int q = 0; q < NG; ++q) {
        if (q == 0 || ngf > 1) {
            const double *G = gf + (size_t)(ngf > 1 ? q : 0) * GF * n_el + e;
#pragma unroll
            for (int k = 0; k < 9; ++k) Ji[k] = G[(size_t)k * n_el];
            cl = G[(size_t)9 * n_el];
            cm = G[(size_t)10 * n_el];
        }
        double gi[3], gj[3];
        elem_grad<NPE>(n_el, e, Ji, q, i, gi);
        elem_grad<NPE>(n_el, e, Ji, q, j, gj);
        const double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                K[a * 3 + b] += cl * gi[a] * gj[b] + cm * (gi[b] * gj[a] + (a == b ? dot : 0.0));
    }
}

// Element matrices for the EBE operator: thread (pair p = blockIdx.y, element
// e), written in FP64 and FP32 in the chunked, element-interleaved layout
// k_ebe_elem reads.
template <int NPE>
__global__ __launch_bounds__(256) void k_elem_mat(int n_el, int ngf, const double *__restrict__ gf,
                                                 double *__restrict__ Ke, float *__restrict__ Ke32) {
    constexpr PairTable<NPE> PT = make_pairs<NPE>();
    const int e = xcd_block() * blockDim.x + threadIdx.x, p = blockIdx.y;
    if (e >= n_el) return;
    double K[9];
    elem_block<NPE>(n_el, ngf, gf, e, PT.i[p], PT.j[p], K);
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        const int q = 9 * p + m;
        Ke[((size_t)(q / 2) * n_el + e) * 2 + q % 2] = K[m];
        Ke32[((size_t)(q / 4) * n_el + e) * 4 + q % 4] = (float)K[m];
    }
}

// --------------------------------------- finite strain (total Lagrangian) --
// Problem{Elasticity} with finite_strain = true (src/problems_elasticity.jl:
// 203-451): at each Gauss point H = grad u, F = I + H, E = (H + H' + H'H)/2,
// S = D E (St. Venant-Kirchhoff: lambda tr(E) I + 2 mu E).  For the isotropic
// D the 3x3 block (i,j) of BL(F)' D BL(F) is
//     lambda a_i (x) a_j + mu ((g_i . g_j) F F' + a_j (x) a_i),   a = F g,
// the block of BNL' S3 BNL is (g_i' S g_j) I, and BL' S at node i is F S g_i
// (all times w detJ in the reference configuration).  which = 0: Km, 1: Kg,
// 2: Km + Kg, into the packed pair blocks `Ke`; f_int (when ye != null) goes
// to the incidence slots for the nodal gather.  One 64-lane block per element.
template <int NPE>
__global__ __launch_bounds__(64) void k_tl_elem(int n_el, int ngf, const double *__restrict__ gf,
                                               const int32_t *__restrict__ conn, const int32_t *__restrict__ slot,
                                               const double *__restrict__ u, int which, double *__restrict__ Ke,
                                               double *__restrict__ ye) {
    using EL = Elem<NPE>;
    constexpr int NG = EL::NGP, NP = EL::NPAIR;
    constexpr PairTable<NPE> PT = make_pairs<NPE>();
    __shared__ double g[NG][NPE][3];
    __shared__ double ue[NPE][3];
    __shared__ double scl[NG], scm[NG];
    __shared__ double Fq[NG][9], FFq[NG][9], Sq[NG][9];
    const int e = blockIdx.x, t = threadIdx.x;
    if (t < NG * NPE) {
        const int q = t / NPE, i = t % NPE, qf = (ngf == 1) ? 0 : q;
        const double *G = gf + (size_t)qf * GF * n_el + e;
        double Ji[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Ji[k] = G[(size_t)k * n_el];
        const double d0 = dn_rt<NPE>(q, i, 0), d1 = dn_rt<NPE>(q, i, 1), d2 = dn_rt<NPE>(q, i, 2);
#pragma unroll
        for (int b = 0; b < 3; ++b) g[q][i][b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
        if (i == 0) {
            scl[q] = G[(size_t)9 * n_el];
            scm[q] = G[(size_t)10 * n_el];
        }
    }
    if (t < NPE) {
        const size_t a = (size_t)conn[(size_t)t * n_el + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[t][c] = u[3 * a + c];
    }
    __syncthreads();
    if (t < NG) {
        const int q = t;
        double H[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) H[m] = 0.0;
        for (int i = 0; i < NPE; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int A = 0; A < 3; ++A) H[c * 3 + A] += ue[i][c] * g[q][i][A];
        double F[9], E[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) F[m] = H[m] + ((m % 4) == 0 ? 1.0 : 0.0);
#pragma unroll
        for (int A = 0; A < 3; ++A)
#pragma unroll
            for (int B = 0; B < 3; ++B) {
                double hh = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) hh += H[c * 3 + A] * H[c * 3 + B];
                E[A * 3 + B] = 0.5 * (H[A * 3 + B] + H[B * 3 + A] + hh);
            }
        const double tr = E[0] + E[4] + E[8];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            Fq[q][m] = F[m];
            Sq[q][m] = 2.0 * scm[q] * E[m] + ((m % 4) == 0 ? scl[q] * tr : 0.0);   // w detJ S
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                FFq[q][c * 3 + d] = F[c * 3] * F[d * 3] + F[c * 3 + 1] * F[d * 3 + 1] + F[c * 3 + 2] * F[d * 3 + 2];
    }
    __syncthreads();
    if (Ke && t < NP) {
        const int i = PT.i[t], j = PT.j[t];
        double K[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) K[m] = 0.0;
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const double gi[3] = {g[q][i][0], g[q][i][1], g[q][i][2]};
            const double gj[3] = {g[q][j][0], g[q][j][1], g[q][j][2]};
            if (which != 1) {
                const double cl = scl[q], cm = scm[q];
                const double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
                double ai[3], aj[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ai[c] = Fq[q][c * 3] * gi[0] + Fq[q][c * 3 + 1] * gi[1] + Fq[q][c * 3 + 2] * gi[2];
                    aj[c] = Fq[q][c * 3] * gj[0] + Fq[q][c * 3 + 1] * gj[1] + Fq[q][c * 3 + 2] * gj[2];
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        K[c * 3 + d] += cl * ai[c] * aj[d] + cm * (dot * FFq[q][c * 3 + d] + aj[c] * ai[d]);
            }
            if (which != 0) {
                double kg = 0.0;
#pragma unroll
                for (int A = 0; A < 3; ++A)
                    kg += gi[A] * (Sq[q][A * 3] * gj[0] + Sq[q][A * 3 + 1] * gj[1] + Sq[q][A * 3 + 2] * gj[2]);
                K[0] += kg;
                K[4] += kg;
                K[8] += kg;
            }
        }
        double *out = Ke + ((size_t)e * NP + t) * 9;
#pragma unroll
        for (int m = 0; m < 9; ++m) out[m] = K[m];
    }
    if (ye && t < NPE) {
        double fi[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            double sg[3];
#pragma unroll
            for (int A = 0; A < 3; ++A)
                sg[A] = Sq[q][A * 3] * g[q][t][0] + Sq[q][A * 3 + 1] * g[q][t][1] + Sq[q][A * 3 + 2] * g[q][t][2];
#pragma unroll
            for (int c = 0; c < 3; ++c) fi[c] += Fq[q][c * 3] * sg[0] + Fq[q][c * 3 + 1] * sg[1] + Fq[q][c * 3 + 2] * sg[2];
        }
        double *o = ye + 3 * (size_t)slot[(size_t)t * n_el + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = fi[c];
    }
}

// ------------------------------------------------- CSR values (phase 2) ----
// One thread per 3x3 block of K: sums its contributing element blocks in
// ascending element id, each recomputed from the element's geometric factors
// (no element matrices are stored or read), and writes the scalar-CSR-ordered
// values in FP64 and FP32.
template <int NPE>
__global__ __launch_bounds__(256) void k_asm_vals(int nnzb, int n_el, int ngf, const double *__restrict__ gf,
                                                 const int32_t *__restrict__ cptr, const int32_t *__restrict__ contrib,
                                                 const int32_t *__restrict__ blk_row,
                                                 const int32_t *__restrict__ brow, double *__restrict__ vals,
                                                 float *__restrict__ vals32) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= nnzb) return;
    double acc[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) acc[m] = 0.0;
    const int k1 = cptr[s + 1];
    for (int k = cptr[s]; k < k1; ++k) {
        const int c = contrib[k];
        const int e = c / (NPE * NPE), rem = c - e * NPE * NPE;
        const int i = rem / NPE, j = rem - i * NPE;
        double B[9];
        elem_block<NPE>(n_el, ngf, gf, e, i, j, B);
#pragma unroll
        for (int m = 0; m < 9; ++m) acc[m] += B[m];
    }
    const int a = blk_row[s];
    const int r0 = brow[a], nb = brow[a + 1] - r0;
    const size_t base = 9 * (size_t)r0 + 3 * (size_t)(s - r0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t o = base + (size_t)c * 3 * nb + d;
            vals[o] = acc[c * 3 + d];
            vals32[o] = (float)acc[c * 3 + d];
        }
}

// The same from stored element matrices, element-major packed pairs
// Ke[(e * NPAIR + p) * 9 + m] (the finite-strain blocks of k_tl_elem).
template <int NPE>
__global__ __launch_bounds__(256) void k_asm_vals_ke(int nnzb, const int32_t *__restrict__ cptr,
                                                 const int32_t *__restrict__ contrib,
                                                 const int32_t *__restrict__ blk_row,
                                                 const int32_t *__restrict__ brow, const double *__restrict__ Ke,
                                                 double *__restrict__ vals, float *__restrict__ vals32) {
    using EL = Elem<NPE>;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnzb) return;
    double acc[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) acc[m] = 0.0;
    const int k1 = cptr[s + 1];
    for (int k = cptr[s]; k < k1; ++k) {
        const int c = contrib[k];
        const int e = c / (NPE * NPE), rem = c - e * NPE * NPE;
        const int i = rem / NPE, j = rem - i * NPE;
        if (i <= j) {
            const double *B = Ke + ((size_t)e * EL::NPAIR + pair_index<NPE>(i, j)) * 9;
#pragma unroll
            for (int m = 0; m < 9; ++m) acc[m] += B[m];
        } else {
            const double *B = Ke + ((size_t)e * EL::NPAIR + pair_index<NPE>(j, i)) * 9;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) acc[a * 3 + b] += B[b * 3 + a];
        }
    }
    const int a = blk_row[s];
    const int r0 = brow[a], nb = brow[a + 1] - r0;
    const size_t base = 9 * (size_t)r0 + 3 * (size_t)(s - r0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t o = base + (size_t)c * 3 * nb + d;
            vals[o] = acc[c * 3 + d];
            vals32[o] = (float)acc[c * 3 + d];
        }
}

// ------------------------------------------------------------ face loads ----
// Consistent traction load per face (src/problems_elasticity.jl:454-502):
// f_i += w |dX/du x dX/dv| t N_i.  Tri6: 3-pt rule (gauss_points.jl:100-106),
// Tri3: 1-pt (:93-97), Quad4: 2x2 (:133-142).
__global__ __launch_bounds__(256) void k_face(int nf, int nfn, const int32_t *__restrict__ fconn,
                                             const double *__restrict__ X, const double *__restrict__ ft,
                                             double *__restrict__ fe) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double Xf[6][3];
    for (int i = 0; i < nfn; ++i) {
        const int nd = fconn[(size_t)i * nf + f];
        for (int c = 0; c < 3; ++c) Xf[i][c] = X[3 * (size_t)nd + c];
    }
    double w[4], pu[4], pv[4];
    const int ng = surf_rule(nfn, w, pu, pv);
    double out[6][3];
    for (int i = 0; i < nfn; ++i) out[i][0] = out[i][1] = out[i][2] = 0.0;
    const double t0 = ft[3 * (size_t)f], t1 = ft[3 * (size_t)f + 1], t2 = ft[3 * (size_t)f + 2];
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
        for (int i = 0; i < nfn; ++i) {
            out[i][0] += wq * t0 * N[i];
            out[i][1] += wq * t1 * N[i];
            out[i][2] += wq * t2 * N[i];
        }
    }
    for (int i = 0; i < nfn; ++i)
        for (int c = 0; c < 3; ++c) fe[((size_t)f * nfn + i) * 3 + c] = out[i][c];
}

// ----------------------------------------------------------- block Jacobi --
// Diagonal 3x3 blocks sum_e K_e(i,i) over the node's incidences (ascending
// element id), fixed rows/columns replaced by identity before inversion and
// zeroed after, so the preconditioner never touches fixed DOFs.
template <int NPE>
__global__ __launch_bounds__(256) void k_block_jacobi(int N, const int32_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ inc, int n_el, int ngf,
                                                     const double *__restrict__ gf,
                                                     const uint8_t *__restrict__ fixed,
                                                     double *__restrict__ M64, float *__restrict__ M32) {
    using EL = Elem<NPE>;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int k1 = ptr[a + 1];
    for (int k = ptr[a]; k < k1; ++k) {
        const int idx = inc[k];
        const int e = idx / NPE, i = idx - e * NPE;
        for (int q = 0; q < EL::NGP; ++q) {
            const double *G = gf + (size_t)(ngf == 1 ? 0 : q) * GF * n_el + e;
            double Ji[9];
            for (int m = 0; m < 9; ++m) Ji[m] = G[(size_t)m * n_el];
            const double cl = G[(size_t)9 * n_el], cm = G[(size_t)10 * n_el];
            const double d0 = dn_rt<NPE>(q, i, 0), d1 = dn_rt<NPE>(q, i, 1), d2 = dn_rt<NPE>(q, i, 2);
            double g[3];
            for (int b = 0; b < 3; ++b) g[b] = Ji[b * 3] * d0 + Ji[b * 3 + 1] * d1 + Ji[b * 3 + 2] * d2;
            const double dot = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) B[r * 3 + c] += cl * g[r] * g[c] + cm * (g[c] * g[r] + (r == c ? dot : 0.0));
        }
    }
    const uint8_t f0 = fixed ? fixed[3 * (size_t)a] : 0, f1 = fixed ? fixed[3 * (size_t)a + 1] : 0,
                  f2 = fixed ? fixed[3 * (size_t)a + 2] : 0;
    const uint8_t fx[3] = {f0, f1, f2};
    for (int c = 0; c < 3; ++c)
        if (fx[c])
            for (int d = 0; d < 3; ++d) { B[c * 3 + d] = (c == d); B[d * 3 + c] = (c == d); }
    const double c00 = B[4] * B[8] - B[5] * B[7];
    const double c01 = B[5] * B[6] - B[3] * B[8];
    const double c02 = B[3] * B[7] - B[4] * B[6];
    const double id = 1.0 / (B[0] * c00 + B[1] * c01 + B[2] * c02);
    double M[9];
    M[0] = c00 * id;
    M[1] = (B[2] * B[7] - B[1] * B[8]) * id;
    M[2] = (B[1] * B[5] - B[2] * B[4]) * id;
    M[3] = c01 * id;
    M[4] = (B[0] * B[8] - B[2] * B[6]) * id;
    M[5] = (B[2] * B[3] - B[0] * B[5]) * id;
    M[6] = c02 * id;
    M[7] = (B[1] * B[6] - B[0] * B[7]) * id;
    M[8] = (B[0] * B[4] - B[1] * B[3]) * id;
    for (int c = 0; c < 3; ++c)
        if (fx[c])
            for (int d = 0; d < 3; ++d) { M[c * 3 + d] = 0.0; M[d * 3 + c] = 0.0; }
    for (int m = 0; m < 9; ++m) {
        M64[9 * (size_t)a + m] = M[m];
        M32[9 * (size_t)a + m] = (float)M[m];
    }
}

// ------------------------------------------------------------- launchers ----
static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_geom_factors(jfm_ctx *ctx, int *d_bad) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    switch (ctx->npe) {
    case 10: k_geom<10><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->ngf, ctx->conn, ctx->X, ctx->E, ctx->nu, ctx->gf64, ctx->gf32, d_bad); break;
    case 4: k_geom<4><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->ngf, ctx->conn, ctx->X, ctx->E, ctx->nu, ctx->gf64, ctx->gf32, d_bad); break;
    case 8: k_geom<8><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->ngf, ctx->conn, ctx->X, ctx->E, ctx->nu, ctx->gf64, ctx->gf32, d_bad); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_element_matrices(jfm_ctx *ctx) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    const dim3 grid(xcd_grid(nblk(n)), (unsigned)ctx->npair);
    // in the patch order when the patch operator is set up (ke_perm), from the
    // factors in that order, so that the matrices sit where it reads them
    const double *gf = ctx->ke_perm ? ctx->gfP64 : ctx->gf64;
    switch (ctx->npe) {
    case 10: k_elem_mat<10><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->ngf, gf, ctx->Ke, ctx->Ke32); break;
    case 4: k_elem_mat<4><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->ngf, gf, ctx->Ke, ctx->Ke32); break;
    case 8: k_elem_mat<8><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->ngf, gf, ctx->Ke, ctx->Ke32); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tl_element_matrices(jfm_ctx *ctx, const double *u, int which, double *Ke, double *ye) {
    const int n = (int)ctx->n_el;
    if (n == 0) return hipSuccess;
    switch (ctx->npe) {
    case 10: k_tl_elem<10><<<n, 64, 0, ctx->stream>>>(n, ctx->ngf, ctx->gf64, ctx->conn, ctx->inc_slot, u, which, Ke, ye); break;
    case 4: k_tl_elem<4><<<n, 64, 0, ctx->stream>>>(n, ctx->ngf, ctx->gf64, ctx->conn, ctx->inc_slot, u, which, Ke, ye); break;
    case 8: k_tl_elem<8><<<n, 64, 0, ctx->stream>>>(n, ctx->ngf, ctx->gf64, ctx->conn, ctx->inc_slot, u, which, Ke, ye); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_assemble_values(jfm_ctx *ctx) {
    const int n = (int)ctx->nnzb;
    if (n == 0) return hipSuccess;
    const int ne = (int)ctx->n_el;
    const unsigned g = xcd_grid(nblk(n));
    switch (ctx->npe) {
    case 10: k_asm_vals<10><<<g, kBlock, 0, ctx->stream>>>(n, ne, ctx->ngf, ctx->gf64, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, ctx->vals64, ctx->vals32); break;
    case 4: k_asm_vals<4><<<g, kBlock, 0, ctx->stream>>>(n, ne, ctx->ngf, ctx->gf64, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, ctx->vals64, ctx->vals32); break;
    case 8: k_asm_vals<8><<<g, kBlock, 0, ctx->stream>>>(n, ne, ctx->ngf, ctx->gf64, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, ctx->vals64, ctx->vals32); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_assemble_values_ke(jfm_ctx *ctx, const double *Ke) {
    const int n = (int)ctx->nnzb;
    if (n == 0) return hipSuccess;
    switch (ctx->npe) {
    case 10: k_asm_vals_ke<10><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, Ke, ctx->vals64, ctx->vals32); break;
    case 4: k_asm_vals_ke<4><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, Ke, ctx->vals64, ctx->vals32); break;
    case 8: k_asm_vals_ke<8><<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->cptr, ctx->contrib, ctx->blk_row, ctx->brow, Ke, ctx->vals64, ctx->vals32); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_face_loads(jfm_ctx *ctx) {
    const int nf = (int)ctx->n_faces;
    if (nf > 0) {
        k_face<<<nblk(nf), kBlock, 0, ctx->stream>>>(nf, ctx->nfn, ctx->face_conn, ctx->X, ctx->face_t, ctx->fe);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return launch_face_gather(ctx);
}

hipError_t launch_face_gather(jfm_ctx *ctx) {
    const int N = (int)ctx->N;
    k_gather_simple<<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->finc_ptr, ctx->finc, ctx->fe, ctx->f_ext);
    return hipGetLastError();
}

hipError_t launch_block_jacobi(jfm_ctx *ctx, double *Minv64, float *Minv32) {
    const int N = (int)ctx->N;
    switch (ctx->npe) {
    case 10: k_block_jacobi<10><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, (int)ctx->n_el, ctx->ngf, ctx->gf64, ctx->fixed, Minv64, Minv32); break;
    case 4: k_block_jacobi<4><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, (int)ctx->n_el, ctx->ngf, ctx->gf64, ctx->fixed, Minv64, Minv32); break;
    case 8: k_block_jacobi<8><<<nblk(N), kBlock, 0, ctx->stream>>>(N, ctx->inc_ptr, ctx->inc, (int)ctx->n_el, ctx->ngf, ctx->gf64, ctx->fixed, Minv64, Minv32); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_gather(jfm_ctx *ctx, const T *ye, const uint8_t *fixed, T *y, const T *dotv, double *part,
                         const int *done, int *npart) {
    const unsigned g = capped_grid(ctx->N, kBlock);
    if (npart) *npart = (int)g;
    k_gather<T><<<g, kBlock, 0, ctx->stream>>>((int)ctx->N, ctx->inc_ptr, ye, fixed, y, dotv, part, done, nullptr);
    return hipGetLastError();
}

// y_j = K x_j (masked, FP64) for the `nv` vectors x_j = x + j xs (y_j, and
// the incidence slots ye_j = ye + j ys, likewise), restricted to the elements
// `elist` and the nodes `nlist`: only those element slots are written and only
// those rows of y_j.  Every other incidence slot must hold zero (the caller
// keeps it so).
template <int NPE>
static void subset_elem(jfm_ctx *ctx, const double *x, double *ye, int nv, int64_t xs, int64_t ys,
                        const int32_t *elist, int ne) {
    const int n = (int)ctx->n_el;
    const dim3 grid(xcd_grid(nblk(ne)), nv);
    if (ctx->affine)
        k_mf_elem<NPE, double, true><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->gf64, x, ye,
                                                                      nullptr, elist, ne, xs, ys);
    else
        k_mf_elem<NPE, double, false><<<grid, kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, ctx->gf64, x, ye,
                                                                       nullptr, elist, ne, xs, ys);
}

hipError_t launch_subset_apply(jfm_ctx *ctx, bool tangent, const double *x, double *y, double *ye, int nv, int64_t xs,
                               int64_t ys, const int32_t *elist, int ne, const int32_t *nlist, int nn) {
    if (ne > 0) {
        if (tangent) {
            hipError_t e = launch_nh_elem<double>(ctx, true, ctx->gf64, ctx->ustate64, x, ye, nullptr, elist, ne, nv, xs, ys);
            if (e != hipSuccess) return e;
        } else {
            switch (ctx->npe) {
            case 10: subset_elem<10>(ctx, x, ye, nv, xs, ys, elist, ne); break;
            case 4: subset_elem<4>(ctx, x, ye, nv, xs, ys, elist, ne); break;
            case 8: subset_elem<8>(ctx, x, ye, nv, xs, ys, elist, ne); break;
            default: return hipErrorInvalidValue;
            }
        }
    }
    if (nn > 0)
        k_gather<double><<<dim3(capped_grid(nn, kBlock), nv), kBlock, 0, ctx->stream>>>(
            nn, ctx->inc_ptr, ye, ctx->fixed, y, nullptr, nullptr, nullptr, nlist, ys, xs);
    return hipGetLastError();
}

template <typename T> static T *ye_of(jfm_ctx *ctx);
template <> double *ye_of<double>(jfm_ctx *ctx) { return ctx->ye64; }
template <> float *ye_of<float>(jfm_ctx *ctx) { return ctx->ye32; }
template <typename T> static const T *ke_of(jfm_ctx *ctx);
template <> const double *ke_of<double>(jfm_ctx *ctx) { return ctx->Ke; }
template <> const float *ke_of<float>(jfm_ctx *ctx) { return ctx->Ke32; }
template <typename T> static const T *gf_of(jfm_ctx *ctx);
template <> const double *gf_of<double>(jfm_ctx *ctx) { return ctx->gf64; }
template <> const float *gf_of<float>(jfm_ctx *ctx) { return ctx->gf32; }

template <int NPE, typename T>
static void launch_elem_phase(jfm_ctx *ctx, int op, const T *x, T *ye, const int *done) {
    const int n = (int)ctx->n_el;
    if (op == JFM_OP_EBE) {
        k_ebe_elem<NPE, T><<<xcd_grid(nblk(n)), kBlock, 0, ctx->stream>>>(
            n, ctx->ke_perm ? ctx->connP : ctx->conn, ctx->ke_perm, ctx->inc_slot, ke_of<T>(ctx), x, ye, done);
    } else if (ctx->affine) {
        k_mf_elem<NPE, T, true><<<xcd_grid(nblk(n)), kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, gf_of<T>(ctx), x, ye, done,
                                                                              nullptr, n);
    } else {
        k_mf_elem<NPE, T, false><<<xcd_grid(nblk(n)), kBlock, 0, ctx->stream>>>(n, ctx->conn, ctx->inc_slot, gf_of<T>(ctx), x, ye, done,
                                                                               nullptr, n);
    }
}

template <typename T>
hipError_t launch_elem_only(jfm_ctx *ctx, int op, const T *x, const int *done) {
    if (ctx->n_el == 0) return hipSuccess;
    if (op == kOpMass) {  // consistent mass (least-squares recovery), FP64 only
        if (sizeof(T) != 8) return hipErrorInvalidValue;
        return launch_mass_elem(ctx, (const double *)x, done);
    }
    if (op == JFM_OP_TANGENT) {  // neo-Hookean K_T(u) x at the stored state
        const T *u = sizeof(T) == 8 ? (const T *)ctx->ustate64 : (const T *)ctx->ustate32;
        return launch_nh_elem<T>(ctx, true, gf_of<T>(ctx), u, x, ye_of<T>(ctx), done);
    }
    switch (ctx->npe) {
    case 10: launch_elem_phase<10, T>(ctx, op, x, ye_of<T>(ctx), done); break;
    case 4: launch_elem_phase<4, T>(ctx, op, x, ye_of<T>(ctx), done); break;
    case 8: launch_elem_phase<8, T>(ctx, op, x, ye_of<T>(ctx), done); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Element phase of the linear operator (JFM_OP_MATFREE or JFM_OP_EBE) into the
// incidence slots `ye` (3 npe n_el values), a buffer the caller provides.
template <typename T>
hipError_t launch_elem_into(jfm_ctx *ctx, int op, const T *x, T *ye) {
    if (op != JFM_OP_MATFREE && op != JFM_OP_EBE) return hipErrorInvalidValue;
    if (op == JFM_OP_EBE && !ctx->Ke) return hipErrorInvalidValue;
    if (ctx->n_el == 0) return hipSuccess;
    switch (ctx->npe) {
    case 10: launch_elem_phase<10, T>(ctx, op, x, ye, nullptr); break;
    case 4: launch_elem_phase<4, T>(ctx, op, x, ye, nullptr); break;
    case 8: launch_elem_phase<8, T>(ctx, op, x, ye, nullptr); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_two_phase(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, double *part, const int *done,
                            int *npart) {
    if (jfem_patch_form(ctx, op)) return launch_patch_apply<T>(ctx, op, x, y, dotv, part, done, npart);
    hipError_t e = launch_elem_only<T>(ctx, op, x, done);
    if (e != hipSuccess) return e;
    return launch_gather<T>(ctx, ye_of<T>(ctx), op == kOpMass ? nullptr : ctx->fixed, y, dotv, part, done, npart);
}

template hipError_t launch_elem_only<double>(jfm_ctx *, int, const double *, const int *);
template hipError_t launch_elem_into<double>(jfm_ctx *, int, const double *, double *);
template hipError_t launch_elem_into<float>(jfm_ctx *, int, const float *, float *);
template hipError_t launch_elem_only<float>(jfm_ctx *, int, const float *, const int *);
template hipError_t launch_two_phase<double>(jfm_ctx *, int, const double *, double *, const double *, double *, const int *, int *);
template hipError_t launch_two_phase<float>(jfm_ctx *, int, const float *, float *, const float *, double *, const int *, int *);
template hipError_t launch_gather<double>(jfm_ctx *, const double *, const uint8_t *, double *, const double *, double *, const int *, int *);
template hipError_t launch_gather<float>(jfm_ctx *, const float *, const uint8_t *, float *, const float *, double *, const int *, int *);

}  // namespace jfem
