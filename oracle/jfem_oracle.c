/*
 * jfem_oracle.c -- CPU FP64 restatement of JuliaFEM's linear-elasticity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  It is never linked into, called by, or used as a fallback for the
 * product library (juliafem.jl_amd/libjfem_hip.so).
 *
 * The reference is Julia and cannot run here, so every routine restates the
 * reference algorithm from its source text (paths relative to the reference
 * repository root):
 *
 *   basis functions      src/basis/lagrange_generated.jl:89-143 (Tri3, Tri6),
 *                        :145-171 (Quad4), :229-310 (Tet4, Tet10, Hex8)
 *   quadrature           src/integration/gauss_points.jl:92-106, 125-142,
 *                        165-182, 207-219; rule choice src/elements/elements.jl:828-874
 *   Jacobian / gradient  src/basis/math.jl:127-212 (J = sum dN_i (x) X_i,
 *                        grad_i = invJ . dN_i, signed detJ)
 *   element stiffness    src/problems_elasticity.jl:203-451 (BL :291-301,
 *                        D :310-317, Km += w BL' D BL :369-375, f_int :406-409)
 *   surface traction     src/problems_elasticity.jl:454-502
 *   global DOFs          src/assembly/problems.jl:466-478 (dof = 3(node-1)+c)
 *   COO add!, sparse()   src/sparse/sparse.jl:54, 121-132
 *   Dirichlet + solve    src/problems_dirichlet.jl:60-89, src/solvers.jl:192-216
 *   node->element map    ext/JuliaFEMCUDAExt.jl:179-193
 *   two-phase K.u        ext/JuliaFEMCUDAExt.jl:222-361, 442-478
 *   neo-Hookean TL       src/materials/neo_hookean.jl:129-137, 205-231;
 *                        src/problems_elasticity.jl:246-289, 369-409
 *
 * Pinning (tests/test_oracle.py): closed forms (quadrature moments,
 * partition of unity, affine-field energy u'Ku = vol * eps:D:eps, flat-face
 * consistent loads), invariants (symmetry, rigid-body null space, equilibrium),
 * the reference demo's single-Tet4 answers (demos/nodal_assembly_cpu.jl), a
 * brute-force sparsity pattern and scipy's SuperLU on the same system.
 * Neo-Hookean (tests/test_nonlinear_oracle.py): S and D against central
 * differences of psi and S, D(C = I) = the linear D, the element tangent
 * against differences of the element force and at u = 0 against the linear
 * element stiffness, zero force under rigid translations and rotations, the
 * patch test (interior nodes in equilibrium under homogeneous F).
 *
 * Deviations, each deliberate: Hex8 integrates with 2x2x2 Gauss points
 * (src/integration/integration.jl:121; the new-API default Gauss{1} at
 * src/elements/elements.jl:857 is hourglass-unstable) and Quad4 faces with
 * 2x2.  The direct solve is a banded Cholesky (CHOLMOD's ldlt is a third-party
 * routine that is not in the reference tree).
 *
 * Indices are 0-based; DOF of (node a, component c) is 3a + c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ET_TET4 = 4, ET_HEX8 = 8, ET_TET10 = 10 };
enum { FT_TRI3 = 3, FT_QUAD4 = 4, FT_TRI6 = 6 };

/* ---------------------------------------------------------------- basis -- */

/* Tet10: src/basis/lagrange_generated.jl:267-282 */
static void tet10_basis(const double *xi, double *N, double (*dN)[3]) {
    const double u = xi[0], v = xi[1], w = xi[2];
    if (N) {
        N[0] = 1 - 3.0 * u - 3.0 * v - 3.0 * w + 2.0 * u * u + 2.0 * v * v + 2.0 * w * w
               + 4.0 * u * v + 4.0 * u * w + 4.0 * v * w;
        N[1] = -u + 2.0 * u * u;
        N[2] = -v + 2.0 * v * v;
        N[3] = -w + 2.0 * w * w;
        N[4] = 4.0 * u - 4.0 * u * u - 4.0 * u * v - 4.0 * u * w;
        N[5] = 4.0 * u * v;
        N[6] = 4.0 * v - 4.0 * v * v - 4.0 * u * v - 4.0 * v * w;
        N[7] = 4.0 * w - 4.0 * w * w - 4.0 * u * w - 4.0 * v * w;
        N[8] = 4.0 * u * w;
        N[9] = 4.0 * v * w;
    }
    if (dN) {
        const double s = -3.0 + 4.0 * u + 4.0 * v + 4.0 * w;
        dN[0][0] = s; dN[0][1] = s; dN[0][2] = s;
        dN[1][0] = -1.0 + 4.0 * u; dN[1][1] = 0; dN[1][2] = 0;
        dN[2][0] = 0; dN[2][1] = -1.0 + 4.0 * v; dN[2][2] = 0;
        dN[3][0] = 0; dN[3][1] = 0; dN[3][2] = -1.0 + 4.0 * w;
        dN[4][0] = 4.0 - 8.0 * u - 4.0 * v - 4.0 * w; dN[4][1] = -4.0 * u; dN[4][2] = -4.0 * u;
        dN[5][0] = 4.0 * v; dN[5][1] = 4.0 * u; dN[5][2] = 0;
        dN[6][0] = -4.0 * v; dN[6][1] = 4.0 - 8.0 * v - 4.0 * u - 4.0 * w; dN[6][2] = -4.0 * v;
        dN[7][0] = -4.0 * w; dN[7][1] = -4.0 * w; dN[7][2] = 4.0 - 8.0 * w - 4.0 * u - 4.0 * v;
        dN[8][0] = 4.0 * w; dN[8][1] = 0; dN[8][2] = 4.0 * u;
        dN[9][0] = 0; dN[9][1] = 4.0 * w; dN[9][2] = 4.0 * v;
    }
}

/* Tet4: src/basis/lagrange_generated.jl:229-256 */
static void tet4_basis(const double *xi, double *N, double (*dN)[3]) {
    if (N) { N[0] = 1 - xi[0] - xi[1] - xi[2]; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2]; }
    if (dN) {
        static const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        memcpy(dN, d, sizeof d);
    }
}

/* Hex8: src/basis/lagrange_generated.jl:285-310, corners (+-1,+-1,+-1) */
static const double HEX8_REF[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static void hex8_basis(const double *xi, double *N, double (*dN)[3]) {
    for (int i = 0; i < 8; ++i) {
        const double a = 1 + HEX8_REF[i][0] * xi[0], b = 1 + HEX8_REF[i][1] * xi[1],
                     c = 1 + HEX8_REF[i][2] * xi[2];
        if (N) N[i] = 0.125 * a * b * c;
        if (dN) {
            dN[i][0] = 0.125 * HEX8_REF[i][0] * b * c;
            dN[i][1] = 0.125 * a * HEX8_REF[i][1] * c;
            dN[i][2] = 0.125 * a * b * HEX8_REF[i][2];
        }
    }
}

/* Tri6: src/basis/lagrange_generated.jl:117-143 */
static void tri6_basis(const double *xi, double *N, double (*dN)[2]) {
    const double u = xi[0], v = xi[1];
    N[0] = 1 - 3.0 * u - 3.0 * v + 2.0 * u * u + 4.0 * u * v + 2.0 * v * v;
    N[1] = -u + 2.0 * u * u;
    N[2] = -v + 2.0 * v * v;
    N[3] = 4.0 * u - 4.0 * u * u - 4.0 * u * v;
    N[4] = 4.0 * u * v;
    N[5] = 4.0 * v - 4.0 * u * v - 4.0 * v * v;
    dN[0][0] = -3.0 + 4.0 * u + 4.0 * v; dN[0][1] = -3.0 + 4.0 * u + 4.0 * v;
    dN[1][0] = -1.0 + 4.0 * u; dN[1][1] = 0;
    dN[2][0] = 0; dN[2][1] = -1.0 + 4.0 * v;
    dN[3][0] = 4.0 - 8.0 * u - 4.0 * v; dN[3][1] = -4.0 * u;
    dN[4][0] = 4.0 * v; dN[4][1] = 4.0 * u;
    dN[5][0] = -4.0 * v; dN[5][1] = 4.0 - 4.0 * u - 8.0 * v;
}

/* Tri3: src/basis/lagrange_generated.jl:89-115 */
static void tri3_basis(const double *xi, double *N, double (*dN)[2]) {
    N[0] = 1 - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1];
    dN[0][0] = -1; dN[0][1] = -1; dN[1][0] = 1; dN[1][1] = 0; dN[2][0] = 0; dN[2][1] = 1;
}

/* Quad4: src/basis/lagrange_generated.jl:145-171 */
static void quad4_basis(const double *xi, double *N, double (*dN)[2]) {
    static const double r[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    for (int i = 0; i < 4; ++i) {
        const double a = 1 + r[i][0] * xi[0], b = 1 + r[i][1] * xi[1];
        N[i] = 0.25 * a * b;
        dN[i][0] = 0.25 * r[i][0] * b;
        dN[i][1] = 0.25 * a * r[i][1];
    }
}

/* ----------------------------------------------------------- quadrature -- */

/* Volume rules.  Tet10 -> Gauss{2} with the reference's 15-digit literals
 * (src/integration/gauss_points.jl:173-182); Tet4 -> Gauss{1} (:166-170);
 * Hex8 -> 2x2x2 (:207-219). Returns the number of points. */
int orc_volume_rule(int etype, double *w, double (*xi)[3]) {
    if (etype == ET_TET10) {
        const double a = 0.585410196624968, b = 0.138196601125011;
        const double p[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
        for (int g = 0; g < 4; ++g) { w[g] = 1.0 / 24.0; memcpy(xi[g], p[g], sizeof p[g]); }
        return 4;
    }
    if (etype == ET_TET4) {
        w[0] = 1.0 / 6.0; xi[0][0] = xi[0][1] = xi[0][2] = 0.25;
        return 1;
    }
    if (etype == ET_HEX8) {
        const double a = 1.0 / sqrt(3.0);
        const double s[2] = {-a, a};
        int g = 0;
        for (int k = 0; k < 2; ++k)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    w[g] = 1.0; xi[g][0] = s[i]; xi[g][1] = s[j]; xi[g][2] = s[k]; ++g;
                }
        return 8;
    }
    return 0;
}

/* Surface rules: Tri6 -> Gauss{2} 3-pt (:100-106), Tri3 -> Gauss{1}
 * (:93-97), Quad4 -> 2x2 (:133-142). */
static int surface_rule(int ftype, double *w, double (*xi)[2]) {
    if (ftype == FT_TRI6) {
        const double p[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
        for (int g = 0; g < 3; ++g) { w[g] = 1.0 / 6; xi[g][0] = p[g][0]; xi[g][1] = p[g][1]; }
        return 3;
    }
    if (ftype == FT_TRI3) { w[0] = 0.5; xi[0][0] = xi[0][1] = 1.0 / 3; return 1; }
    if (ftype == FT_QUAD4) {
        const double a = 1.0 / sqrt(3.0);
        const double p[4][2] = {{-a, -a}, {a, -a}, {-a, a}, {a, a}};
        for (int g = 0; g < 4; ++g) { w[g] = 1.0; xi[g][0] = p[g][0]; xi[g][1] = p[g][1]; }
        return 4;
    }
    return 0;
}

static void volume_basis(int etype, const double *xi, double *N, double (*dN)[3]) {
    if (etype == ET_TET10) tet10_basis(xi, N, dN);
    else if (etype == ET_TET4) tet4_basis(xi, N, dN);
    else hex8_basis(xi, N, dN);
}

/* eval_basis!: J = sum_i dN_i (x) X_i  (J[a][b] = sum_i dN_i[a] X_i[b]),
 * invJ = inv(J), grad_i = invJ . dN_i, detJ signed (src/basis/math.jl:185-212). */
static double eval_basis(int npe, const double (*dN)[3], const double *Xe, double (*grad)[3]) {
    double J[3][3] = {{0}};
    for (int i = 0; i < npe; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) J[a][b] += dN[i][a] * Xe[3 * i + b];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                     - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                     + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int i = 0; i < npe; ++i)
        for (int a = 0; a < 3; ++a)
            grad[i][a] = inv[a][0] * dN[i][0] + inv[a][1] * dN[i][1] + inv[a][2] * dN[i][2];
    return det;
}

/* Isotropic D in Voigt order 11,22,33,12,23,13 (src/problems_elasticity.jl:310-317). */
static void iso_D(double E, double nu, double D[6][6]) {
    const double la = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2.0 * (1.0 + nu));
    memset(D, 0, 36 * sizeof(double));
    D[0][0] = D[1][1] = D[2][2] = 2 * mu + la;
    D[3][3] = D[4][4] = D[5][5] = mu;
    D[0][1] = D[1][0] = D[1][2] = D[2][1] = D[0][2] = D[2][0] = la;
}

/* BL (6 x 3npe), small strain (src/problems_elasticity.jl:291-301). */
static void build_BL(int npe, const double (*g)[3], double *BL /*6*3npe*/) {
    const int nd = 3 * npe;
    memset(BL, 0, 6 * nd * sizeof(double));
    for (int i = 0; i < npe; ++i) {
        const int c = 3 * i;
        BL[0 * nd + c + 0] = g[i][0];
        BL[1 * nd + c + 1] = g[i][1];
        BL[2 * nd + c + 2] = g[i][2];
        BL[3 * nd + c + 0] = g[i][1];
        BL[3 * nd + c + 1] = g[i][0];
        BL[4 * nd + c + 1] = g[i][2];
        BL[4 * nd + c + 2] = g[i][1];
        BL[5 * nd + c + 0] = g[i][2];
        BL[5 * nd + c + 2] = g[i][0];
    }
}

/* ------------------------------------------------------ element kernels -- */

/* Element stiffness Km = sum_gp w BL' D BL, w = ip.weight * detJ
 * (src/problems_elasticity.jl:246-248, 369-375).  Ke is (3npe)^2 row-major.
 * Returns the smallest detJ seen (<= 0 means an inverted element). */
double orc_element_stiffness(int etype, const double *Xe, double E, double nu, double *Ke) {
    const int npe = etype;
    const int nd = 3 * npe;
    double w[8], xi[8][3], dN[10][3], g[10][3], D[6][6];
    double BL[6 * 30], BtD[30 * 6];
    const int ng = orc_volume_rule(etype, w, xi);
    iso_D(E, nu, D);
    memset(Ke, 0, nd * nd * sizeof(double));
    double min_det = INFINITY;
    for (int q = 0; q < ng; ++q) {
        volume_basis(etype, xi[q], NULL, dN);
        const double det = eval_basis(npe, (const double (*)[3])dN, Xe, g);
        if (det < min_det) min_det = det;
        const double wq = w[q] * det;
        build_BL(npe, (const double (*)[3])g, BL);
        /* Bt_mul_D = BL' * D */
        for (int r = 0; r < nd; ++r)
            for (int s = 0; s < 6; ++s) {
                double acc = 0;
                for (int t = 0; t < 6; ++t) acc += BL[t * nd + r] * D[t][s];
                BtD[r * 6 + s] = acc;
            }
        /* Km += w * (Bt_mul_D * BL) */
        for (int r = 0; r < nd; ++r)
            for (int c = 0; c < nd; ++c) {
                double acc = 0;
                for (int t = 0; t < 6; ++t) acc += BtD[r * 6 + t] * BL[t * nd + c];
                Ke[r * nd + c] += wq * acc;
            }
    }
    return min_det;
}

/* Internal force f_int = sum_gp w BL' (D eps), eps = BL u (small strain,
 * src/problems_elasticity.jl:251-265, 406-409). */
void orc_element_internal_force(int etype, const double *Xe, const double *ue, double E,
                                double nu, double *fe) {
    const int npe = etype, nd = 3 * npe;
    double w[8], xi[8][3], dN[10][3], g[10][3], D[6][6], BL[6 * 30];
    const int ng = orc_volume_rule(etype, w, xi);
    iso_D(E, nu, D);
    memset(fe, 0, nd * sizeof(double));
    for (int q = 0; q < ng; ++q) {
        volume_basis(etype, xi[q], NULL, dN);
        const double det = eval_basis(npe, (const double (*)[3])dN, Xe, g);
        build_BL(npe, (const double (*)[3])g, BL);
        double eps[6] = {0}, sig[6] = {0};
        for (int t = 0; t < 6; ++t)
            for (int c = 0; c < nd; ++c) eps[t] += BL[t * nd + c] * ue[c];
        for (int t = 0; t < 6; ++t)
            for (int s = 0; s < 6; ++s) sig[t] += D[t][s] * eps[s];
        for (int c = 0; c < nd; ++c) {
            double acc = 0;
            for (int t = 0; t < 6; ++t) acc += BL[t * nd + c] * sig[t];
            fe[c] += w[q] * det * acc;
        }
    }
}

/* Consistent surface load f += w T N, w = ip.weight * detJ with
 * detJ = |dX/du x dX/dv| (src/problems_elasticity.jl:454-502). */
void orc_surface_load(int ftype, const double *Xf, const double *t, double *fe) {
    const int nfn = ftype;
    double w[4], xi[4][2], N[6], dN[6][2];
    const int ng = surface_rule(ftype, w, xi);
    memset(fe, 0, 3 * nfn * sizeof(double));
    for (int q = 0; q < ng; ++q) {
        if (ftype == FT_TRI6) tri6_basis(xi[q], N, dN);
        else if (ftype == FT_TRI3) tri3_basis(xi[q], N, dN);
        else quad4_basis(xi[q], N, dN);
        double J[2][3] = {{0}};
        for (int i = 0; i < nfn; ++i)
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 3; ++b) J[a][b] += dN[i][a] * Xf[3 * i + b];
        const double cx = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        const double cy = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        const double cz = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double wq = w[q] * sqrt(cx * cx + cy * cy + cz * cz);
        for (int i = 0; i < nfn; ++i)
            for (int c = 0; c < 3; ++c) fe[3 * i + c] += wq * t[c] * N[i];
    }
}

static void gather_X(const double *X, const int64_t *cn, int npe, double *Xe) {
    for (int i = 0; i < npe; ++i)
        for (int c = 0; c < 3; ++c) Xe[3 * i + c] = X[3 * cn[i] + c];
}

/* ---------------------------------------------------- global assembly -- */

/* Number of COO triplets the element loop pushes (npe^2 * 9 per element). */
int64_t orc_coo_size(int64_t n_el, int etype) { return n_el * (int64_t)(9 * etype * etype); }

/* Element loop: add!(K, gdofs, gdofs, Km) (src/sparse/sparse.jl:121-132) pushes
 * Km column by column, so triplet k of element e is (gdofs[i], gdofs[j], Km[i,j])
 * with k = j*nd + i.  Returns min detJ over all elements. */
double orc_assemble_coo(const double *X, int64_t n_el, int etype, const int64_t *conn,
                        const double *E, const double *nu, int64_t *I, int64_t *Jc, double *V,
                        int nthreads) {
    const int npe = etype, nd = 3 * npe;
    double min_det = INFINITY;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(min : min_det)
#endif
    for (int64_t e = 0; e < n_el; ++e) {
        double Xe[30], Ke[900];
        int64_t gd[30];
        const int64_t *cn = conn + e * npe;
        gather_X(X, cn, npe, Xe);
        const double d = orc_element_stiffness(etype, Xe, E[e], nu[e], Ke);
        if (d < min_det) min_det = d;
        for (int i = 0; i < npe; ++i)
            for (int c = 0; c < 3; ++c) gd[3 * i + c] = 3 * cn[i] + c;
        int64_t k = e * nd * nd;
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i < nd; ++i, ++k) { I[k] = gd[i]; Jc[k] = gd[j]; V[k] = Ke[i * nd + j]; }
    }
    return min_det;
}

/* sparse(I, J, V, n, n) semantics (SparseArrays): duplicates are summed in the
 * order they appear in the COO arrays and explicitly stored zeros are kept.
 * Output is CSR (rows sorted, columns sorted within a row); K is structurally
 * symmetric, so this is also the CSC pattern of the reference.
 * rowptr has n+1 entries; colind/vals need room for the returned nnz, which
 * is at most m (the COO length). Scratch `perm` needs m entries. */
int64_t orc_coo_to_csr(int64_t n, int64_t m, const int64_t *I, const int64_t *Jc, const double *V,
                       int64_t *rowptr, int64_t *colind, double *vals, int64_t *perm) {
    /* stable counting sort by row */
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < m; ++k) cnt[I[k] + 1]++;
    for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    int64_t *pos = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    memcpy(pos, cnt, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < m; ++k) perm[pos[I[k]]++] = k;
    free(pos);
    int64_t *tmp = (int64_t *)malloc((size_t)m * sizeof(int64_t) + 1);
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        int64_t *seg = perm + cnt[r];
        const int64_t len = cnt[r + 1] - cnt[r];
        /* stable bottom-up merge sort of seg by column */
        int64_t *src = seg, *dst = tmp;
        for (int64_t width = 1; width < len; width *= 2) {
            for (int64_t lo = 0; lo < len; lo += 2 * width) {
                int64_t mid = lo + width < len ? lo + width : len;
                int64_t hi = lo + 2 * width < len ? lo + 2 * width : len;
                int64_t a = lo, b = mid, o = lo;
                while (a < mid && b < hi) dst[o++] = (Jc[src[b]] < Jc[src[a]]) ? src[b++] : src[a++];
                while (a < mid) dst[o++] = src[a++];
                while (b < hi) dst[o++] = src[b++];
            }
            int64_t *t = src; src = dst; dst = t;
        }
        if (src != seg) memcpy(seg, src, (size_t)len * sizeof(int64_t));
        for (int64_t q = 0; q < len;) {
            const int64_t col = Jc[seg[q]];
            double acc = 0.0;
            while (q < len && Jc[seg[q]] == col) acc += V[seg[q++]];
            colind[nnz] = col;
            vals[nnz] = acc;
            ++nnz;
        }
        rowptr[r + 1] = nnz;
    }
    free(tmp);
    free(cnt);
    return nnz;
}

/* External load vector from surface elements, pushed face by face and summed
 * in that order (add!(assembly.f, gdofs, f), src/sparse/sparse.jl:142-150). */
void orc_load_vector(int64_t n_dofs, const double *X, int64_t n_faces, int ftype,
                     const int64_t *faces, const double *traction /*3 per face*/, double *f) {
    memset(f, 0, (size_t)n_dofs * sizeof(double));
    for (int64_t k = 0; k < n_faces; ++k) {
        double Xf[18], fe[18];
        const int64_t *cn = faces + k * ftype;
        gather_X(X, cn, ftype, Xf);
        orc_surface_load(ftype, Xf, traction + 3 * k, fe);
        for (int i = 0; i < ftype; ++i)
            for (int c = 0; c < 3; ++c) f[3 * cn[i] + c] += fe[3 * i + c];
    }
}

/* y = K x */
void orc_spmv(int64_t n, const int64_t *rowptr, const int64_t *colind, const double *vals,
              const double *x, double *y, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t r = 0; r < n; ++r) {
        double acc = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) acc += vals[k] * x[colind[k]];
        y[r] = acc;
    }
}

/* --------------------------------------------------------- direct solve -- */

/* solve!(..., Val{1}) (src/solvers.jl:192-216) with C1 = C2 = I on the fixed
 * DOFs B (src/problems_dirichlet.jl:60-89):
 *   u[B] = g[B];  K[I,I] u[I] = f[I] - K[I,B] u[B];  la[B] = f[B] - K[B,:] u.
 * K[I,I] is factorised by a banded Cholesky (LAPACK dpbtrf-style, lower band).
 * Returns 0 on success, -1 if K[I,I] is not positive definite, -2 on OOM.
 * *bandwidth_out receives the half-bandwidth used. */
int orc_solve_direct(int64_t n, const int64_t *rowptr, const int64_t *colind, const double *vals,
                     const uint8_t *fixed, const double *g, const double *f, double *u,
                     double *la, int64_t *bandwidth_out) {
    int64_t *map = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t nf = 0;
    for (int64_t i = 0; i < n; ++i) map[i] = fixed[i] ? -1 : nf++;
    int64_t bw = 0;
    for (int64_t r = 0; r < n; ++r) {
        if (map[r] < 0) continue;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int64_t c = colind[k];
            if (map[c] >= 0 && map[r] - map[c] > bw) bw = map[r] - map[c];
        }
    }
    if (bandwidth_out) *bandwidth_out = bw;
    const int64_t ld = bw + 1;
    double *AB = (double *)calloc((size_t)(ld * nf), sizeof(double));
    double *rhs = (double *)malloc((size_t)(nf ? nf : 1) * sizeof(double));
    if (!AB || !rhs) { free(map); free(AB); free(rhs); return -2; }
    for (int64_t i = 0; i < n; ++i) u[i] = fixed[i] ? g[i] : 0.0;
    for (int64_t r = 0; r < n; ++r) {
        if (map[r] < 0) continue;
        double b = f[r];
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int64_t c = colind[k];
            if (map[c] < 0) b -= vals[k] * u[c];
            else if (map[c] <= map[r]) AB[map[c] * ld + (map[r] - map[c])] += vals[k];
        }
        rhs[map[r]] = b;
    }
    /* A = L L', column j of L stored in AB[j*ld + 0..bw] */
    for (int64_t j = 0; j < nf; ++j) {
        double *cj = AB + j * ld;
        if (!(cj[0] > 0.0)) { free(map); free(AB); free(rhs); return -1; }
        const double d = sqrt(cj[0]);
        cj[0] = d;
        const int64_t m = (bw < nf - 1 - j) ? bw : nf - 1 - j;
        const double inv = 1.0 / d;
        for (int64_t k = 1; k <= m; ++k) cj[k] *= inv;
        for (int64_t c = 1; c <= m; ++c) {
            const double lc = cj[c];
            if (lc == 0.0) continue;
            double *cc = AB + (j + c) * ld;
            for (int64_t r = c; r <= m; ++r) cc[r - c] -= cj[r] * lc;
        }
    }
    /* L y = rhs, then L' x = y */
    for (int64_t j = 0; j < nf; ++j) {
        const double *cj = AB + j * ld;
        rhs[j] /= cj[0];
        const int64_t m = (bw < nf - 1 - j) ? bw : nf - 1 - j;
        for (int64_t k = 1; k <= m; ++k) rhs[j + k] -= cj[k] * rhs[j];
    }
    for (int64_t j = nf - 1; j >= 0; --j) {
        const double *cj = AB + j * ld;
        const int64_t m = (bw < nf - 1 - j) ? bw : nf - 1 - j;
        double acc = rhs[j];
        for (int64_t k = 1; k <= m; ++k) acc -= cj[k] * rhs[j + k];
        rhs[j] = acc / cj[0];
    }
    for (int64_t i = 0; i < n; ++i)
        if (map[i] >= 0) u[i] = rhs[map[i]];
    for (int64_t r = 0; r < n; ++r) {
        la[r] = 0.0;
        if (map[r] >= 0) continue;
        double acc = f[r];
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) acc -= vals[k] * u[colind[k]];
        la[r] = acc;
    }
    free(map); free(AB); free(rhs);
    return 0;
}

/* -------------------------------------------------- block-Jacobi PCG ----- */

/* 3x3 node-block Jacobi PCG on the Dirichlet-eliminated system, written as a
 * masked operator on all DOFs (fixed rows/columns removed).  Stops when
 * |r|/|b| <= rtol (recursive residual).  x holds the initial guess on entry
 * (its fixed entries are ignored).  There is no preconditioned CG in the
 * reference (its CG is plain, ext/JuliaFEMCUDAExt.jl:531-577); this is the
 * CPU counterpart of the library's FP64 method. */
int orc_pcg_bj(int64_t n, const int64_t *rowptr, const int64_t *colind, const double *vals,
               const uint8_t *fixed, const double *b_in, double *x, double rtol, int maxit,
               int *iters_out, double *relres_out, int nthreads) {
    const int64_t nn = n / 3;
    double *Minv = (double *)malloc((size_t)nn * 9 * sizeof(double));
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *q = (double *)malloc((size_t)n * sizeof(double));
    /* diagonal blocks with fixed rows/cols replaced by identity, inverted,
     * then fixed rows/cols zeroed */
    for (int64_t a = 0; a < nn; ++a) {
        double B[3][3];
        for (int c = 0; c < 3; ++c) {
            const int64_t row = 3 * a + c;
            for (int d = 0; d < 3; ++d) B[c][d] = 0.0;
            for (int64_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
                const int64_t col = colind[k];
                if (col >= 3 * a && col < 3 * a + 3) B[c][col - 3 * a] = vals[k];
            }
        }
        for (int c = 0; c < 3; ++c)
            if (fixed[3 * a + c])
                for (int d = 0; d < 3; ++d) { B[c][d] = (c == d); B[d][c] = (c == d); }
        const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1])
                         - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0])
                         + B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
        double *M = Minv + 9 * a;
        M[0] = (B[1][1] * B[2][2] - B[1][2] * B[2][1]) / det;
        M[1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) / det;
        M[2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) / det;
        M[3] = (B[1][2] * B[2][0] - B[1][0] * B[2][2]) / det;
        M[4] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) / det;
        M[5] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) / det;
        M[6] = (B[1][0] * B[2][1] - B[1][1] * B[2][0]) / det;
        M[7] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) / det;
        M[8] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) / det;
        for (int c = 0; c < 3; ++c)
            if (fixed[3 * a + c])
                for (int d = 0; d < 3; ++d) { M[3 * c + d] = 0.0; M[3 * d + c] = 0.0; }
    }
    for (int64_t i = 0; i < n; ++i) if (fixed[i]) x[i] = 0.0;
    orc_spmv(n, rowptr, colind, vals, x, q, nthreads);
    double bb = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double bi = fixed[i] ? 0.0 : b_in[i];
        r[i] = fixed[i] ? 0.0 : bi - q[i];
        bb += bi * bi;
    }
    const double bnorm = sqrt(bb);
    double rr = 0.0, rz = 0.0;
    for (int64_t a = 0; a < nn; ++a) {
        const double *M = Minv + 9 * a;
        for (int c = 0; c < 3; ++c) {
            z[3 * a + c] = M[3 * c] * r[3 * a] + M[3 * c + 1] * r[3 * a + 1] + M[3 * c + 2] * r[3 * a + 2];
        }
    }
    for (int64_t i = 0; i < n; ++i) { p[i] = z[i]; rz += r[i] * z[i]; rr += r[i] * r[i]; }
    int it = 0;
    double rel = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    while (rel > rtol && it < maxit) {
        orc_spmv(n, rowptr, colind, vals, p, q, nthreads);
        double pq = 0.0;
        for (int64_t i = 0; i < n; ++i) { if (fixed[i]) q[i] = 0.0; pq += p[i] * q[i]; }
        const double alpha = rz / pq;
        rr = 0.0;
        for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rr += r[i] * r[i]; }
        double rz_new = 0.0;
        for (int64_t a = 0; a < nn; ++a) {
            const double *M = Minv + 9 * a;
            for (int c = 0; c < 3; ++c)
                z[3 * a + c] = M[3 * c] * r[3 * a] + M[3 * c + 1] * r[3 * a + 1] + M[3 * c + 2] * r[3 * a + 2];
        }
        for (int64_t i = 0; i < n; ++i) rz_new += r[i] * z[i];
        const double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
        rel = sqrt(rr) / bnorm;
    }
    if (iters_out) *iters_out = it;
    if (relres_out) *relres_out = rel;
    free(Minv); free(r); free(z); free(p); free(q);
    return rel <= rtol ? 0 : 1;
}

/* ------------------------------------------- two-phase matrix-free K.u -- */

/* Node -> element CSR in ascending element id (ext/JuliaFEMCUDAExt.jl:179-193).
 * ptr has N+1 entries, data n_el*npe entries (element ids). */
void orc_node_to_elem(int64_t N, int64_t n_el, int npe, const int64_t *conn, int64_t *ptr,
                      int64_t *data) {
    memset(ptr, 0, ((size_t)N + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < n_el; ++e)
        for (int i = 0; i < npe; ++i) ptr[conn[e * npe + i] + 1]++;
    for (int64_t a = 0; a < N; ++a) ptr[a + 1] += ptr[a];
    int64_t *pos = (int64_t *)malloc((size_t)N * sizeof(int64_t));
    memcpy(pos, ptr, (size_t)N * sizeof(int64_t));
    for (int64_t e = 0; e < n_el; ++e)
        for (int i = 0; i < npe; ++i) data[pos[conn[e * npe + i]]++] = e;
    free(pos);
}

/* Two-phase internal force (ext/JuliaFEMCUDAExt.jl:222-361): phase 1 stores
 * the stress at every Gauss point; phase 2 loops nodes, visits the node's
 * elements in ascending id, recomputes grad N of the local node and adds
 * w detJ (grad N . sigma).  For linear elasticity this is y = K u.
 * Rows of fixed DOFs are zeroed when `fixed` is non-NULL. */
void orc_matfree_apply(int64_t N, const double *X, int64_t n_el, int etype, const int64_t *conn,
                       const double *E, const double *nu, const uint8_t *fixed, const double *u,
                       double *y) {
    const int npe = etype;
    double w[8], xi[8][3];
    const int ng = orc_volume_rule(etype, w, xi);
    double *sig = (double *)malloc((size_t)n_el * ng * 6 * sizeof(double));
    for (int64_t e = 0; e < n_el; ++e) {
        double Xe[30], ue[30], dN[10][3], g[10][3], D[6][6], BL[6 * 30];
        const int64_t *cn = conn + e * npe;
        gather_X(X, cn, npe, Xe);
        gather_X(u, cn, npe, ue);
        iso_D(E[e], nu[e], D);
        for (int q = 0; q < ng; ++q) {
            volume_basis(etype, xi[q], NULL, dN);
            eval_basis(npe, (const double (*)[3])dN, Xe, g);
            build_BL(npe, (const double (*)[3])g, BL);
            double eps[6] = {0};
            for (int t = 0; t < 6; ++t)
                for (int c = 0; c < 3 * npe; ++c) eps[t] += BL[t * 3 * npe + c] * ue[c];
            for (int t = 0; t < 6; ++t) {
                double acc = 0;
                for (int s = 0; s < 6; ++s) acc += D[t][s] * eps[s];
                sig[(e * ng + q) * 6 + t] = acc;
            }
        }
    }
    int64_t *ptr = (int64_t *)malloc(((size_t)N + 1) * sizeof(int64_t));
    int64_t *data = (int64_t *)malloc((size_t)n_el * npe * sizeof(int64_t));
    orc_node_to_elem(N, n_el, npe, conn, ptr, data);
    for (int64_t a = 0; a < N; ++a) {
        double fa[3] = {0, 0, 0};
        for (int64_t k = ptr[a]; k < ptr[a + 1]; ++k) {
            const int64_t e = data[k];
            const int64_t *cn = conn + e * npe;
            int loc = 0;
            while (cn[loc] != a) ++loc;
            double Xe[30], dN[10][3], g[10][3];
            gather_X(X, cn, npe, Xe);
            for (int q = 0; q < ng; ++q) {
                volume_basis(etype, xi[q], NULL, dN);
                const double det = eval_basis(npe, (const double (*)[3])dN, Xe, g);
                const double *s = sig + (e * ng + q) * 6;
                /* sigma as a symmetric 3x3 from Voigt 11,22,33,12,23,13 */
                const double S[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
                for (int c = 0; c < 3; ++c)
                    fa[c] += (g[loc][0] * S[0][c] + g[loc][1] * S[1][c] + g[loc][2] * S[2][c]) * (w[q] * det);
            }
        }
        for (int c = 0; c < 3; ++c) y[3 * a + c] = (fixed && fixed[3 * a + c]) ? 0.0 : fa[c];
    }
    free(sig); free(ptr); free(data);
}

/* ------------------------------------------------------------ exports ---- */

/* Basis values N and derivatives dN (npe x dim) at xi for volume (4,8,10) or
 * surface (3,4,6 with surface=1) elements; exported for the closed-form tests. */
void orc_basis(int type, int surface, const double *xi, double *N, double *dN) {
    if (!surface) {
        volume_basis(type, xi, N, (double (*)[3])dN);
    } else if (type == FT_TRI6) {
        tri6_basis(xi, N, (double (*)[2])dN);
    } else if (type == FT_TRI3) {
        tri3_basis(xi, N, (double (*)[2])dN);
    } else {
        quad4_basis(xi, N, (double (*)[2])dN);
    }
}

int orc_surface_rule(int ftype, double *w, double *xi) { return surface_rule(ftype, w, (double (*)[2])xi); }

/* ------------------------------------- neo-Hookean total Lagrangian ---- */
/* Compressible neo-Hookean (src/materials/neo_hookean.jl:129-137):
 *   psi = mu/2 (tr C - 3) - mu ln J + lambda/2 ln^2 J,  J = sqrt(det C),
 * evaluated like the reference's compute_stress(material, E) (:205-231) from
 * the Green-Lagrange strain E (3x3), C = I + 2E.  The reference differentiates
 * psi by AD; here S = dpsi/dE and D = d2psi/dE2 are written in closed form with
 * A = C^-1:
 *   S    = mu A (C - I) + lambda ln J A        (= mu (I - A) + lambda ln J A),
 *   D_ijkl = lambda A_ij A_kl + (mu - lambda ln J)(A_ik A_jl + A_il A_jk),
 * D in Voigt order 11,22,33,12,23,13 (engineering shear).  mu A (C - I) and
 * ln J = log1p(det C - 1)/2 with det C - 1 = I1 + I2 + I3 of 2E keep full
 * relative precision at small strain, where mu (I - A) cancels.  Returns psi,
 * or NAN when det C <= 0.  S (3x3) and D (6x6) may be NULL. */
static const int VI[6] = {0, 1, 2, 0, 1, 0}, VJ[6] = {0, 1, 2, 1, 2, 2};

static double det3(const double *M) {
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6])
         + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* det(I + M) - 1 = I1(M) + I2(M) + I3(M) */
static double det_ipm_minus1(const double *M) {
    const double i1 = M[0] + M[4] + M[8];
    const double i2 = M[0] * M[4] - M[1] * M[3] + M[4] * M[8] - M[5] * M[7] + M[0] * M[8] - M[2] * M[6];
    return i1 + i2 + det3(M);
}

double orc_nh_material(const double *Eg, double mu, double la, double *S, double *D) {
    double E2[9], C[9];
    for (int k = 0; k < 9; ++k) {
        E2[k] = 2.0 * Eg[k];
        C[k] = E2[k] + (k % 4 == 0 ? 1.0 : 0.0);
    }
    const double dm1 = det_ipm_minus1(E2);
    const double det = 1.0 + dm1;
    if (!(det > 0.0)) return NAN;
    const double lnJ = 0.5 * log1p(dm1);
    double A[9];
    A[0] = (C[4] * C[8] - C[5] * C[7]) / det;
    A[1] = (C[2] * C[7] - C[1] * C[8]) / det;
    A[2] = (C[1] * C[5] - C[2] * C[4]) / det;
    A[3] = (C[5] * C[6] - C[3] * C[8]) / det;
    A[4] = (C[0] * C[8] - C[2] * C[6]) / det;
    A[5] = (C[2] * C[3] - C[0] * C[5]) / det;
    A[6] = (C[3] * C[7] - C[4] * C[6]) / det;
    A[7] = (C[1] * C[6] - C[0] * C[7]) / det;
    A[8] = (C[0] * C[4] - C[1] * C[3]) / det;
    if (S) {
        double AE[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                AE[i * 3 + j] = A[i * 3] * E2[j] + A[i * 3 + 1] * E2[3 + j] + A[i * 3 + 2] * E2[6 + j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                S[i * 3 + j] = 0.5 * mu * (AE[i * 3 + j] + AE[j * 3 + i]) + la * lnJ * A[i * 3 + j];
    }
    if (D)
        for (int I = 0; I < 6; ++I)
            for (int J = 0; J < 6; ++J) {
                const int i = VI[I], j = VJ[I], k = VI[J], l = VJ[J];
                D[I * 6 + J] = la * A[i * 3 + j] * A[k * 3 + l]
                             + (mu - la * lnJ) * (A[i * 3 + k] * A[j * 3 + l] + A[i * 3 + l] * A[j * 3 + k]);
            }
    return 0.5 * mu * (E2[0] + E2[4] + E2[8]) - mu * lnJ + 0.5 * la * lnJ * lnJ;
}

/* Total-Lagrangian element (src/problems_elasticity.jl:246-289 for F, E and
 * BL(F); :369-375 material stiffness; :378-404 geometric stiffness with
 * BNL and S3 = blockdiag(S, S, S); :406-409 f_int), Lame parameters from E and
 * nu as in :310-317:
 *   f  = sum_gp w BL(F)' S_voigt,
 *   Kt = sum_gp w (BL(F)' D BL(F) + BNL' S3 BNL),   w = ip.weight * detJ.
 * ue is (3npe) in node-major order; Kt (row-major) may be NULL.  Returns the
 * smallest det F seen (<= 0 or NAN: the state is not admissible). */
double orc_nh_element(int etype, const double *Xe, const double *ue, double E, double nu, double *fe, double *Ke) {
    const int npe = etype, nd = 3 * npe;
    const double la = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2.0 * (1.0 + nu));
    double w[8], xi[8][3], dN[10][3], g[10][3];
    double BL[6 * 30], S[9], D[36];
    const int ng = orc_volume_rule(etype, w, xi);
    memset(fe, 0, nd * sizeof(double));
    if (Ke) memset(Ke, 0, nd * nd * sizeof(double));
    double min_detF = INFINITY;
    for (int q = 0; q < ng; ++q) {
        volume_basis(etype, xi[q], NULL, dN);
        const double det = eval_basis(npe, (const double (*)[3])dN, Xe, g);
        const double wq = w[q] * det;
        double H[9] = {0};  /* H[a][b] = du_a / dX_b */
        for (int i = 0; i < npe; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) H[a * 3 + b] += ue[3 * i + a] * g[i][b];
        /* F = I + H, E = (H + H' + H'H) / 2 (src/problems_elasticity.jl:258-260) */
        double F[9], Eg[9];
        for (int k = 0; k < 9; ++k) F[k] = H[k] + (k % 4 == 0 ? 1.0 : 0.0);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                Eg[a * 3 + b] = 0.5 * (H[a * 3 + b] + H[b * 3 + a] + H[a] * H[b] + H[3 + a] * H[3 + b] + H[6 + a] * H[6 + b]);
        const double detF = 1.0 + det_ipm_minus1(H);
        if (!(detF >= min_detF)) min_detF = detF;
        if (!(detF > 0.0)) continue;
        orc_nh_material(Eg, mu, la, S, Ke ? D : NULL);
        const double Sv[6] = {S[0], S[4], S[8], S[1], S[5], S[2]};
        /* BL(F): row I of strain I, column 3i + a (F[a][b] = F[a * 3 + b]) */
        memset(BL, 0, 6 * nd * sizeof(double));
        for (int i = 0; i < npe; ++i)
            for (int a = 0; a < 3; ++a) {
                const int c = 3 * i + a;
                BL[0 * nd + c] = F[a * 3 + 0] * g[i][0];
                BL[1 * nd + c] = F[a * 3 + 1] * g[i][1];
                BL[2 * nd + c] = F[a * 3 + 2] * g[i][2];
                BL[3 * nd + c] = F[a * 3 + 0] * g[i][1] + F[a * 3 + 1] * g[i][0];
                BL[4 * nd + c] = F[a * 3 + 1] * g[i][2] + F[a * 3 + 2] * g[i][1];
                BL[5 * nd + c] = F[a * 3 + 2] * g[i][0] + F[a * 3 + 0] * g[i][2];
            }
        for (int c = 0; c < nd; ++c) {
            double acc = 0.0;
            for (int t = 0; t < 6; ++t) acc += BL[t * nd + c] * Sv[t];
            fe[c] += wq * acc;
        }
        if (!Ke) continue;
        for (int r = 0; r < nd; ++r) {
            double BtD[6];
            for (int s = 0; s < 6; ++s) {
                double acc = 0.0;
                for (int t = 0; t < 6; ++t) acc += BL[t * nd + r] * D[t * 6 + s];
                BtD[s] = acc;
            }
            for (int c = 0; c < nd; ++c) {
                double acc = 0.0;
                for (int t = 0; t < 6; ++t) acc += BtD[t] * BL[t * nd + c];
                Ke[r * nd + c] += wq * acc;
            }
        }
        /* BNL' S3 BNL couples equal components: (3i+a, 3j+a) += g_i . S g_j */
        for (int i = 0; i < npe; ++i)
            for (int j = 0; j < npe; ++j) {
                double acc = 0.0;
                for (int b = 0; b < 3; ++b)
                    for (int d = 0; d < 3; ++d) acc += g[i][b] * S[b * 3 + d] * g[j][d];
                for (int a = 0; a < 3; ++a) Ke[(3 * i + a) * nd + 3 * j + a] += wq * acc;
            }
    }
    return min_detF;
}

/* Global f_int(u) (n_dofs) and, when I is not NULL, the tangent as COO
 * triplets in the order of orc_assemble_coo.  Returns the smallest det F. */
double orc_nh_assemble(const double *X, int64_t n_el, int etype, const int64_t *conn, const double *E,
                       const double *nu, const double *u, double *fint, int64_t n_dofs, int64_t *I, int64_t *Jc,
                       double *V) {
    const int npe = etype, nd = 3 * npe;
    double min_detF = INFINITY;
    memset(fint, 0, (size_t)n_dofs * sizeof(double));
    for (int64_t e = 0; e < n_el; ++e) {   /* serial: f_int sums in element order */
        double Xe[30], ue[30], fe[30], Ke[900];
        int64_t gd[30];
        const int64_t *cn = conn + e * npe;
        gather_X(X, cn, npe, Xe);
        for (int i = 0; i < npe; ++i)
            for (int c = 0; c < 3; ++c) { gd[3 * i + c] = 3 * cn[i] + c; ue[3 * i + c] = u[3 * cn[i] + c]; }
        const double d = orc_nh_element(etype, Xe, ue, E[e], nu[e], fe, I ? Ke : NULL);
        if (!(d >= min_detF)) min_detF = d;
        for (int c = 0; c < nd; ++c) fint[gd[c]] += fe[c];
        if (!I) continue;
        int64_t k = e * nd * nd;
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i < nd; ++i, ++k) { I[k] = gd[i]; Jc[k] = gd[j]; V[k] = Ke[i * nd + j]; }
    }
    return min_detF;
}

/* ---------------------------------------------- learned preconditioner ---- */
/* CPU restatement of the learned 3x3 block preconditioner (no reference
 * counterpart; the library's learned.hip).  Features are built from the
 * oracle's own BL'DBL element matrices; the network runs in FP64 with the
 * library's BF16 rounding points (network inputs and hidden activations on
 * the forward pass, output and hidden gradients on the backward pass). */

#define LF 24
#define LH 64
#define LO 6

static int g_emulate_bf16 = 1;  /* round where the library rounds to BF16 */

static double bf16_round(double v) {
    if (!g_emulate_bf16) return v;
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}
static double gelu_d(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_grad_d(double x) { return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) * 0.39894228040143268; }

/* 24 node features (feat N x 24).  Nodes are visited with their elements in
 * ascending id (the incidence order of the nodal assembly). */
void orc_learned_features(int64_t N, const double *X, int64_t n_el, int etype, const int64_t *conn, const double *E,
                          const double *nu, const uint8_t *fixed, const uint8_t *onface, double *feat) {
    const int npe = etype, nd = 3 * npe;
    const int ncorner = etype == ET_TET10 ? 4 : npe;
    int64_t *ptr = (int64_t *)malloc(((size_t)N + 1) * sizeof(int64_t));
    int64_t *data = (int64_t *)malloc((size_t)n_el * npe * sizeof(int64_t));
    orc_node_to_elem(N, n_el, npe, conn, ptr, data);
    /* per element: K_e, quality q_e, has-fixed flag */
    double *Kd = (double *)malloc((size_t)n_el * npe * 9 * sizeof(double)); /* diagonal blocks */
    double *qe = (double *)malloc((size_t)n_el * sizeof(double));
    uint8_t *fe = (uint8_t *)malloc((size_t)n_el);
    for (int64_t e = 0; e < n_el; ++e) {
        double Xe[30], Ke[900], w[8], xi[8][3], dN[10][3], g[10][3];
        const int64_t *cn = conn + e * npe;
        gather_X(X, cn, npe, Xe);
        orc_element_stiffness(etype, Xe, E[e], nu[e], Ke);
        for (int i = 0; i < npe; ++i)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) Kd[((size_t)e * npe + i) * 9 + r * 3 + c] = Ke[(3 * i + r) * nd + 3 * i + c];
        orc_volume_rule(etype, w, xi);
        volume_basis(etype, xi[0], NULL, dN);
        double J[3][3] = {{0}};
        for (int i = 0; i < npe; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) J[a][b] += dN[i][a] * Xe[3 * i + b];
        const double det = eval_basis(npe, (const double (*)[3])dN, Xe, g);
        /* |invJ|_F from the adjugate */
        double fro = 0;
        const double c[9] = {J[1][1] * J[2][2] - J[1][2] * J[2][1], J[0][2] * J[2][1] - J[0][1] * J[2][2],
                             J[0][1] * J[1][2] - J[0][2] * J[1][1], J[1][2] * J[2][0] - J[1][0] * J[2][2],
                             J[0][0] * J[2][2] - J[0][2] * J[2][0], J[0][2] * J[1][0] - J[0][0] * J[1][2],
                             J[1][0] * J[2][1] - J[1][1] * J[2][0], J[0][1] * J[2][0] - J[0][0] * J[2][1],
                             J[0][0] * J[1][1] - J[0][1] * J[1][0]};
        for (int m = 0; m < 9; ++m) fro += (c[m] / det) * (c[m] / det);
        qe[e] = sqrt(fro) * cbrt(det);
        uint8_t f = 0;
        for (int i = 0; i < npe; ++i) f |= fixed[3 * cn[i]] | fixed[3 * cn[i] + 1] | fixed[3 * cn[i] + 2];
        fe[e] = f;
    }
    for (int64_t a = 0; a < N; ++a) {
        double D[9] = {0}, tmin = 1e300, tmax = 0, tsum = 0, tsq = 0, emin = 1e300, emax = 0, esum = 0, lesum = 0,
               tle = 0, qsum = 0, qmin = 1e300, qmax = 0;
        int nc = 0, nfx = 0;
        const int64_t n = ptr[a + 1] - ptr[a];
        double *f = feat + a * LF;
        if (n == 0) { for (int k = 0; k < LF; ++k) f[k] = 0.0; continue; }
        for (int64_t k = ptr[a]; k < ptr[a + 1]; ++k) {
            const int64_t e = data[k];
            int i = 0;
            while (conn[e * npe + i] != a) ++i;
            const double *B = Kd + ((size_t)e * npe + i) * 9;
            for (int m = 0; m < 9; ++m) D[m] += B[m];
            const double te = B[0] + B[4] + B[8];
            tmin = fmin(tmin, te); tmax = fmax(tmax, te); tsum += te; tsq += te * te;
            emin = fmin(emin, E[e]); emax = fmax(emax, E[e]); esum += E[e]; lesum += log(E[e]); tle += te * log(E[e]);
            qsum += qe[e]; qmin = fmin(qmin, qe[e]); qmax = fmax(qmax, qe[e]);
            if (i < ncorner) ++nc;
            if (fe[e]) ++nfx;
        }
        const double emean = esum / n;
        int nab = 0;
        for (int64_t k = ptr[a]; k < ptr[a + 1]; ++k) nab += E[data[k]] > emean;
        const double T = D[0] + D[4] + D[8], s = 3.0 / T;
        const double d00 = D[0] * s, d11 = D[4] * s, d22 = D[8] * s, d01 = D[1] * s, d12 = D[5] * s, d02 = D[2] * s;
        const double tmean = tsum / n;
        f[0] = d00; f[1] = d11; f[2] = d22; f[3] = d01; f[4] = d12; f[5] = d02;
        f[6] = log((double)n);
        f[7] = tmax / tsum;
        f[8] = n * tmin / tsum;
        f[9] = sqrt(fmax(tsq / n - tmean * tmean, 0.0)) / tmean;
        f[10] = log(emax / emin);
        f[11] = log(emean) - lesum / n;
        f[12] = (double)nc / n;
        f[13] = (fixed[3 * a] + fixed[3 * a + 1] + fixed[3 * a + 2]) / 3.0;
        f[14] = (double)nfx / n;
        f[15] = onface[a] ? 1.0 : 0.0;
        f[16] = d00 * (d11 * d22 - d12 * d12) - d01 * (d01 * d22 - d12 * d02) + d02 * (d01 * d12 - d11 * d02);
        f[17] = sqrt(2.0 * (d01 * d01 + d12 * d12 + d02 * d02));
        f[18] = qsum / n; f[19] = qmin; f[20] = qmax;
        f[21] = tle / tsum - log(emax);
        f[22] = (tmax - tmin) / tsum;
        f[23] = (double)nab / n;
    }
    free(ptr); free(data); free(Kd); free(qe); free(fe);
}

typedef struct {
    const double *W1, *W2, *W3;     /* [out][in] (BF16 values when emulating the library) */
    const double *b1, *b2, *b3, *shift, *scale;
} orc_net;

/* forward for one node; keeps pre-activations for the backward pass */
static void net_forward(const orc_net *nt, const double *feat, double *x, double *z1, double *h1, double *z2,
                        double *h2, double *o) {
    for (int k = 0; k < LF; ++k)
        x[k] = bf16_round(g_emulate_bf16 ? (double)(((float)feat[k] - (float)nt->shift[k]) * (float)nt->scale[k])
                                         : (feat[k] - nt->shift[k]) * nt->scale[k]);
    for (int j = 0; j < LH; ++j) {
        double s = 0;
        for (int k = 0; k < LF; ++k) s += nt->W1[j * LF + k] * x[k];
        z1[j] = s + nt->b1[j];
        h1[j] = bf16_round(gelu_d(z1[j]));
    }
    for (int j = 0; j < LH; ++j) {
        double s = 0;
        for (int k = 0; k < LH; ++k) s += nt->W2[j * LH + k] * h1[k];
        z2[j] = s + nt->b2[j];
        h2[j] = bf16_round(gelu_d(z2[j]));
    }
    for (int j = 0; j < LO; ++j) {
        double s = 0;
        for (int k = 0; k < LH; ++k) s += nt->W3[j * LH + k] * h2[k];
        o[j] = s + nt->b3[j];
    }
}

static double clampo(double v) { return v < -5 ? -5 : (v > 5 ? 5 : v); }

static void chol3(const double *S, double *C) {
    const double c00 = sqrt(S[0]), c10 = S[3] / c00, c20 = S[6] / c00;
    const double c11 = sqrt(S[4] - c10 * c10), c21 = (S[7] - c20 * c10) / c11;
    const double c22 = sqrt(S[8] - c20 * c20 - c21 * c21);
    const double t[9] = {c00, 0, 0, c10, c11, 0, c20, c21, c22};
    memcpy(C, t, sizeof t);
}

static void lower_of(const double *o, double *L) {
    const double t[9] = {exp(clampo(o[0])), 0, 0, o[1], exp(clampo(o[2])), 0, o[3], o[4], exp(clampo(o[5]))};
    memcpy(L, t, sizeof t);
}

/* network outputs (N x 6) and blocks M = C (L L' + eps I) C' (N x 9); nodes
 * with a fixed DOF or a zero block keep Minv */
void orc_learned_blocks(int64_t N, const double *feat, const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, const double *b3, const double *shift, const double *scale,
                        double eps, const double *Minv, const uint8_t *fixed, double *out, double *M, int emulate_bf16) {
    g_emulate_bf16 = emulate_bf16;
    orc_net nt = {W1, W2, W3, b1, b2, b3, shift, scale};
    for (int64_t a = 0; a < N; ++a) {
        double x[LF], z1[LH], h1[LH], z2[LH], h2[LH], o[LO];
        net_forward(&nt, feat + a * LF, x, z1, h1, z2, h2, o);
        for (int j = 0; j < LO; ++j) out[a * LO + j] = o[j];
        const double *Mi = Minv + 9 * a;
        double *Ma = M + 9 * a;
        if (fixed[3 * a] || fixed[3 * a + 1] || fixed[3 * a + 2] || !(Mi[0] > 0)) {
            memcpy(Ma, Mi, 9 * sizeof(double));
            continue;
        }
        double C[9], L[9], A[9];
        chol3(Mi, C);
        lower_of(o, L);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = (i == j) ? eps : 0.0;
                for (int k = 0; k < 3; ++k) s += L[i * 3 + k] * L[j * 3 + k];
                A[i * 3 + j] = s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k)
                    for (int m = 0; m < 3; ++m) s += C[i * 3 + k] * A[k * 3 + m] * C[j * 3 + m];
                Ma[i * 3 + j] = s;
            }
    }
}

/* Loss over the pairs and its gradient (6150 doubles, order W1 b1 W2 b2 W3 b3).
 * K is the CSR of the whole system; rows/columns of fixed DOFs are masked. */
double orc_learned_loss_grad(int64_t N, const int64_t *rowptr, const int64_t *colind, const double *vals,
                             const uint8_t *fixed, const double *Minv, const double *feat, const double *W1,
                             const double *b1, const double *W2, const double *b2, const double *W3, const double *b3,
                             const double *shift, const double *scale, double eps, int n_pairs, const double *R,
                             const double *Er, const double *Cc, double *grad, int emulate_bf16) {
    const int64_t n = 3 * N;
    double *out = (double *)malloc((size_t)N * LO * sizeof(double));
    double *M = (double *)malloc((size_t)N * 9 * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *G = (double *)calloc((size_t)N * 9, sizeof(double));
    orc_learned_blocks(N, feat, W1, b1, W2, b2, W3, b3, shift, scale, eps, Minv, fixed, out, M, emulate_bf16);
    double loss = 0;
    for (int p = 0; p < n_pairs; ++p) {
        const double *r = R + (size_t)p * n, *e = Er + (size_t)p * n;
        for (int64_t a = 0; a < N; ++a)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int j = 0; j < 3; ++j) s += M[9 * a + 3 * i + j] * r[3 * a + j];
                z[3 * a + i] = s + (Cc ? Cc[(size_t)p * n + 3 * a + i] : 0.0);
            }
        orc_spmv(n, rowptr, colind, vals, z, w, 1);
        double zr = 0, zw = 0, er = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (fixed[i]) w[i] = 0;
            zr += z[i] * r[i]; zw += z[i] * w[i]; er += e[i] * r[i];
        }
        const double c = zr / zw, coef = -2.0 * c / (er * n_pairs);
        loss += 1.0 - zr * c / er;
        for (int64_t a = 0; a < N; ++a)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) G[9 * a + 3 * i + j] += coef * (r[3 * a + i] - c * w[3 * a + i]) * r[3 * a + j];
    }
    loss /= n_pairs;
    orc_net nt = {W1, W2, W3, b1, b2, b3, shift, scale};
    memset(grad, 0, 6150 * sizeof(double));
    double *gW1 = grad, *gb1 = grad + 1536, *gW2 = grad + 1600, *gb2 = grad + 5696, *gW3 = grad + 5760, *gb3 = grad + 6144;
    for (int64_t a = 0; a < N; ++a) {
        const double *Mi = Minv + 9 * a;
        if (fixed[3 * a] || fixed[3 * a + 1] || fixed[3 * a + 2] || !(Mi[0] > 0)) continue;
        double x[LF], z1[LH], h1[LH], z2[LH], h2[LH], o[LO];
        net_forward(&nt, feat + a * LF, x, z1, h1, z2, h2, o);
        double C[9], L[9], Gs[9], Q[9], dL[9];
        chol3(Mi, C);
        lower_of(o, L);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Gs[i * 3 + j] = G[9 * a + i * 3 + j] + G[9 * a + j * 3 + i];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k)
                    for (int m = 0; m < 3; ++m) s += C[k * 3 + i] * Gs[k * 3 + m] * C[m * 3 + j];
                Q[i * 3 + j] = s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += Q[i * 3 + k] * L[k * 3 + j];
                dL[i * 3 + j] = s;
            }
        double dO[LO];
        dO[0] = (o[0] > -5 && o[0] < 5) ? dL[0] * L[0] : 0.0;
        dO[1] = dL[3];
        dO[2] = (o[2] > -5 && o[2] < 5) ? dL[4] * L[4] : 0.0;
        dO[3] = dL[6];
        dO[4] = dL[7];
        dO[5] = (o[5] > -5 && o[5] < 5) ? dL[8] * L[8] : 0.0;
        double dOb[LO], dz2[LH], dz2b[LH], dz1[LH], dz1b[LH];
        for (int j = 0; j < LO; ++j) { dOb[j] = bf16_round(dO[j]); gb3[j] += dO[j]; }
        for (int j = 0; j < LO; ++j)
            for (int k = 0; k < LH; ++k) gW3[j * LH + k] += dOb[j] * h2[k];
        for (int k = 0; k < LH; ++k) {
            double s = 0;
            for (int j = 0; j < LO; ++j) s += dOb[j] * W3[j * LH + k];
            dz2[k] = s * gelu_grad_d(z2[k]);
            dz2b[k] = bf16_round(dz2[k]);
            gb2[k] += dz2[k];
        }
        for (int j = 0; j < LH; ++j)
            for (int k = 0; k < LH; ++k) gW2[j * LH + k] += dz2b[j] * h1[k];
        for (int k = 0; k < LH; ++k) {
            double s = 0;
            for (int j = 0; j < LH; ++j) s += dz2b[j] * W2[j * LH + k];
            dz1[k] = s * gelu_grad_d(z1[k]);
            dz1b[k] = bf16_round(dz1[k]);
            gb1[k] += dz1[k];
        }
        for (int j = 0; j < LH; ++j)
            for (int k = 0; k < LF; ++k) gW1[j * LF + k] += dz1b[j] * x[k];
    }
    free(out); free(M); free(z); free(w); free(G);
    return loss;
}
