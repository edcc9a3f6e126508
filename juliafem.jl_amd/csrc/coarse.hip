// coarse.hip -- two-level preconditioner: rigid-body modes on node aggregates.
//
// No reference counterpart (the reference solves with CHOLMOD or plain CG).
// Block Jacobi and per-node learned blocks act on one node at a time; the
// smooth bending modes of a slender body, which set the PCG iteration count,
// need a global correction.  This level adds one:
//
//   M^-1 r = B r + Q A_c^-1 Q' r,    A_c = Q' K Q,
//
// additive and symmetric, so PCG stays valid.  B is block Jacobi (or the
// learned or user blocks).  Q holds, per aggregate of nodes, an orthonormal
// basis of the six rigid-body modes (three translations, three rotations
// about the aggregate centroid) restricted to its free DOFs: Q_g = B_g R_g^-1
// with B_g the masked modes and R_g from the Cholesky factor of B_g' B_g.
//
// Setup (once per operator):
//  * host: aggregates from a coordinate grid sized for `coarse_nodes` nodes
//    each, tiny aggregates merged into a neighbour; aggregate adjacency from
//    the elements; a greedy distance-2 colouring; R_g^-1 per aggregate;
//  * device: A_c by probing -- for every colour c and mode j, K applied to
//    the sum of column j of Q over the aggregates of colour c.  Aggregates of
//    one colour are at distance >= 3, so the restriction of that product to
//    aggregate g is Q_g' K Q_{p,j} for the single colour-c aggregate p coupled
//    to g.  No atomics; the result is deterministic.
//  * a block Gauss-Jordan sweep inverts A_c (dense, a few thousand rows; one
//    launch per 32-row pivot block, between two copies of the matrix); the
//    inverse is folded with
//    the R_g^-1 into Ahat = R^-1 A_c^-1 R^-T (stored FP32), so the iteration
//    needs only the raw modes: Q A_c^-1 Q' r = Bm Ahat Bm' r.
//  * the whole device part is one launch sequence per aggregation and
//    operator, captured once as a HIP graph and replayed.
//
// Per PCG iteration:
//  k_pcg_update_agg   r, z = B r and the partials of r.Br and r.r, one
//                     block per aggregate over its nodes; the block also
//                     restricts the new residual, w_g = sum_{a in g}
//                     [m_a r_a ; d_a x m_a r_a], from the values it holds
//  k_coarse_gemv      zc = Ahat w, partials of w.zc             (wave per row)
//  k_pcg_dir_coarse   x += alpha p, then p_a = z_a + m_a o (t_g + omega_g x d_a) + beta p_a
// (JFEM_FUSED_RESTRICTION=0: the node-ordered update writes m o r in
// aggregate order and k_coarse_restrict sums it after the update instead.)
// The prolongation is folded into the direction update, and r.z of the full
// preconditioned residual is r.Br + w.zc (r.(m o Q zc) = (Q' m r).zc), so the
// corrected z is never stored.  At the start of PCG, k_coarse_prolong adds
// the correction to z explicitly.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "jfem_device.h"
#include "jfem_learned.h"
#include "mf_element.h"

using namespace jfem;

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// v_j = sum over aggregates of colour c of Q_g e_j (masked) on the listed
// nodes (a superset of the colour's nodes; v_j is zero elsewhere and stays
// so); mode j = blockIdx.y, v_j = v + j vs
__global__ __launch_bounds__(256) void k_coarse_probe(int nn, const int32_t *__restrict__ nlist,
                                                     const int32_t *__restrict__ node_agg,
                                                     const int32_t *__restrict__ agg_color, int c,
                                                     const double *__restrict__ cd, const double *__restrict__ Rinv,
                                                     const uint8_t *__restrict__ fixed, double *__restrict__ v,
                                                     int64_t vs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int j = blockIdx.y;
    v += j * vs;
    const int a = nlist[k];
    const int g = node_agg[a];
    double out[3] = {0.0, 0.0, 0.0};
    if (agg_color[g] == c) {
        const double *R = Rinv + 36 * (size_t)g;
        double q[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = R[k * 6 + j];
        const double d0 = cd[3 * (size_t)a], d1 = cd[3 * (size_t)a + 1], d2 = cd[3 * (size_t)a + 2];
        // t + omega x d
        out[0] = q[0] + q[4] * d2 - q[5] * d1;
        out[1] = q[1] + q[5] * d0 - q[3] * d2;
        out[2] = q[2] + q[3] * d1 - q[4] * d0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) v[3 * (size_t)a + k] = fixed[3 * (size_t)a + k] ? 0.0 : out[k];
}

// A_c rows of aggregate g for probe (c, j): Rinv_g' w~_g into column 6p + j,
// with w~_g summed over g's nodes next to colour c only (segment blockIdx.x of
// the colour's node list; the product is zero on g's other nodes).  Mode j =
// blockIdx.y, its product at y + j ys.
__global__ __launch_bounds__(256) void k_coarse_probe_restrict(const int32_t *__restrict__ segb,
                                                              const int32_t *__restrict__ sega,
                                                              const int32_t *__restrict__ ncol,
                                                              const double *__restrict__ cd, const double *__restrict__ y,
                                                              int64_t ys, const int32_t *__restrict__ partner_c,
                                                              const double *__restrict__ Rinv, double *__restrict__ Ac,
                                                              int nc) {
    const int sg = blockIdx.x, j = blockIdx.y;
    y += j * ys;
    const int g = sega[sg];
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = segb[sg] + threadIdx.x; k < segb[sg + 1]; k += blockDim.x) {
        const size_t o = 3 * (size_t)ncol[k];
        const double y0 = y[o], y1 = y[o + 1], y2 = y[o + 2];
        const double d0 = cd[o], d1 = cd[o + 1], d2 = cd[o + 2];
        w[0] += y0; w[1] += y1; w[2] += y2;
        w[3] += d1 * y2 - d2 * y1;
        w[4] += d2 * y0 - d0 * y2;
        w[5] += d0 * y1 - d1 * y0;
    }
    block_sum_all<6>(w);
    // an aggregate next to colour c always has its colour-c partner
    const int p = partner_c[g];
    if (p >= 0 && threadIdx.x < 6) {
        const int i = threadIdx.x;
        const double *R = Rinv + 36 * (size_t)g;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += R[k * 6 + i] * w[k];
        Ac[(size_t)(6 * g + i) * nc + 6 * p + j] = s;
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(double *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Rows of the colour's nodes of y_j = K x_j: the sums of their incidence
// slots in the colour's elements only (ascending element id), so the slots of
// other elements, stale from earlier colours, are never read; mode j =
// blockIdx.y (slots at ye + j ys, rows at y + j vs).  Fixed rows are zero.
__global__ __launch_bounds__(256) void k_coarse_gather(int nn, const int32_t *__restrict__ nlist,
                                                      const int32_t *__restrict__ sptr, const int32_t *__restrict__ sidx,
                                                      const double *__restrict__ ye, int64_t ys,
                                                      const uint8_t *__restrict__ fixed, double *__restrict__ y,
                                                      int64_t vs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    ye += blockIdx.y * ys;
    y += blockIdx.y * vs;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int r = sptr[k]; r < sptr[k + 1]; ++r) {
        const size_t o = 3 * (size_t)sidx[r];
        s0 += ye[o];
        s1 += ye[o + 1];
        s2 += ye[o + 2];
    }
    const size_t o = 3 * (size_t)nlist[k];
    y[o] = fixed[o] ? 0.0 : s0;
    y[o + 1] = fixed[o + 1] ? 0.0 : s1;
    y[o + 2] = fixed[o + 2] ? 0.0 : s2;
}

// symmetrise A_c; unit diagonal for dropped (all-zero) modes
__global__ __launch_bounds__(256) void k_coarse_fix(int nc, double *__restrict__ Ac) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nc * nc) return;
    const int i = (int)(idx / nc), j = (int)(idx - (int64_t)i * nc);
    if (j > i) return;
    if (i == j) {
        if (Ac[idx] == 0.0) Ac[idx] = 1.0;
        return;
    }
    const double s = 0.5 * (Ac[(size_t)i * nc + j] + Ac[(size_t)j * nc + i]);
    Ac[(size_t)i * nc + j] = s;
    Ac[(size_t)j * nc + i] = s;
}

// ---- A_c = Q' K Q element by element (the linear operator) ----
// Element e, listed as m, holds the distinct aggregates h_0 < ... < h_{k-1}
// of its nodes (glist[gptr[m] ..]).  Item (m, a), mode i (thread 6 item + i,
// so that the six threads of an item load the same element data in one
// request per wave):
// w = K_e Q_{g,i}|_e with g = h_a (the mode on e's nodes in g, fixed DOFs
// zero), then for every b the six numbers
//     c_j = Q_{h_b,j}|_e' w = sum_k Rinv_{h_b}[k][j] t_k,
//     t = sum over e's nodes in h_b of [w ; d x w]   (fixed DOFs zero),
// the contribution of e to A_c[(h_b, j), (g, i)].  They go to entry
// gpos[gq[m] + a k + b] of the pair (h_b, g), whose entries k_gal_sum adds
// in ascending element order: deterministic, no atomics.
// Elements whose nodes all lie in one aggregate and hold no fixed DOF are
// not listed: the mode is a rigid motion of the whole element, which K_e
// annihilates (zero strain), so their contribution is zero.
template <int NPE, bool AFF>
__global__ __launch_bounds__(256) void k_gal_elem(int n_items, const int32_t *__restrict__ item,
                                                 const int32_t *__restrict__ gel, const int32_t *__restrict__ gptr,
                                                 const int32_t *__restrict__ glist, const int32_t *__restrict__ gq,
                                                 const int32_t *__restrict__ gpos, int n_el,
                                                 const int32_t *__restrict__ conn, const double *__restrict__ gf,
                                                 const int32_t *__restrict__ node_agg, const double *__restrict__ cd,
                                                 const double *__restrict__ Rinv, const uint8_t *__restrict__ fixed,
                                                 double *__restrict__ C) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= 6 * n_items) return;
    const int k = tid / 6, i = tid - 6 * k;
    const int it = item[k], m = it >> 4, a = it & 15;
    const int e = gel[m], g0 = gptr[m], ke = gptr[m + 1] - g0;
    const int g = glist[g0 + a];
    double q[6];
#pragma unroll
    for (int kk = 0; kk < 6; ++kk) q[kk] = Rinv[36 * (size_t)g + kk * 6 + i];
    int nd[NPE], na[NPE];
    double w[NPE][3];
    mf_element_x<NPE, double, AFF>(
        n_el, e, gf,
        [&](int l, double (&v)[3]) {
            const int an = conn[(size_t)l * n_el + e];
            nd[l] = an;
            na[l] = node_agg[an];
            const double *d = cd + 3 * (size_t)an;
            const double o[3] = {q[0] + q[4] * d[2] - q[5] * d[1], q[1] + q[5] * d[0] - q[3] * d[2],
                                 q[2] + q[3] * d[1] - q[4] * d[0]};
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = (na[l] == g && !fixed[3 * (size_t)an + c]) ? o[c] : 0.0;
        },
        [&](int l, double f0, double f1, double f2) {
            w[l][0] = f0;
            w[l][1] = f1;
            w[l][2] = f2;
        });
#pragma unroll
    for (int l = 0; l < NPE; ++l)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (fixed[3 * (size_t)nd[l] + c]) w[l][c] = 0.0;
    for (int b = 0; b < ke; ++b) {
        const int h = glist[g0 + b];
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int l = 0; l < NPE; ++l) {
            if (na[l] != h) continue;
            const double *d = cd + 3 * (size_t)nd[l];
            const double y0 = w[l][0], y1 = w[l][1], y2 = w[l][2];
            t[0] += y0;
            t[1] += y1;
            t[2] += y2;
            t[3] += d[1] * y2 - d[2] * y1;
            t[4] += d[2] * y0 - d[0] * y2;
            t[5] += d[0] * y1 - d[1] * y0;
        }
        const double *R = Rinv + 36 * (size_t)h;
        double *o = C + 36 * (size_t)gpos[gq[m] + a * ke + b] + 6 * i;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int kk = 0; kk < 6; ++kk) s += R[kk * 6 + j] * t[kk];
            o[j] = s;
        }
    }
}

// One block per unordered aggregate pair {h, g} (h <= g), with its two
// ordered pairs' entry ranges.  Entry value l (i = l / 6, j = l % 6) of the
// pair (h, g) is A_c[(h, j), (g, i)]; the pair (g, h) holds the transposed
// value at 6 j + i.  Thread t < 1008 takes value l = t % 36 of every 28th
// entry from t / 36 on (eight loads in flight), the 28 group sums are added
// in group order, and the mean of the two orientations goes to both symmetric
// places, so A_c is symmetric by construction.  A zero diagonal (a mode
// dropped from its aggregate) becomes 1.
constexpr int kGalGroups = 28;
__device__ __forceinline__ double gal_range_sum(const double *__restrict__ C, int r0, int r1, int grp, int l) {
    double s = 0.0;
    int r = r0 + grp;
    for (; r + 7 * kGalGroups < r1; r += 8 * kGalGroups) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = C[36 * (size_t)(r + u * kGalGroups) + l];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; r < r1; r += kGalGroups) s += C[36 * (size_t)r + l];
    return s;
}

__global__ __launch_bounds__(1024) void k_gal_sum(const int32_t *__restrict__ upair, const int32_t *__restrict__ pptr,
                                                const double *__restrict__ C, double *__restrict__ Ac, int nc) {
    __shared__ double sh[2][kGalGroups][36];
    const int u = blockIdx.x, t = threadIdx.x;
    const int P = upair[4 * u], Pt = upair[4 * u + 1], h = upair[4 * u + 2], g = upair[4 * u + 3];
    if (t < 36 * kGalGroups) {
        const int grp = t / 36, l = t - 36 * grp;
        const int i = l / 6, j = l - 6 * (l / 6);
        sh[0][grp][l] = gal_range_sum(C, pptr[P], pptr[P + 1], grp, l);
        sh[1][grp][l] = gal_range_sum(C, pptr[Pt], pptr[Pt + 1], grp, 6 * j + i);
    }
    __syncthreads();
    if (t >= 36) return;
    const int i = t / 6, j = t - 6 * (t / 6);
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int grp = 0; grp < kGalGroups; ++grp) {
        s1 += sh[0][grp][t];
        s2 += sh[1][grp][t];
    }
    double v = 0.5 * (s1 + s2);
    if (h == g && i == j && v == 0.0) v = 1.0;
    Ac[(size_t)(6 * h + j) * nc + 6 * g + i] = v;
    Ac[(size_t)(6 * g + i) * nc + 6 * h + j] = v;
}

// ---- inverse of A_c by block Gauss-Jordan (in place, no pivoting) ----
// A_c is SPD, so every pivot block -- a Schur complement of A_c -- is SPD and
// needs no pivoting.  Step K (rows/columns k0..k0+bk-1), with P = A_KK:
//   A_KK <- P^-1,  A_Kj <- P^-1 A_Kj,  A_iK <- -A_iK P^-1,
//   A_ij <- A_ij - (A_iK P^-1) A_Kj          (i, j outside K)
// After the last step A holds A_c^-1.  A pivot of the small in-LDS sweep that
// is not positive and finite flags the matrix as not positive definite.
#ifndef JFEM_GJ_WIDTH
#define JFEM_GJ_WIDTH 32
#endif
constexpr int kGJ = JFEM_GJ_WIDTH;  // pivot block (32 or 64; 64 measured slower: setup 2.52 ms against 1.98 ms)
static_assert(kGJ == 32 || kGJ == 64, "pivot blocks of 32 or 64");
constexpr int kGJSub = kGJ / 16;    // look-ahead entries per thread and direction

// P^-1 of a bk x bk pivot block (padded to the identity), entry (i, j) read
// by ld(i, j), by scalar Gauss-Jordan over a block of 256 threads: thread t
// holds column j = t % kGJ of rows t / kGJ + RG u (RG = 256 / kGJ, u < kGJ / RG).  At step k the
// threads holding column k (the multipliers) and row k (the pivot row)
// write them to LDS (double-buffered by step parity: one barrier per step),
// and every thread updates its four entries.  The result goes to P.
// Returns false for a pivot that is not positive and finite (every thread
// sees the same pivots).  Called by the whole block.
template <typename Ld>
__device__ __forceinline__ bool gj_pivot_inverse(int bk, Ld &&ld, double (*P)[kGJ + 1]) {
    constexpr int RG = 256 / kGJ, R = kGJ / RG;  // row groups, rows per thread
    __shared__ double colk[2][kGJ], rowk[2][kGJ];
    const int t = threadIdx.x, j = t % kGJ, i0 = t / kGJ;
    double a[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int i = i0 + RG * u;
        a[u] = (i < bk && j < bk) ? ld(i, j) : (i == j ? 1.0 : 0.0);
    }
    bool ok = true;
#pragma unroll
    for (int k = 0; k < kGJ; ++k) {
        if (k < bk) {  // uniform
            double *ck = colk[k & 1], *rw = rowk[k & 1];
            const int uk = k / RG, ik = k - RG * (k / RG);  // row k is row group ik, entry uk
            if (j == k)
#pragma unroll
                for (int u = 0; u < R; ++u) ck[i0 + RG * u] = a[u];
            if (i0 == ik) rw[j] = a[uk];
            __syncthreads();
            // every LDS read of the step at once, then 1/p
            const double p = ck[k], rwj = rw[j];
            double c[R];
#pragma unroll
            for (int u = 0; u < R; ++u) c[u] = ck[i0 + RG * u];
            ok = ok && p > 0.0 && p < INFINITY;
            const double ip = 1.0 / p;
            const double r = rwj * ip;  // a_kj / p
            const bool piv = j == k;
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const double v = piv ? -c[u] * ip : a[u] - c[u] * r;
                a[u] = (u == uk && i0 == ik) ? (piv ? ip : r) : v;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < R; ++u) P[i0 + RG * u][j] = a[u];
    __syncthreads();
    return ok;
}

// The first step's pivot inverse (rows and columns 0 .. bk), one block.
__global__ __launch_bounds__(256) void k_gj_first(int n, int bk, const double *__restrict__ A,
                                                 double *__restrict__ Pout, int *__restrict__ flag) {
    __shared__ double P[kGJ][kGJ + 1];
    const bool ok = gj_pivot_inverse(bk, [&](int i, int j) { return A[(size_t)i * n + j]; }, P);
    for (int t = threadIdx.x; t < kGJ * kGJ; t += blockDim.x) Pout[t] = P[t / kGJ][t - kGJ * (t / kGJ)];
    if (threadIdx.x == 0 && !ok) flag[0] = 1;
}

// One block Gauss-Jordan step K (pivot rows and columns k0 .. k0 + bk) from
// Ao into An, one 64 x 64 tile on or below the diagonal per block (diagonal
// tiles whole), 4 x 4 entries per thread:
//   A_ij -= C_i R_j  (i, j not in K),   A_iK = -C_i,   A_Kj = P^-1 R_j,
//   A_KK = P^-1,
// with P^-1 the pivot block's inverse (Pbuf, formed by the previous step),
// C = A_:K P^-1 (every tile forms its own 64 rows of it) and R = row K.  An
// is a second copy of the matrix, so no tile overwrites what another still
// reads, and the step needs no panel kernel before it.  The sweep keeps
// M_ij = +-M_ji' (minus when exactly one of i, j is pivoted), so only the
// tiles on and below the diagonal are stored and updated: A_:K of a row
// above the pivot's tile is read as -A_K:', and row K right of the pivot's
// tile as A_:K'.  With nbk > 0, block 0 instead forms the next pivot block
// after this update, with the same arithmetic as the tile that holds it (so
// the same values), and its inverse (Pnext), beside the tiles.
__global__ __launch_bounds__(256) void k_gj_step(int n, int k0, int bk, const double *__restrict__ Ao,
                                                double *__restrict__ An, const double *__restrict__ Pbuf, int nk0,
                                                int nbk, double *__restrict__ Pnext, int *__restrict__ flag) {
    __shared__ double Cs[64][kGJ + 1], Rs[kGJ][65], Ps[kGJ][kGJ + 1], As[64][kGJ + 1];
    const int TK = k0 / 64;
    const int jmax = min(n, 64 * (TK + 1));  // row K is stored as such up to the end of the pivot's tile
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    // the panel's 64 rows into Cs, from As staged by the caller:
    // Cs[a][m] = sum_l As[a][l] P[l][m] (As holds -A_K:' above the pivot's
    // tile, so the sum there is c -= A P with A as stored), rows ty + 16 u
    // and columns tx + 16 v per thread
    auto panel = [&]() {
        constexpr int V = kGJ / 16;
        double c[4][V];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < V; ++v) c[u][v] = 0.0;
        for (int l = 0; l < bk; ++l) {
            double av[4], pv[V];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[ty + 16 * u][l];
#pragma unroll
            for (int v = 0; v < V; ++v) pv[v] = Ps[l][tx + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < V; ++v) c[u][v] += av[u] * pv[v];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < V; ++v) Cs[ty + 16 * u][tx + 16 * v] = c[u][v];
    };
    // As rows a of rows r0 + a (nr of them): A_:K as stored at or below the
    // pivot's tile (consecutive threads along a row of A), or -A_K:' above
    // it (consecutive threads along a row of A_K:)
    auto stage_col = [&](int r0, int nr) {
        if (r0 / 64 >= TK) {
#pragma unroll
            for (int u = 0; u < 64 * kGJ / 256; ++u) {
                const int t = (int)threadIdx.x + 256 * u;
                const int a = t / kGJ, m = t - kGJ * (t / kGJ);
                As[a][m] = (a < nr && m < bk) ? Ao[(size_t)(r0 + a) * n + k0 + m] : 0.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < 64 * kGJ / 256; ++u) {
                const int t = (int)threadIdx.x + 256 * u;
                const int m = t / 64, a = t - 64 * (t / 64);
                As[a][m] = (a < nr && m < bk) ? -Ao[(size_t)(k0 + m) * n + r0 + a] : 0.0;
            }
        }
    };
    // row K entry (k0 + m, j): as stored up to the end of the pivot's tile,
    // else as the transposed entry of column K
    auto rowK = [&](int m, int j) -> double {
        return j < jmax ? Ao[(size_t)(k0 + m) * n + j] : Ao[(size_t)j * n + k0 + m];
    };
    if (nbk > 0 && blockIdx.x == 0) {
        // the next pivot block: rows and columns nk0 .. nk0 + nbk, below the
        // pivot rows, so A_:K is read as stored
        double av[kGJSub][kGJSub];
#pragma unroll
        for (int a = 0; a < kGJSub; ++a)
#pragma unroll
            for (int b = 0; b < kGJSub; ++b) {
                const int i = ty + 16 * a, j = tx + 16 * b;
                av[a][b] = (i < nbk && j < nbk) ? Ao[(size_t)(nk0 + i) * n + nk0 + j] : 0.0;
            }
        stage_col(nk0, nbk);
        for (int t = threadIdx.x; t < kGJ * kGJ; t += blockDim.x) {
            const int a = t / kGJ, m = t - kGJ * (t / kGJ);
            Rs[a][m] = (a < bk && m < nbk) ? rowK(a, nk0 + m) : 0.0;
            Ps[a][m] = Pbuf[t];
        }
        __syncthreads();
        panel();
        __syncthreads();
        double acc[kGJSub][kGJSub];
#pragma unroll
        for (int a = 0; a < kGJSub; ++a)
#pragma unroll
            for (int b = 0; b < kGJSub; ++b) acc[a][b] = 0.0;
        for (int m = 0; m < bk; ++m) {
            double cv[kGJSub], rv[kGJSub];
#pragma unroll
            for (int a = 0; a < kGJSub; ++a) cv[a] = Cs[ty + 16 * a][m];
#pragma unroll
            for (int b = 0; b < kGJSub; ++b) rv[b] = Rs[m][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < kGJSub; ++a)
#pragma unroll
                for (int b = 0; b < kGJSub; ++b) acc[a][b] += cv[a] * rv[b];
        }
        __syncthreads();  // Ps is overwritten below
#pragma unroll
        for (int a = 0; a < kGJSub; ++a)
#pragma unroll
            for (int b = 0; b < kGJSub; ++b) Ps[ty + 16 * a][tx + 16 * b] = av[a][b] - acc[a][b];
        __syncthreads();
        const bool ok = gj_pivot_inverse(nbk, [&](int i, int j) { return Ps[i][j]; }, As);
        for (int t = threadIdx.x; t < kGJ * kGJ; t += blockDim.x) Pnext[t] = As[t / kGJ][t - kGJ * (t / kGJ)];
        if (threadIdx.x == 0 && !ok) flag[0] = 1;
        return;
    }
    // tiles on and below the diagonal: bid = ti (ti + 1) / 2 + tj, tj <= ti
    const int bid = (int)blockIdx.x - (nbk > 0 ? 1 : 0);
    int ti = (int)((sqrtf(8.f * (float)bid + 1.f) - 1.f) * 0.5f);
    while ((ti + 1) * (ti + 2) / 2 <= bid) ++ti;
    while (ti * (ti + 1) / 2 > bid) --ti;
    const int tj = bid - ti * (ti + 1) / 2;
    const int i0 = 64 * ti, j0 = 64 * tj;
    double av[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            av[a][b] = (i < n && j < n) ? Ao[(size_t)i * n + j] : 0.0;
        }
    stage_col(i0, min(64, n - i0));
#pragma unroll
    for (int u = 0; u < 64 * kGJ / 256; ++u) {
        const int t = (int)threadIdx.x + 256 * u;
        const int mm = t / 64, b = t - 64 * (t / 64);
        Rs[mm][b] = (mm < bk && j0 + b < n) ? rowK(mm, j0 + b) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kGJ * kGJ / 256; ++u) {
        const int t = (int)threadIdx.x + 256 * u;
        Ps[t / kGJ][t - kGJ * (t / kGJ)] = Pbuf[t];
    }
    __syncthreads();
    panel();
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int m = 0; m < bk; ++m) {
        double cv[4], rv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) cv[a] = Cs[ty + 16 * a][m];
#pragma unroll
        for (int b = 0; b < 4; ++b) rv[b] = Rs[m][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += cv[a] * rv[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + ty + 16 * a;
        if (i >= n) continue;
        const bool inI = i >= k0 && i < k0 + bk;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + tx + 16 * b;
            if (j >= n) continue;
            const bool inJ = j >= k0 && j < k0 + bk;
            double *o = An + (size_t)i * n + j;
            if (!inI && !inJ) {
                *o = av[a][b] - acc[a][b];
            } else if (!inI) {
                *o = -Cs[ty + 16 * a][j - k0];
            } else if (!inJ) {
                double v = 0.0;
                for (int l = 0; l < bk; ++l) v += Ps[i - k0][l] * Rs[l][tx + 16 * b];
                *o = v;
            } else {
                *o = Ps[i - k0][j - k0];
            }
        }
    }
}

__global__ void k_flag_clear(int *flag) { flag[0] = 0; }

// Ahat = R^-1 A^-1 R^-T from the lower triangle (row-major) of the inverse:
// entry (r, s) of A^-1 with r >= s is read at L[r * nc + s]; rows of Ahat
// have stride ld and columns nc..ld-1 are zero.  One thread per 6 x 6 block
// (g, h) of aggregate pair reads the block of A^-1 once, forms T = A^-1_gh
// R_h^-T, then R_g^-1 T (a dropped mode's row and column are zero); the
// threads of the last column block also zero the padding columns of their
// rows.  (One thread per entry, reading its 6 x 6 block each time, took
// 79 us at nc = 1,512 and 0.05 ms more per setup.)
__global__ __launch_bounds__(256) void k_coarse_fold_blocks(int n_agg, int nc, int ld, const double *__restrict__ L,
                                                           const double *__restrict__ Rinv,
                                                           const int32_t *__restrict__ dropped,
                                                           float *__restrict__ A32) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n_agg * n_agg) return;
    const int g = (int)(idx / n_agg), h = (int)(idx - (int64_t)g * n_agg);
    const double *Rg = Rinv + 36 * (size_t)g, *Rh = Rinv + 36 * (size_t)h;
    double a[6][6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int l = 0; l < 6; ++l) {
            const int rr = 6 * g + k, ss = 6 * h + l;
            const double v = rr >= ss ? L[(size_t)rr * nc + ss] : L[(size_t)ss * nc + rr];
            a[k][l] = (dropped[rr] || dropped[ss]) ? 0.0 : v;
        }
    double t[6][6];  // t[k][j] = sum_l a[k][l] Rh[j][l]
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double inner = 0.0;
#pragma unroll
            for (int l = 0; l < 6; ++l) inner += a[k][l] * Rh[j * 6 + l];
            t[k][j] = inner;
        }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        float *row = A32 + (size_t)(6 * g + i) * ld + 6 * h;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double rik = Rg[i * 6 + k];
                if (rik != 0.0) acc += rik * t[k][j];
            }
            row[j] = (float)acc;
        }
        if (h == n_agg - 1)
            for (int c = nc; c < ld; ++c) A32[(size_t)(6 * g + i) * ld + c] = 0.f;
    }
}

// w_g = sum_{a in g} [m_a r_a ; d_a x (m_a r_a)], one 1024-thread block per
// aggregate (about two nodes per thread); cdA is cd in agg_nodes order so its
// reads are contiguous.  Written in FP32 (padded row of the gemv) and FP64
// (for w.zc).
// With AGG, r is already m o r in aggregate order (rA, position k), so the
// reads are contiguous and need no node list or mask.
template <typename T, bool AGG>
__global__ __launch_bounds__(1024) void k_coarse_restrict(int n_agg, const int32_t *__restrict__ agg_ptr,
                                                         const int32_t *__restrict__ agg_nodes, const T *__restrict__ cdA,
                                                         const uint8_t *__restrict__ fixed, const T *__restrict__ r,
                                                         float *__restrict__ wc, double *__restrict__ wc64,
                                                         const int *__restrict__ done) {
    // the reads below are harmless past convergence, so the flag is tested
    // after them and its latency overlaps theirs
    const int dn = done ? *done : 0;
    const int g = blockIdx.x;
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = agg_ptr[g] + threadIdx.x; k < agg_ptr[g + 1]; k += blockDim.x) {
        double y0, y1, y2;
        if (AGG) {
            y0 = (double)r[3 * (size_t)k];
            y1 = (double)r[3 * (size_t)k + 1];
            y2 = (double)r[3 * (size_t)k + 2];
        } else {
            const size_t o = 3 * (size_t)agg_nodes[k];
            y0 = fixed[o] ? 0.0 : (double)r[o];
            y1 = fixed[o + 1] ? 0.0 : (double)r[o + 1];
            y2 = fixed[o + 2] ? 0.0 : (double)r[o + 2];
        }
        const double d0 = cdA[3 * (size_t)k], d1 = cdA[3 * (size_t)k + 1], d2 = cdA[3 * (size_t)k + 2];
        w[0] += y0; w[1] += y1; w[2] += y2;
        w[3] += d1 * y2 - d2 * y1;
        w[4] += d2 * y0 - d0 * y2;
        w[5] += d0 * y1 - d1 * y0;
    }
    if (dn) return;  // uniform
    block_sum_all<6>(w);
    if (threadIdx.x < 6) {
        wc[6 * g + threadIdx.x] = (float)w[threadIdx.x];
        wc64[6 * g + threadIdx.x] = w[threadIdx.x];
    }
}

// zc = Ahat w, one 64-lane wave per row (grid-stride over rows), float4 loads
// of the zero-padded rows, FP64 lane sums added in a fixed order.  Ahat is
// kept in FP32 for both precisions: it only scales the coarse correction of a
// fixed SPD preconditioner, so its rounding changes the iteration count at
// most, never the answer, and it halves the dominant read of the level.
// Block partials of w.zc (w in FP64) go to `part`.
__global__ __launch_bounds__(256) void k_coarse_gemv(int nc, int ld4, const float4 *__restrict__ A,
                                                    const float4 *__restrict__ w, const double *__restrict__ w64,
                                                    float *__restrict__ zc, double *__restrict__ part,
                                                    const int *__restrict__ done) {
    const int dn = done ? *done : 0;  // tested once the first row is in flight
    const int lane = threadIdx.x & 63;
    double v[1] = {0.0};
    for (int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; row < nc; row += gridDim.x * (blockDim.x / 64)) {
        const float4 *Ar = A + (size_t)row * ld4;
        double s = 0.0;
#pragma unroll 8
        for (int k = lane; k < ld4; k += 64) {
            const float4 a = Ar[k], b = w[k];
            s += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
        }
        if (dn) return;  // uniform
        const double sd = wave_sum(s);
        const float z = (float)sd;
        if (lane == 0) {
            zc[row] = z;
            v[0] += w64[row] * (double)z;
        }
    }
    if (dn) return;  // every thread leaves before the block reduction
    write_partials<1>(v, part);
}


// z_a += m_a o (t_g + omega_g x d_a); optionally p = z; partials (r.z, r.r)
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_prolong(int N, const int32_t *__restrict__ node_agg, const T *__restrict__ cd,
                                                       const uint8_t *__restrict__ fixed, const float *__restrict__ zc,
                                                       const T *__restrict__ r, T *__restrict__ z, T *__restrict__ p,
                                                       double *__restrict__ part, const int *__restrict__ done) {
    if (done && *done) return;
    double v[2] = {0.0, 0.0};
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += gridDim.x * blockDim.x) {
        const size_t o = 3 * (size_t)a;
        const float *yf = zc + 6 * (size_t)node_agg[a];
        const T y[6] = {(T)yf[0], (T)yf[1], (T)yf[2], (T)yf[3], (T)yf[4], (T)yf[5]};
        const T d0 = cd[o], d1 = cd[o + 1], d2 = cd[o + 2];
        const T c[3] = {y[0] + y[4] * d2 - y[5] * d1, y[1] + y[5] * d0 - y[3] * d2, y[2] + y[3] * d1 - y[4] * d0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T rv = r[o + k];
            const T zv = z[o + k] + (fixed[o + k] ? T(0) : c[k]);
            z[o + k] = zv;
            if (p) p[o + k] = zv;
            v[0] += (double)rv * (double)zv;
            v[1] += (double)rv * (double)rv;
        }
    }
    write_partials<2>(v, part);
}

// ------------------------------------------------------------ host setup ---
namespace {

struct Aggregation {
    int n_agg = 0;
    std::vector<int32_t> node_agg, agg_ptr, agg_nodes, color, partner;  // partner: ncolor x n_agg
    int ncolor = 0;
    std::vector<double> cd, Rinv;   // N x 3, n_agg x 36
    std::vector<int32_t> dropped;   // nc
    // per colour: the elements touching its aggregates and those elements' nodes
    std::vector<int32_t> ecol_ptr, ecol, ncol_ptr, ncol;
    // a colour's nodes are listed by aggregate: segment s (colours' segments
    // at cseg[c]..cseg[c+1]) is ncol[segb[s]..segb[s+1]) of aggregate sega[s]
    std::vector<int32_t> cseg, segb, sega;
    // per listed node (ncol order): its incidence slots in the colour's
    // elements, ascending element id: csub[csub_ptr[k]..csub_ptr[k+1])
    std::vector<int32_t> csub_ptr, csub;
    // element-level Galerkin lists (k_gal_elem, k_gal_sum): listed elements,
    // their aggregates, entry positions, items (m << 4 | a), entry ranges of
    // the ordered aggregate pairs, unordered pairs (P, P', h, g)
    std::vector<int32_t> gal_el, gal_gptr, gal_glist, gal_gq, gal_gpos, gal_item, gal_pptr, gal_upair;
};

// grid binning of the coordinates, about `target` nodes per bin
void bin_nodes(const std::vector<double> &X, int64_t N, int target, std::vector<int32_t> &agg, int &n_agg) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t a = 0; a < N; ++a)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], X[3 * a + c]);
            hi[c] = std::max(hi[c], X[3 * a + c]);
        }
    double ext[3], vol = 1.0, maxe = 0.0;
    for (int c = 0; c < 3; ++c) maxe = std::max(maxe, hi[c] - lo[c]);
    for (int c = 0; c < 3; ++c) {
        ext[c] = std::max(hi[c] - lo[c], 1e-12 * std::max(maxe, 1e-300));
        vol *= ext[c];
    }
    const double nbins = std::max(1.0, (double)N / std::max(target, 1));
    const double cell = std::cbrt(vol / nbins);
    int nb[3];
    for (int c = 0; c < 3; ++c) nb[c] = std::max(1, (int)std::lround(ext[c] / cell));
    std::vector<int64_t> key(N);
    for (int64_t a = 0; a < N; ++a) {
        int b[3];
        for (int c = 0; c < 3; ++c) {
            b[c] = (int)std::floor((X[3 * a + c] - lo[c]) / ext[c] * nb[c]);
            b[c] = std::min(std::max(b[c], 0), nb[c] - 1);
        }
        key[a] = ((int64_t)b[0] * nb[1] + b[1]) * nb[2] + b[2];
    }
    std::vector<int64_t> uk(key);
    std::sort(uk.begin(), uk.end());
    uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
    agg.resize(N);
    for (int64_t a = 0; a < N; ++a) agg[a] = (int32_t)(std::lower_bound(uk.begin(), uk.end(), key[a]) - uk.begin());
    n_agg = (int)uk.size();
}

std::vector<std::vector<int32_t>> agg_adjacency(const std::vector<int32_t> &conn, int64_t n_el, int npe,
                                                const std::vector<int32_t> &agg, int n_agg) {
    std::vector<std::vector<int32_t>> adj(n_agg);
    int32_t loc[32];
    for (int64_t e = 0; e < n_el; ++e) {
        int m = 0;
        for (int i = 0; i < npe; ++i) {
            const int32_t g = agg[conn[e * npe + i]];
            bool seen = false;
            for (int k = 0; k < m; ++k) seen = seen || loc[k] == g;
            if (!seen) loc[m++] = g;
        }
        for (int k = 0; k < m; ++k)
            for (int l = 0; l < m; ++l)
                if (k != l) adj[loc[k]].push_back(loc[l]);
    }
    for (auto &v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    return adj;
}

// The lists of the element-level Galerkin product.  Ordered pair (h, g) is
// pair_base[h] + (position of g in nbr[h]), nbr[h] = {h} and h's neighbours,
// ascending.  Each pair's entries are numbered in ascending element order.
void galerkin_lists(const jfm_ctx *ctx, const std::vector<int32_t> &agg, int n_agg,
                    const std::vector<std::vector<int32_t>> &adj, const std::vector<uint8_t> &fixed, Aggregation &A) {
    const int64_t n_el = ctx->n_el;
    const int npe = ctx->npe;
    std::vector<std::vector<int32_t>> nbr(n_agg);
    std::vector<int64_t> pair_base(n_agg + 1, 0);
    for (int h = 0; h < n_agg; ++h) {
        nbr[h] = adj[h];
        nbr[h].insert(std::lower_bound(nbr[h].begin(), nbr[h].end(), h), h);
        pair_base[h + 1] = pair_base[h] + (int64_t)nbr[h].size();
    }
    auto pair_of = [&](int h, int g) -> int64_t {
        return pair_base[h] + (std::lower_bound(nbr[h].begin(), nbr[h].end(), g) - nbr[h].begin());
    };
    const int64_t n_pairs = pair_base[n_agg];
    std::vector<int32_t> cnt(n_pairs + 1, 0);
    A.gal_gptr.assign(1, 0);
    A.gal_gq.clear();
    int32_t loc[32];
    int64_t q2 = 0;
    for (int64_t e = 0; e < n_el; ++e) {
        int k = 0;
        bool has_fixed = false;
        for (int l = 0; l < npe; ++l) {
            const int32_t a = ctx->h_conn[e * npe + l];
            has_fixed = has_fixed || fixed[3 * (size_t)a] || fixed[3 * (size_t)a + 1] || fixed[3 * (size_t)a + 2];
            const int32_t g = agg[a];
            bool seen = false;
            for (int t = 0; t < k; ++t) seen = seen || loc[t] == g;
            if (!seen) loc[k++] = g;
        }
        if (k < 2 && !has_fixed) continue;  // a rigid motion of the element: no contribution
        std::sort(loc, loc + k);
        A.gal_el.push_back((int32_t)e);
        A.gal_gq.push_back((int32_t)q2);
        q2 += (int64_t)k * k;
        for (int t = 0; t < k; ++t) A.gal_glist.push_back(loc[t]);
        A.gal_gptr.push_back((int32_t)A.gal_glist.size());
        for (int a = 0; a < k; ++a)
            for (int b = 0; b < k; ++b) ++cnt[pair_of(loc[b], loc[a]) + 1];
    }
    for (int64_t P = 0; P < n_pairs; ++P) cnt[P + 1] += cnt[P];
    A.gal_pptr.assign(cnt.begin(), cnt.end());
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    A.gal_gpos.resize(q2);
    A.gal_item.clear();
    const int64_t M = (int64_t)A.gal_el.size();
    for (int64_t m = 0; m < M; ++m) {
        const int g0 = A.gal_gptr[m], k = A.gal_gptr[m + 1] - g0;
        for (int a = 0; a < k; ++a) {
            A.gal_item.push_back((int32_t)(m * 16 + a));
            for (int b = 0; b < k; ++b)
                A.gal_gpos[A.gal_gq[m] + a * k + b] = fill[pair_of(A.gal_glist[g0 + b], A.gal_glist[g0 + a])]++;
        }
    }
    A.gal_upair.clear();
    for (int h = 0; h < n_agg; ++h)
        for (int32_t g : nbr[h])
            if (g >= h) {
                A.gal_upair.push_back((int32_t)pair_of(h, g));
                A.gal_upair.push_back((int32_t)pair_of(g, h));
                A.gal_upair.push_back(h);
                A.gal_upair.push_back(g);
            }
}

}  // namespace

// Host-side aggregation, colouring and per-aggregate mode bases.
static int coarse_host_setup(jfm_ctx *ctx, int coarse_nodes, const std::vector<uint8_t> &fixed,
                             const std::vector<int32_t> &slot, Aggregation &A) {
    const int64_t N = ctx->N;
    const std::vector<double> &X = ctx->h_X;
    int n_agg = 0;
    std::vector<int32_t> agg;
    bin_nodes(X, N, coarse_nodes, agg, n_agg);
    // merge aggregates far below the target size into their largest neighbour
    {
        auto adj = agg_adjacency(ctx->h_conn, ctx->n_el, ctx->npe, agg, n_agg);
        std::vector<int64_t> cnt(n_agg, 0);
        for (int64_t a = 0; a < N; ++a) cnt[agg[a]]++;
        const int64_t small = std::max<int64_t>(8, coarse_nodes / 8);
        std::vector<int32_t> into(n_agg);
        std::iota(into.begin(), into.end(), 0);
        for (int g = 0; g < n_agg; ++g) {
            if (cnt[g] >= small) continue;
            int best = -1;
            for (int32_t h : adj[g])
                if (cnt[h] >= small && (best < 0 || cnt[h] > cnt[best])) best = h;
            if (best >= 0) into[g] = best;
        }
        std::vector<int32_t> remap(n_agg, -1);
        int m = 0;
        for (int g = 0; g < n_agg; ++g)
            if (into[g] == g) remap[g] = m++;
        for (int64_t a = 0; a < N; ++a) agg[a] = remap[into[agg[a]]];
        n_agg = m;
    }
    A.n_agg = n_agg;
    A.node_agg = agg;
    A.agg_ptr.assign(n_agg + 1, 0);
    for (int64_t a = 0; a < N; ++a) A.agg_ptr[agg[a] + 1]++;
    for (int g = 0; g < n_agg; ++g) A.agg_ptr[g + 1] += A.agg_ptr[g];
    A.agg_nodes.resize(N);
    {
        std::vector<int32_t> pos(A.agg_ptr.begin(), A.agg_ptr.end() - 1);
        for (int64_t a = 0; a < N; ++a) A.agg_nodes[pos[agg[a]]++] = (int32_t)a;  // ascending node id
    }
    // centroids and offsets
    std::vector<double> cent(3 * (size_t)n_agg, 0.0);
    for (int64_t a = 0; a < N; ++a)
        for (int c = 0; c < 3; ++c) cent[3 * agg[a] + c] += X[3 * a + c];
    for (int g = 0; g < n_agg; ++g) {
        const double k = (double)(A.agg_ptr[g + 1] - A.agg_ptr[g]);
        for (int c = 0; c < 3; ++c) cent[3 * g + c] /= k;
    }
    A.cd.resize(3 * (size_t)N);
    for (int64_t a = 0; a < N; ++a)
        for (int c = 0; c < 3; ++c) A.cd[3 * a + c] = X[3 * a + c] - cent[3 * agg[a] + c];
    // Gram matrix of the masked modes per aggregate -> Cholesky -> R^-1
    A.Rinv.assign(36 * (size_t)n_agg, 0.0);
    A.dropped.assign(6 * (size_t)n_agg, 0);
    for (int g = 0; g < n_agg; ++g) {
        double G[36] = {0};
        for (int k = A.agg_ptr[g]; k < A.agg_ptr[g + 1]; ++k) {
            const int64_t a = A.agg_nodes[k];
            const double *d = &A.cd[3 * a];
            // rows of the mode matrix for the three DOFs of node a
            double Bm[3][6] = {{1, 0, 0, 0, d[2], -d[1]}, {0, 1, 0, -d[2], 0, d[0]}, {0, 0, 1, d[1], -d[0], 0}};
            for (int c = 0; c < 3; ++c) {
                if (fixed[3 * a + c]) continue;
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 6; ++j) G[i * 6 + j] += Bm[c][i] * Bm[c][j];
            }
        }
        // Cholesky G = R'R with column dropping
        double R[36] = {0};
        double gmax = 0.0;
        for (int i = 0; i < 6; ++i) gmax = std::max(gmax, G[i * 6 + i]);
        bool keep[6];
        for (int i = 0; i < 6; ++i) {
            double s = G[i * 6 + i];
            for (int k = 0; k < i; ++k) s -= R[k * 6 + i] * R[k * 6 + i];
            keep[i] = gmax > 0.0 && s > 1e-10 * gmax;
            if (!keep[i]) continue;
            R[i * 6 + i] = std::sqrt(s);
            for (int j = i + 1; j < 6; ++j) {
                double t = G[i * 6 + j];
                for (int k = 0; k < i; ++k) t -= R[k * 6 + i] * R[k * 6 + j];
                R[i * 6 + j] = t / R[i * 6 + i];
            }
        }
        // R^-1 (upper triangular, dropped rows/columns zero)
        double *Ri = &A.Rinv[36 * (size_t)g];
        for (int j = 5; j >= 0; --j) {
            if (!keep[j]) continue;
            Ri[j * 6 + j] = 1.0 / R[j * 6 + j];
            for (int i = j - 1; i >= 0; --i) {
                if (!keep[i]) continue;
                double s = 0.0;
                for (int k = i + 1; k <= j; ++k) s += R[i * 6 + k] * Ri[k * 6 + j];
                Ri[i * 6 + j] = -s / R[i * 6 + i];
            }
        }
        for (int i = 0; i < 6; ++i) A.dropped[6 * g + i] = keep[i] ? 0 : 1;
    }
    // distance-2 greedy colouring of the aggregate graph
    auto adj = agg_adjacency(ctx->h_conn, ctx->n_el, ctx->npe, agg, n_agg);
    A.color.assign(n_agg, -1);
    int ncolor = 0;
    std::vector<int> mark;
    for (int g = 0; g < n_agg; ++g) {
        mark.assign(ncolor + 1, 0);
        auto forbid = [&](int h) {
            if (A.color[h] >= 0) mark[A.color[h]] = 1;
        };
        for (int32_t h : adj[g]) {
            forbid(h);
            for (int32_t k : adj[h])
                if (k != g) forbid(k);
        }
        int c = 0;
        while (c < ncolor && mark[c]) ++c;
        A.color[g] = c;
        ncolor = std::max(ncolor, c + 1);
    }
    A.ncolor = ncolor;
    galerkin_lists(ctx, agg, n_agg, adj, fixed, A);
    A.partner.assign((size_t)ncolor * n_agg, -1);
    for (int g = 0; g < n_agg; ++g) {
        A.partner[(size_t)A.color[g] * n_agg + g] = g;
        for (int32_t h : adj[g]) {
            int32_t &slot = A.partner[(size_t)A.color[h] * n_agg + g];
            if (slot >= 0 && slot != h) return jfm_fail(ctx, JFM_EINVAL, "aggregate colouring is not distance-2");
            slot = h;
        }
    }
    // probing touches, for colour c, only the elements with a node in a
    // colour-c aggregate and those elements' nodes
    const int64_t n_el = ctx->n_el;
    const int npe = ctx->npe;
    std::vector<std::vector<int32_t>> el(ncolor), nd(ncolor);
    std::vector<int32_t> mark_e(ncolor, -1);
    for (int64_t e = 0; e < n_el; ++e) {
        for (int i = 0; i < npe; ++i) {
            const int c = A.color[agg[ctx->h_conn[e * npe + i]]];
            if (mark_e[c] != (int32_t)e) {
                mark_e[c] = (int32_t)e;
                el[c].push_back((int32_t)e);
            }
        }
    }
    std::vector<int32_t> seen(N, -1);
    A.ecol_ptr.assign(ncolor + 1, 0);
    A.ncol_ptr.assign(ncolor + 1, 0);
    for (int c = 0; c < ncolor; ++c) {
        for (int32_t e : el[c])
            for (int i = 0; i < npe; ++i) {
                const int32_t a = ctx->h_conn[(int64_t)e * npe + i];
                if (seen[a] != c) {
                    seen[a] = c;
                    nd[c].push_back(a);
                }
            }
        std::sort(nd[c].begin(), nd[c].end(), [&](int32_t u, int32_t v) {
            return agg[u] != agg[v] ? agg[u] < agg[v] : u < v;
        });
        A.ecol_ptr[c + 1] = A.ecol_ptr[c] + (int32_t)el[c].size();
        A.ncol_ptr[c + 1] = A.ncol_ptr[c] + (int32_t)nd[c].size();
    }
    A.ecol.reserve(A.ecol_ptr[ncolor]);
    A.ncol.reserve(A.ncol_ptr[ncolor]);
    A.cseg.assign(1, 0);
    for (int c = 0; c < ncolor; ++c) {
        A.ecol.insert(A.ecol.end(), el[c].begin(), el[c].end());
        for (size_t k = 0; k < nd[c].size(); ++k) {
            if (k == 0 || agg[nd[c][k]] != agg[nd[c][k - 1]]) {
                A.segb.push_back((int32_t)A.ncol.size());
                A.sega.push_back(agg[nd[c][k]]);
            }
            A.ncol.push_back(nd[c][k]);
        }
        A.cseg.push_back((int32_t)A.sega.size());
    }
    A.segb.push_back((int32_t)A.ncol.size());
    // per colour and listed node: its incidence slots in the colour's
    // elements, ascending element id (slot[i * n_el + e] is incidence (e, i))
    A.csub_ptr.assign(1, 0);
    std::vector<int32_t> posn(N, -1), cnt, fill;
    for (int c = 0; c < ncolor; ++c) {
        const int32_t k0 = A.ncol_ptr[c], k1 = A.ncol_ptr[c + 1];
        for (int32_t k = k0; k < k1; ++k) posn[A.ncol[k]] = k - k0;
        cnt.assign(k1 - k0 + 1, 0);
        for (int32_t e : el[c])
            for (int i = 0; i < npe; ++i) ++cnt[posn[ctx->h_conn[(int64_t)e * npe + i]] + 1];
        for (int32_t k = 0; k < k1 - k0; ++k) cnt[k + 1] += cnt[k];
        const size_t base = A.csub.size();
        A.csub.resize(base + cnt[k1 - k0]);
        fill.assign(cnt.begin(), cnt.end() - 1);
        for (int32_t e : el[c])  // ascending element id
            for (int i = 0; i < npe; ++i) {
                const int32_t a = ctx->h_conn[(int64_t)e * npe + i];
                A.csub[base + fill[posn[a]]++] = slot[(size_t)i * n_el + e];
            }
        for (int32_t k = 0; k < k1 - k0; ++k) A.csub_ptr.push_back((int32_t)(base + cnt[k + 1]));
    }
    return JFM_OK;
}

template <typename T> static void free_dev(jfm_ctx *ctx, T *&p) {
    if (p) jfm_free(ctx, p);
    p = nullptr;
}

void jfem_coarse_release(jfm_ctx *ctx) {
    free_dev(ctx, ctx->node_agg); free_dev(ctx, ctx->agg_ptr); free_dev(ctx, ctx->agg_nodes);
    free_dev(ctx, ctx->agg_color); free_dev(ctx, ctx->partner); free_dev(ctx, ctx->cd64); free_dev(ctx, ctx->cd32);
    free_dev(ctx, ctx->Rinv); free_dev(ctx, ctx->Ac); free_dev(ctx, ctx->Ac2); free_dev(ctx, ctx->Ahat32);
    free_dev(ctx, ctx->wc32); free_dev(ctx, ctx->zc32); free_dev(ctx, ctx->wc64);
    free_dev(ctx, ctx->cdA64); free_dev(ctx, ctx->cdA32); free_dev(ctx, ctx->apos); free_dev(ctx, ctx->rA);
    free_dev(ctx, ctx->gjP);
    free_dev(ctx, ctx->pv6); free_dev(ctx, ctx->qv6); free_dev(ctx, ctx->ye6);
    free_dev(ctx, ctx->gal_el); free_dev(ctx, ctx->gal_gptr); free_dev(ctx, ctx->gal_glist);
    free_dev(ctx, ctx->gal_gq); free_dev(ctx, ctx->gal_gpos); free_dev(ctx, ctx->gal_item);
    free_dev(ctx, ctx->gal_pptr); free_dev(ctx, ctx->gal_upair); free_dev(ctx, ctx->gal_C);
    ctx->gal_nitems = ctx->gal_nupair = 0;
    free_dev(ctx, ctx->dropped); free_dev(ctx, ctx->ecol); free_dev(ctx, ctx->ncol);
    ctx->h_ecol_ptr.clear(); ctx->h_ncol_ptr.clear(); ctx->h_cseg.clear();
    free_dev(ctx, ctx->segb); free_dev(ctx, ctx->sega); free_dev(ctx, ctx->csub_ptr); free_dev(ctx, ctx->csub);
    for (int k = 0; k < 2; ++k)
        if (ctx->coarse_graph[k]) {
            (void)hipGraphExecDestroy(ctx->coarse_graph[k]);
            ctx->coarse_graph[k] = nullptr;
        }
    if (ctx->coarse_inv_graph) {
        (void)hipGraphExecDestroy(ctx->coarse_inv_graph);
        ctx->coarse_inv_graph = nullptr;
    }
    ctx->coarse_ready = false;
    ctx->coarse_tag = -1;
    ctx->coarse_built_nodes = 0;
    ctx->n_agg = ctx->nc = 0;
}

// Zero A_c, then one batched product of the six modes per colour over the
// colour's elements and nodes only.  Each colour writes every row of pv6 and
// qv6 and every slot of ye6 that it reads (the probe covers the colour's
// nodes, which hold all nodes of its elements; the gather reads only the
// colour's element slots), so nothing needs clearing between colours.  No
// memset nodes, so the sequence can be captured.
static hipError_t probe_sequence(jfm_ctx *ctx, bool tangent) {
    hipStream_t s = ctx->stream;
    const int64_t N3 = 3 * ctx->N;
    const int nc = (int)ctx->nc, n_agg = (int)ctx->n_agg;
    const int64_t nslots = 3 * (int64_t)ctx->npe * ctx->n_el;
    k_fill_zero<<<1024, kBlock, 0, s>>>(ctx->Ac, (int64_t)nc * nc);
    hipError_t e = hipGetLastError();
    for (int c = 0; c < ctx->ncolor && e == hipSuccess; ++c) {
        const int e0 = ctx->h_ecol_ptr[c], ne = ctx->h_ecol_ptr[c + 1] - e0;
        const int n0 = ctx->h_ncol_ptr[c], nn = ctx->h_ncol_ptr[c + 1] - n0;
        if (nn == 0) continue;
        k_coarse_probe<<<dim3(nblk(nn), 6), kBlock, 0, s>>>(nn, ctx->ncol + n0, ctx->node_agg, ctx->agg_color, c,
                                                            ctx->cd64, ctx->Rinv, ctx->fixed, ctx->pv6, N3);
        // element products only: the gather below reads just this colour's slots
        e = launch_subset_apply(ctx, tangent, ctx->pv6, ctx->qv6, ctx->ye6, 6, N3, nslots, ctx->ecol + e0, ne,
                                nullptr, 0);
        if (e != hipSuccess) break;
        k_coarse_gather<<<dim3(nblk(nn), 6), kBlock, 0, s>>>(nn, ctx->ncol + n0, ctx->csub_ptr + n0, ctx->csub,
                                                             ctx->ye6, nslots, ctx->fixed, ctx->qv6, N3);
        const int s0 = ctx->h_cseg[c], ns = ctx->h_cseg[c + 1] - s0;
        k_coarse_probe_restrict<<<dim3(ns, 6), kBlock, 0, s>>>(ctx->segb + s0, ctx->sega + s0, ctx->ncol, ctx->cd64,
                                                               ctx->qv6, N3, ctx->partner + (size_t)c * n_agg,
                                                               ctx->Rinv, ctx->Ac, nc);
        e = hipGetLastError();
    }
    return e;
}

// Zero A_c, the element entries, and their sums per aggregate pair (written
// symmetric, unit diagonal for dropped modes).
template <int NPE>
static void gal_elem(jfm_ctx *ctx) {
    const unsigned grid = nblk(6 * (int64_t)ctx->gal_nitems);
    if (ctx->affine)
        k_gal_elem<NPE, true><<<grid, kBlock, 0, ctx->stream>>>(
            ctx->gal_nitems, ctx->gal_item, ctx->gal_el, ctx->gal_gptr, ctx->gal_glist, ctx->gal_gq, ctx->gal_gpos,
            (int)ctx->n_el, ctx->conn, ctx->gf64, ctx->node_agg, ctx->cd64, ctx->Rinv, ctx->fixed, ctx->gal_C);
    else
        k_gal_elem<NPE, false><<<grid, kBlock, 0, ctx->stream>>>(
            ctx->gal_nitems, ctx->gal_item, ctx->gal_el, ctx->gal_gptr, ctx->gal_glist, ctx->gal_gq, ctx->gal_gpos,
            (int)ctx->n_el, ctx->conn, ctx->gf64, ctx->node_agg, ctx->cd64, ctx->Rinv, ctx->fixed, ctx->gal_C);
}

static hipError_t galerkin_sequence(jfm_ctx *ctx) {
    hipStream_t s = ctx->stream;
    const int nc = (int)ctx->nc;
    k_fill_zero<<<1024, kBlock, 0, s>>>(ctx->Ac, (int64_t)nc * nc);
    if (ctx->gal_nitems > 0) {
        switch (ctx->npe) {
        case 10: gal_elem<10>(ctx); break;
        case 4: gal_elem<4>(ctx); break;
        case 8: gal_elem<8>(ctx); break;
        default: return hipErrorInvalidValue;
        }
    }
    if (ctx->gal_nupair > 0)
        k_gal_sum<<<ctx->gal_nupair, 1024, 0, s>>>(ctx->gal_upair, ctx->gal_pptr, ctx->gal_C, ctx->Ac, nc);
    return hipGetLastError();
}

// A_c by the element-level product for K, by probing for the tangent (whose
// element matrices do not annihilate rigid rotations), or by probing for
// both with JFEM_COARSE_PROBE=1.
static bool coarse_probe(bool tangent) {
    static const bool env = [] {
        const char *v = std::getenv("JFEM_COARSE_PROBE");
        return v && v[0] == '1';
    }();
    return tangent || env;
}

// A_c (symmetric; unit diagonal for dropped modes) into ctx->Ac
static hipError_t coarse_matrix_sequence(jfm_ctx *ctx, bool tangent) {
    if (!coarse_probe(tangent)) return galerkin_sequence(ctx);
    hipError_t e = probe_sequence(ctx, tangent);
    if (e != hipSuccess) return e;
    const int nc = (int)ctx->nc;
    k_coarse_fix<<<nblk((int64_t)nc * nc), kBlock, 0, ctx->stream>>>(nc, ctx->Ac);
    return hipGetLastError();
}

// A_c, its inverse and the fold into Ahat32; a non-SPD pivot sets the flag
// at ctx->counter + 48.
// A_c (the not-SPD flag cleared first) ...
static hipError_t setup_matrix_sequence(jfm_ctx *ctx, bool tangent) {
    k_flag_clear<<<1, 1, 0, ctx->stream>>>((int *)(ctx->counter + 48));
    return coarse_matrix_sequence(ctx, tangent);
}

// ... then its inverse and the fold into Ahat
static hipError_t setup_inverse_sequence(jfm_ctx *ctx) {
    hipStream_t s = ctx->stream;
    const int nc = (int)ctx->nc;
    int *flag = (int *)(ctx->counter + 48);
    const int tiles = (nc + 63) / 64;
    // gjP: this step's P^-1, the next step's as the update forms it, and the
    // next pivot block before the update
    const double *inv = ctx->Ac;  // where the inverse ends
    // one kernel per step, from one copy of the matrix into the other; P^-1
    // alternates between two slots of gjP
    double *Ab[2] = {ctx->Ac, ctx->Ac2}, *Pb[2] = {ctx->gjP, ctx->gjP + kGJ * kGJ};
    k_gj_first<<<1, 256, 0, s>>>(nc, std::min(kGJ, nc), ctx->Ac, Pb[0], flag);
    int st = 0;
    for (int k0 = 0; k0 < nc; k0 += kGJ, ++st) {
        const int bk = std::min(kGJ, nc - k0);
        const int nk0 = k0 + kGJ, nbk = nk0 < nc ? std::min(kGJ, nc - nk0) : 0;
        k_gj_step<<<tiles * (tiles + 1) / 2 + (nbk > 0 ? 1 : 0), 256, 0, s>>>(nc, k0, bk, Ab[st & 1], Ab[(st + 1) & 1],
                                                                           Pb[st & 1], nk0, nbk, Pb[(st + 1) & 1], flag);
    }
    inv = Ab[st & 1];
    k_coarse_fold_blocks<<<nblk((int64_t)ctx->n_agg * ctx->n_agg), kBlock, 0, s>>>(
        (int)ctx->n_agg, nc, (int)ctx->ldc, inv, ctx->Rinv, ctx->dropped, ctx->Ahat32);
    return hipGetLastError();
}

// The aggregation for coarse_nodes (kept while the mesh, the fixed set and
// coarse_nodes stay the same) and, for a probed operator, the probe vectors.
static int coarse_prepare(jfm_ctx *ctx, int op, int coarse_nodes) {
    hipStream_t s = ctx->stream;
    const int64_t N = ctx->N, n = ctx->ndof;
    if (ctx->coarse_built_nodes != coarse_nodes || !ctx->node_agg) {
        jfem_coarse_release(ctx);
        std::vector<uint8_t> fx(n);
        JFM_HIP(ctx, hipMemcpyAsync(fx.data(), ctx->fixed, n, hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipStreamSynchronize(s));
        std::vector<int32_t> slot((size_t)ctx->npe * ctx->n_el);
        JFM_HIP(ctx, hipMemcpyAsync(slot.data(), ctx->inc_slot, 4 * slot.size(), hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipStreamSynchronize(s));
        Aggregation A;
        if (int rc = coarse_host_setup(ctx, coarse_nodes, fx, slot, A)) return rc;
        const int64_t nc = 6 * (int64_t)A.n_agg;
        if (nc > 20000) return jfm_fail(ctx, JFM_EINVAL, "coarse level too large: raise coarse_nodes");
        ctx->n_agg = A.n_agg;
        ctx->nc = nc;
        ctx->ncolor = A.ncolor;
        std::vector<float> cd32(A.cd.begin(), A.cd.end());
#define CALLOC(ptr, T, cnt)                                   \
    do {                                                      \
        ptr = jfm_alloc<T>(ctx, cnt);                         \
        if (!ptr) return jfm_fail(ctx, JFM_ENOMEM, "coarse level allocation failed"); \
    } while (0)
        CALLOC(ctx->node_agg, int32_t, N);
        CALLOC(ctx->agg_ptr, int32_t, A.n_agg + 1);
        CALLOC(ctx->agg_nodes, int32_t, N);
        CALLOC(ctx->agg_color, int32_t, A.n_agg);
        CALLOC(ctx->partner, int32_t, (int64_t)A.ncolor * A.n_agg);
        CALLOC(ctx->cd64, double, 3 * N);
        CALLOC(ctx->cd32, float, 3 * N);
        CALLOC(ctx->Rinv, double, 36 * (int64_t)A.n_agg);
        CALLOC(ctx->dropped, int32_t, nc);
        CALLOC(ctx->ecol, int32_t, std::max<int64_t>(1, (int64_t)A.ecol.size()));
        CALLOC(ctx->ncol, int32_t, std::max<int64_t>(1, (int64_t)A.ncol.size()));
        CALLOC(ctx->segb, int32_t, (int64_t)A.segb.size());
        CALLOC(ctx->csub_ptr, int32_t, (int64_t)A.csub_ptr.size());
        CALLOC(ctx->csub, int32_t, std::max<int64_t>(1, (int64_t)A.csub.size()));
        CALLOC(ctx->sega, int32_t, std::max<int64_t>(1, (int64_t)A.sega.size()));
        CALLOC(ctx->Ac, double, nc * nc);
        CALLOC(ctx->Ac2, double, nc * nc);
        CALLOC(ctx->gjP, double, 2 * kGJ * kGJ);
        if ((int64_t)A.gal_el.size() >= (int64_t(1) << 27))
            return jfm_fail(ctx, JFM_EINVAL, "too many elements for the coarse level");
        ctx->gal_nitems = (int)A.gal_item.size();
        ctx->gal_nupair = (int)(A.gal_upair.size() / 4);
        const int64_t n_entries = A.gal_pptr.back();
        CALLOC(ctx->gal_el, int32_t, std::max<int64_t>(1, (int64_t)A.gal_el.size()));
        CALLOC(ctx->gal_gptr, int32_t, (int64_t)A.gal_gptr.size());
        CALLOC(ctx->gal_glist, int32_t, std::max<int64_t>(1, (int64_t)A.gal_glist.size()));
        CALLOC(ctx->gal_gq, int32_t, std::max<int64_t>(1, (int64_t)A.gal_gq.size()));
        CALLOC(ctx->gal_gpos, int32_t, std::max<int64_t>(1, (int64_t)A.gal_gpos.size()));
        CALLOC(ctx->gal_item, int32_t, std::max<int64_t>(1, (int64_t)A.gal_item.size()));
        CALLOC(ctx->gal_pptr, int32_t, (int64_t)A.gal_pptr.size());
        CALLOC(ctx->gal_upair, int32_t, std::max<int64_t>(4, (int64_t)A.gal_upair.size()));
        CALLOC(ctx->gal_C, double, std::max<int64_t>(36, 36 * n_entries));
        auto up = [&](int32_t *d, const std::vector<int32_t> &v) -> hipError_t {
            return v.empty() ? hipSuccess : hipMemcpyAsync(d, v.data(), 4 * v.size(), hipMemcpyHostToDevice, s);
        };
        JFM_HIP(ctx, up(ctx->gal_el, A.gal_el));
        JFM_HIP(ctx, up(ctx->gal_gptr, A.gal_gptr));
        JFM_HIP(ctx, up(ctx->gal_glist, A.gal_glist));
        JFM_HIP(ctx, up(ctx->gal_gq, A.gal_gq));
        JFM_HIP(ctx, up(ctx->gal_gpos, A.gal_gpos));
        JFM_HIP(ctx, up(ctx->gal_item, A.gal_item));
        JFM_HIP(ctx, up(ctx->gal_pptr, A.gal_pptr));
        JFM_HIP(ctx, up(ctx->gal_upair, A.gal_upair));
        const int64_t ldc = (nc + 255) / 256 * 256;
        ctx->ldc = ldc;
        CALLOC(ctx->Ahat32, float, nc * ldc);
        CALLOC(ctx->wc32, float, ldc);
        CALLOC(ctx->wc64, double, 2 * nc);
        CALLOC(ctx->zc32, float, nc);
        CALLOC(ctx->cdA64, double, 3 * N);
        CALLOC(ctx->apos, int32_t, N);
        CALLOC(ctx->rA, double, 3 * N);
        CALLOC(ctx->cdA32, float, 3 * N);
#undef CALLOC
        JFM_HIP(ctx, hipMemcpyAsync(ctx->node_agg, A.node_agg.data(), 4 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_ptr, A.agg_ptr.data(), 4 * (A.n_agg + 1), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_nodes, A.agg_nodes.data(), 4 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_color, A.color.data(), 4 * A.n_agg, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->partner, A.partner.data(), 4 * A.partner.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cd64, A.cd.data(), 8 * 3 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cd32, cd32.data(), 4 * 3 * N, hipMemcpyHostToDevice, s));
        std::vector<double> cdA(3 * N);
        for (int64_t k = 0; k < N; ++k)
            for (int c = 0; c < 3; ++c) cdA[3 * k + c] = A.cd[3 * (size_t)A.agg_nodes[k] + c];
        std::vector<float> cdA32(cdA.begin(), cdA.end());
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cdA64, cdA.data(), 8 * 3 * N, hipMemcpyHostToDevice, s));
        {
            std::vector<int32_t> apos(N);
            for (int64_t k = 0; k < N; ++k) apos[A.agg_nodes[k]] = (int32_t)k;
            JFM_HIP(ctx, hipMemcpyAsync(ctx->apos, apos.data(), 4 * N, hipMemcpyHostToDevice, s));
            JFM_HIP(ctx, hipStreamSynchronize(s));  // apos goes out of scope
        }
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cdA32, cdA32.data(), 4 * 3 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemsetAsync(ctx->wc32, 0, 4 * ldc, s));  // padding stays zero
        JFM_HIP(ctx, hipMemcpyAsync(ctx->Rinv, A.Rinv.data(), 8 * A.Rinv.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->dropped, A.dropped.data(), 4 * nc, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->ecol, A.ecol.data(), 4 * A.ecol.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->ncol, A.ncol.data(), 4 * A.ncol.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->segb, A.segb.data(), 4 * A.segb.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->csub_ptr, A.csub_ptr.data(), 4 * A.csub_ptr.size(), hipMemcpyHostToDevice, s));
        if (!A.csub.empty())
            JFM_HIP(ctx, hipMemcpyAsync(ctx->csub, A.csub.data(), 4 * A.csub.size(), hipMemcpyHostToDevice, s));
        if (!A.sega.empty())
            JFM_HIP(ctx, hipMemcpyAsync(ctx->sega, A.sega.data(), 4 * A.sega.size(), hipMemcpyHostToDevice, s));
        ctx->h_ecol_ptr = A.ecol_ptr;
        ctx->h_ncol_ptr = A.ncol_ptr;
        ctx->h_cseg = A.cseg;
        JFM_HIP(ctx, hipStreamSynchronize(s));
        ctx->coarse_built_nodes = coarse_nodes;
        ++ctx->coarse_gen;
    }
    const bool tangent = op == JFM_OP_TANGENT;
    if (tangent && !ctx->has_state) return jfm_fail(ctx, JFM_ESTATE, "tangent coarse level needs a state");
    if (coarse_probe(tangent) && !ctx->pv6) {
        // probe vectors: zero outside the rows and slots each colour writes
        const int64_t nslots = 3 * (int64_t)ctx->npe * ctx->n_el;
        ctx->pv6 = jfm_alloc<double>(ctx, 6 * 3 * N);
        ctx->qv6 = jfm_alloc<double>(ctx, 6 * 3 * N);
        ctx->ye6 = jfm_alloc<double>(ctx, 6 * nslots);
        if (!ctx->pv6 || !ctx->qv6 || !ctx->ye6) {
            jfem_coarse_release(ctx);
            return jfm_fail(ctx, JFM_ENOMEM, "coarse level allocation failed");
        }
        JFM_HIP(ctx, hipMemsetAsync(ctx->pv6, 0, sizeof(double) * 18 * N, s));
        JFM_HIP(ctx, hipMemsetAsync(ctx->qv6, 0, sizeof(double) * 18 * N, s));
        JFM_HIP(ctx, hipMemsetAsync(ctx->ye6, 0, sizeof(double) * 6 * nslots, s));
    }
    return JFM_OK;
}

// Builds the coarse level for operator `op` (FP64, masked); A_c is rebuilt on
// every call.
int jfem_coarse_setup(jfm_ctx *ctx, int op, int coarse_nodes, double *ms, bool learned_beside) {
    hipStream_t s = ctx->stream;
    hipEvent_t e0, e1;
    JFM_HIP(ctx, hipEventCreate(&e0));
    JFM_HIP(ctx, hipEventCreate(&e1));
    JFM_HIP(ctx, hipEventRecord(e0, s));
    if (int rc = coarse_prepare(ctx, op, coarse_nodes)) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    // each part is the same launch sequence every time for a given
    // aggregation and operator, so it is captured once as a HIP graph
    const bool tangent = op == JFM_OP_TANGENT;
    auto run = [&](hipGraphExec_t &ex, int &gen, auto &&seq) -> hipError_t {
        if (!jfem_use_graphs()) return seq();
        if (ex && gen != ctx->coarse_gen) {
            (void)hipGraphExecDestroy(ex);
            ex = nullptr;
        }
        if (!ex) {
            hipGraph_t graph = nullptr;
            hipError_t le = hipSuccess;
            bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                le = seq();
                ok = hipStreamEndCapture(s, &graph) == hipSuccess && le == hipSuccess;
            }
            if (ok) ok = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess;
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            if (!ok) {
                if (ex) (void)hipGraphExecDestroy(ex);
                ex = nullptr;
            } else {
                gen = ctx->coarse_gen;
            }
        }
        return ex ? hipGraphLaunch(ex, s) : seq();
    };
    hipError_t seq_err = run(ctx->coarse_graph[tangent ? 1 : 0], ctx->coarse_graph_gen[tangent ? 1 : 0],
                             [&] { return setup_matrix_sequence(ctx, tangent); });
    // the learned blocks beside the inversion, whose steps leave most CUs
    // idle (beside the element-level products they slowed those by half)
    int lrc = JFM_OK;
    if (seq_err == hipSuccess && learned_beside) lrc = jfem_learned_launch(ctx);
    if (seq_err == hipSuccess)
        seq_err = run(ctx->coarse_inv_graph, ctx->coarse_inv_gen, [&] { return setup_inverse_sequence(ctx); });
    if (seq_err != hipSuccess) {
        // the probe buffers may no longer be zero: drop the level
        jfem_learned_drop(ctx);
        jfem_coarse_release(ctx);
        JFM_HIP(ctx, seq_err);
    }
    if (lrc) {
        (void)hipStreamSynchronize(s);
        return lrc;
    }
    int hflag = 0;
    JFM_HIP(ctx, hipMemcpyAsync(&hflag, ctx->counter + 48, sizeof(int), hipMemcpyDeviceToHost, s));
    JFM_HIP(ctx, hipEventRecord(e1, s));
    JFM_HIP(ctx, hipEventSynchronize(e1));
    float t = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms += t;
    if (hflag != 0) return jfm_fail(ctx, JFM_ENOTSPD, "coarse matrix is not positive definite");
    ctx->coarse_ready = true;
    ctx->coarse_tag = op == JFM_OP_TANGENT ? 1 : 0;
    return JFM_OK;
}

// A_c itself (before inversion), for checking: nc, then A_c (nc x nc, row
// major) and the node aggregates when the pointers are given.  The coarse
// level has to be rebuilt afterwards.
int jfem_coarse_matrix(jfm_ctx *ctx, int op, int coarse_nodes, int64_t *nc, double *Ac, int32_t *node_agg) {
    if (int rc = coarse_prepare(ctx, op, coarse_nodes)) return rc;
    const bool tangent = op == JFM_OP_TANGENT;
    hipStream_t s = ctx->stream;
    if (nc) *nc = ctx->nc;
    if (node_agg)
        JFM_HIP(ctx, hipMemcpyAsync(node_agg, ctx->node_agg, 4 * ctx->N, hipMemcpyDeviceToHost, s));
    if (Ac) {
        ctx->coarse_ready = false;
        hipError_t e = coarse_matrix_sequence(ctx, tangent);
        if (e != hipSuccess) {
            jfem_coarse_release(ctx);
            JFM_HIP(ctx, e);
        }
        JFM_HIP(ctx, hipMemcpyAsync(Ac, ctx->Ac, 8 * ctx->nc * ctx->nc, hipMemcpyDeviceToHost, s));
    }
    JFM_HIP(ctx, hipStreamSynchronize(s));
    return JFM_OK;
}

// coarse correction after the update kernel of a PCG iteration (or after the
// init kernel with set_p): z += Q A_c^-1 Q' r, partials (r.z, r.r) to `part`
#ifndef JFEM_GEMV_ROWS
#define JFEM_GEMV_ROWS 4
#endif
constexpr int kGemvRows = JFEM_GEMV_ROWS;  // rows (waves) per block of the coarse GEMV
static_assert(64 * kGemvRows <= 256, "k_coarse_gemv is compiled for blocks of at most 256 threads");
static unsigned gemv_grid(int64_t nc) {
    return (unsigned)std::min<int64_t>((nc + kGemvRows - 1) / kGemvRows, kMaxPart);
}

template <typename T, bool AGG>
static void coarse_product(jfm_ctx *ctx, const T *r, const int *done, double *part_wz) {
    const T *cdA = sizeof(T) == 8 ? (const T *)ctx->cdA64 : (const T *)ctx->cdA32;
    const int n_agg = (int)ctx->n_agg, nc = (int)ctx->nc;
    hipStream_t s = ctx->stream;
    k_coarse_restrict<T, AGG><<<n_agg, 1024, 0, s>>>(n_agg, ctx->agg_ptr, ctx->agg_nodes, cdA, ctx->fixed, r, ctx->wc32,
                                                    ctx->wc64, done);
    k_coarse_gemv<<<gemv_grid(nc), 64 * kGemvRows, 0, s>>>(nc, (int)(ctx->ldc / 4), (const float4 *)ctx->Ahat32,
                                                  (const float4 *)ctx->wc32, ctx->wc64, ctx->zc32, part_wz, done);
}

template <typename T>
hipError_t launch_coarse_product(jfm_ctx *ctx, const T *r, const int *done, int *npart_wz) {
    coarse_product<T, false>(ctx, r, done, ctx->partials + kPartCoarse);
    *npart_wz = (int)gemv_grid(ctx->nc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_coarse_product_agg(jfm_ctx *ctx, const T *rA, const int *done, int *npart_wz) {
    coarse_product<T, true>(ctx, rA, done, ctx->partials + kPartCoarse);
    *npart_wz = (int)gemv_grid(ctx->nc);
    return hipGetLastError();
}

bool jfem_coarse_fused_restriction(const jfm_ctx *ctx) {
    static const bool on = [] {
        const char *v = std::getenv("JFEM_FUSED_RESTRICTION");
        return !(v && v[0] == '0');
    }();
    return on && ctx->n_agg > 0;
}

// zc = Ahat w for the iteration of parity `slot`, w as the aggregate-ordered
// update kernel wrote it (wc32, and wc64 + (slot ^ 1) nc)
hipError_t launch_coarse_gemv_fused(jfm_ctx *ctx, int slot, int *npart_wz) {
    const int nc = (int)ctx->nc;
    k_coarse_gemv<<<gemv_grid(nc), 64 * kGemvRows, 0, ctx->stream>>>(
        nc, (int)(ctx->ldc / 4), (const float4 *)ctx->Ahat32, (const float4 *)ctx->wc32,
        ctx->wc64 + (size_t)(slot ^ 1) * nc, ctx->zc32, ctx->partials + kPartCoarse, &ctx->S->done[slot]);
    *npart_wz = (int)gemv_grid(nc);
    return hipGetLastError();
}

template <> double *coarse_rA<double>(jfm_ctx *ctx) { return ctx->rA; }
template <> float *coarse_rA<float>(jfm_ctx *ctx) { return (float *)ctx->rA; }

// coarse correction outside the iteration (PCG start, training):
// z += Q A_c^-1 Q' r, optionally p = z; partials (r.z, r.r) to `part`
template <typename T>
hipError_t launch_coarse_correct(jfm_ctx *ctx, const T *r, T *z, T *p, double *part, const int *done, int *npart) {
    const T *cd = sizeof(T) == 8 ? (const T *)ctx->cd64 : (const T *)ctx->cd32;
    hipStream_t s = ctx->stream;
    coarse_product<T, false>(ctx, r, done, ctx->partials + kPartCoarse);
    const unsigned g = std::min<unsigned>(capped_grid(ctx->N, kBlock), 512);
    *npart = (int)g;
    k_coarse_prolong<T><<<g, kBlock, 0, s>>>((int)ctx->N, ctx->node_agg, cd, ctx->fixed, ctx->zc32, r, z, p, part, done);
    return hipGetLastError();
}

template hipError_t launch_coarse_product<double>(jfm_ctx *, const double *, const int *, int *);
template hipError_t launch_coarse_product<float>(jfm_ctx *, const float *, const int *, int *);
template hipError_t launch_coarse_product_agg<double>(jfm_ctx *, const double *, const int *, int *);
template hipError_t launch_coarse_product_agg<float>(jfm_ctx *, const float *, const int *, int *);
template hipError_t launch_coarse_correct<double>(jfm_ctx *, const double *, double *, double *, double *, const int *, int *);
template hipError_t launch_coarse_correct<float>(jfm_ctx *, const float *, float *, float *, double *, const int *, int *);

}  // namespace jfem
