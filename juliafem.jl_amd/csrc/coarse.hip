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
//  * rocSOLVER potrf + potri invert A_c (dense, a few thousand rows); the
//    inverse is folded with the R_g^-1 into Ahat = R^-1 A_c^-1 R^-T (stored
//    FP32), so the iteration needs only the raw modes: Q A_c^-1 Q' r =
//    Bm Ahat Bm' r.
//
// Per PCG iteration (after the update kernel has written z = B r):
//  k_coarse_restrict  w_g = sum_{a in g} [r_a ; d_a x r_a]   (one block per g)
//  k_coarse_gemv      zc = Ahat w                           (one wave per row)
//  k_coarse_prolong   z_a += m_a o (t_g + omega_g x d_a), partials of r.z, r.r
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "jfem_device.h"

using namespace jfem;

namespace jfem {

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// v = sum over aggregates of colour c of Q_g e_j (masked), 0 elsewhere
__global__ __launch_bounds__(256) void k_coarse_probe(int N, const int32_t *__restrict__ node_agg,
                                                     const int32_t *__restrict__ agg_color, int c, int j,
                                                     const double *__restrict__ cd, const double *__restrict__ Rinv,
                                                     const uint8_t *__restrict__ fixed, double *__restrict__ v) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
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

// w~_g = sum_{a in g} [y_a ; d_a x y_a] in a fixed order (block per aggregate)
template <typename T>
__device__ __forceinline__ void agg_sums(int g, const int32_t *__restrict__ agg_ptr, const int32_t *__restrict__ agg_nodes,
                                         const T *__restrict__ cd, const T *__restrict__ y, double (&w)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = 0.0;
    for (int k = agg_ptr[g] + threadIdx.x; k < agg_ptr[g + 1]; k += blockDim.x) {
        const size_t o = 3 * (size_t)agg_nodes[k];
        const double y0 = y[o], y1 = y[o + 1], y2 = y[o + 2];
        const double d0 = cd[o], d1 = cd[o + 1], d2 = cd[o + 2];
        w[0] += y0;
        w[1] += y1;
        w[2] += y2;
        w[3] += d1 * y2 - d2 * y1;
        w[4] += d2 * y0 - d0 * y2;
        w[5] += d0 * y1 - d1 * y0;
    }
    block_sum_all<6>(w);
}

// A_c rows of aggregate g for probe (c, j): Rinv_g' w~_g into column 6p + j
__global__ __launch_bounds__(256) void k_coarse_probe_restrict(int n_agg, const int32_t *__restrict__ agg_ptr,
                                                              const int32_t *__restrict__ agg_nodes,
                                                              const double *__restrict__ cd, const double *__restrict__ y,
                                                              const int32_t *__restrict__ partner_c,
                                                              const double *__restrict__ Rinv, int j, double *__restrict__ Ac,
                                                              int nc) {
    const int g = blockIdx.x;
    const int p = partner_c[g];
    if (p < 0) return;  // uniform per block
    double w[6];
    agg_sums<double>(g, agg_ptr, agg_nodes, cd, y, w);
    if (threadIdx.x < 6) {
        const int i = threadIdx.x;
        const double *R = Rinv + 36 * (size_t)g;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += R[k * 6 + i] * w[k];
        Ac[(size_t)(6 * g + i) * nc + 6 * p + j] = s;
    }
}

// zero the incidence slots of the listed elements and the listed rows of y
__global__ __launch_bounds__(256) void k_coarse_clear(int ne, const int32_t *__restrict__ elist, int npe, int n_el,
                                                     const int32_t *__restrict__ slot, double *__restrict__ ye, int nn,
                                                     const int32_t *__restrict__ nlist, double *__restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ne) {
        const int e = elist[k];
        for (int i = 0; i < npe; ++i) {
            double *o = ye + 3 * (size_t)slot[(size_t)i * n_el + e];
            o[0] = 0.0; o[1] = 0.0; o[2] = 0.0;
        }
    }
    if (k < nn) {
        const size_t o = 3 * (size_t)nlist[k];
        y[o] = 0.0; y[o + 1] = 0.0; y[o + 2] = 0.0;
    }
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

// Ahat = R^-1 A^-1 R^-T from the lower triangle (column-major) that potri
// leaves: entry (r, s) of A^-1 with r >= s is at L[s * nc + r].
__global__ __launch_bounds__(256) void k_coarse_fold(int nc, const double *__restrict__ L, const double *__restrict__ Rinv,
                                                    const int32_t *__restrict__ dropped, float *__restrict__ A32) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nc * nc) return;
    const int r = (int)(idx / nc), s = (int)(idx - (int64_t)r * nc);
    const int g = r / 6, i = r - 6 * g, h = s / 6, j = s - 6 * h;
    const double *Rg = Rinv + 36 * (size_t)g, *Rh = Rinv + 36 * (size_t)h;
    double acc = 0.0;
    for (int k = 0; k < 6; ++k) {
        const double rik = Rg[i * 6 + k];
        if (rik == 0.0) continue;
        const int rr = 6 * g + k;
        double inner = 0.0;
        for (int l = 0; l < 6; ++l) {
            const int ss = 6 * h + l;
            double a = rr >= ss ? L[(size_t)ss * nc + rr] : L[(size_t)rr * nc + ss];
            if (dropped[rr] || dropped[ss]) a = 0.0;
            inner += a * Rh[j * 6 + l];
        }
        acc += rik * inner;
    }
    A32[idx] = (float)acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_coarse_restrict(int n_agg, const int32_t *__restrict__ agg_ptr,
                                                        const int32_t *__restrict__ agg_nodes, const T *__restrict__ cd,
                                                        const T *__restrict__ r, float *__restrict__ wc,
                                                        const int *__restrict__ done) {
    if (done && *done) return;
    const int g = blockIdx.x;
    double w[6];
    agg_sums<T>(g, agg_ptr, agg_nodes, cd, r, w);
    if (threadIdx.x < 6) wc[6 * g + threadIdx.x] = (float)w[threadIdx.x];
}

// zc = Ahat w, one 64-lane wave per row, fixed-order lane sums.  Ahat is
// kept in FP32 for both precisions: it only scales the coarse correction of a
// fixed SPD preconditioner, so its rounding changes the iteration count at
// most, never the answer, and it halves the dominant read of the level.
__global__ __launch_bounds__(256) void k_coarse_gemv(int nc, const float *__restrict__ A, const float *__restrict__ w,
                                                    float *__restrict__ zc, const int *__restrict__ done) {
    if (done && *done) return;
    const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (row >= nc) return;  // whole wave leaves together
    const float *Ar = A + (size_t)row * nc;
    double s = 0.0;
    for (int k = lane; k < nc; k += 64) s += (double)Ar[k] * (double)w[k];
    s = wave_sum(s);
    if (lane == 0) zc[row] = (float)s;
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

}  // namespace

// Host-side aggregation, colouring and per-aggregate mode bases.
static int coarse_host_setup(jfm_ctx *ctx, int coarse_nodes, const std::vector<uint8_t> &fixed, Aggregation &A) {
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
        std::sort(nd[c].begin(), nd[c].end());
        A.ecol_ptr[c + 1] = A.ecol_ptr[c] + (int32_t)el[c].size();
        A.ncol_ptr[c + 1] = A.ncol_ptr[c] + (int32_t)nd[c].size();
    }
    A.ecol.reserve(A.ecol_ptr[ncolor]);
    A.ncol.reserve(A.ncol_ptr[ncolor]);
    for (int c = 0; c < ncolor; ++c) {
        A.ecol.insert(A.ecol.end(), el[c].begin(), el[c].end());
        A.ncol.insert(A.ncol.end(), nd[c].begin(), nd[c].end());
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
    free_dev(ctx, ctx->Rinv); free_dev(ctx, ctx->Ac); free_dev(ctx, ctx->Ahat32);
    free_dev(ctx, ctx->wc32); free_dev(ctx, ctx->zc32);
    free_dev(ctx, ctx->dropped); free_dev(ctx, ctx->ecol); free_dev(ctx, ctx->ncol);
    ctx->h_ecol_ptr.clear(); ctx->h_ncol_ptr.clear();
    ctx->coarse_ready = false;
    ctx->coarse_tag = -1;
    ctx->coarse_built_nodes = 0;
    ctx->n_agg = ctx->nc = 0;
}

void jfem_coarse_destroy_handle(jfm_ctx *ctx) {
    if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
}

// Builds the coarse level for operator `op` (FP64, masked).  The aggregation
// is kept while the mesh, the fixed set and coarse_nodes stay the same; A_c is
// rebuilt on every call.
int jfem_coarse_setup(jfm_ctx *ctx, int op, int coarse_nodes, double *ms) {
    hipStream_t s = ctx->stream;
    hipEvent_t e0, e1;
    JFM_HIP(ctx, hipEventCreate(&e0));
    JFM_HIP(ctx, hipEventCreate(&e1));
    JFM_HIP(ctx, hipEventRecord(e0, s));
    const int64_t N = ctx->N, n = ctx->ndof;
    if (ctx->coarse_built_nodes != coarse_nodes || !ctx->node_agg) {
        jfem_coarse_release(ctx);
        std::vector<uint8_t> fx(n);
        JFM_HIP(ctx, hipMemcpyAsync(fx.data(), ctx->fixed, n, hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipStreamSynchronize(s));
        Aggregation A;
        if (int rc = coarse_host_setup(ctx, coarse_nodes, fx, A)) return rc;
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
        CALLOC(ctx->Ac, double, nc * nc);
        CALLOC(ctx->Ahat32, float, nc * nc);
        CALLOC(ctx->wc32, float, nc);
        CALLOC(ctx->zc32, float, nc);
#undef CALLOC
        JFM_HIP(ctx, hipMemcpyAsync(ctx->node_agg, A.node_agg.data(), 4 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_ptr, A.agg_ptr.data(), 4 * (A.n_agg + 1), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_nodes, A.agg_nodes.data(), 4 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->agg_color, A.color.data(), 4 * A.n_agg, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->partner, A.partner.data(), 4 * A.partner.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cd64, A.cd.data(), 8 * 3 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->cd32, cd32.data(), 4 * 3 * N, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->Rinv, A.Rinv.data(), 8 * A.Rinv.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->dropped, A.dropped.data(), 4 * nc, hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->ecol, A.ecol.data(), 4 * A.ecol.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->ncol, A.ncol.data(), 4 * A.ncol.size(), hipMemcpyHostToDevice, s));
        ctx->h_ecol_ptr = A.ecol_ptr;
        ctx->h_ncol_ptr = A.ncol_ptr;
        JFM_HIP(ctx, hipStreamSynchronize(s));
        ctx->coarse_built_nodes = coarse_nodes;
        ++ctx->coarse_gen;
    }
    const int nc = (int)ctx->nc, n_agg = (int)ctx->n_agg;
    // A_c by probing; the FP64 PCG direction vectors are free before a solve
    double *pv = (double *)ctx->ws64.p, *qv = (double *)ctx->ws64.q;
    if (!pv || !qv) return jfm_fail(ctx, JFM_ESTATE, "coarse setup needs the solver workspaces");
    JFM_HIP(ctx, hipMemsetAsync(ctx->Ac, 0, sizeof(double) * (size_t)nc * nc, s));
    // The products touch only the elements next to colour c and their nodes:
    // every other incidence slot and row of qv is kept at zero (cleared once
    // here, and after each product for the rows and slots it wrote).
    const bool tangent = op == JFM_OP_TANGENT;
    if (tangent && !ctx->has_state) return jfm_fail(ctx, JFM_ESTATE, "tangent coarse level needs a state");
    JFM_HIP(ctx, hipMemsetAsync(ctx->ye64, 0, sizeof(double) * 3 * (size_t)ctx->npe * ctx->n_el, s));
    JFM_HIP(ctx, hipMemsetAsync(qv, 0, sizeof(double) * 3 * (size_t)N, s));
    for (int c = 0; c < ctx->ncolor; ++c) {
        const int e0 = ctx->h_ecol_ptr[c], ne = ctx->h_ecol_ptr[c + 1] - e0;
        const int n0 = ctx->h_ncol_ptr[c], nn = ctx->h_ncol_ptr[c + 1] - n0;
        for (int j = 0; j < 6; ++j) {
            k_coarse_probe<<<nblk(N), kBlock, 0, s>>>((int)N, ctx->node_agg, ctx->agg_color, c, j, ctx->cd64, ctx->Rinv,
                                                      ctx->fixed, pv);
            JFM_HIP(ctx, hipGetLastError());
            JFM_HIP(ctx, launch_subset_apply(ctx, tangent, pv, qv, ctx->ecol + e0, ne, ctx->ncol + n0, nn));
            k_coarse_probe_restrict<<<n_agg, kBlock, 0, s>>>(n_agg, ctx->agg_ptr, ctx->agg_nodes, ctx->cd64, qv,
                                                             ctx->partner + (size_t)c * n_agg, ctx->Rinv, j, ctx->Ac, nc);
            JFM_HIP(ctx, hipGetLastError());
            if (std::max(ne, nn) > 0) {
                k_coarse_clear<<<nblk(std::max(ne, nn)), kBlock, 0, s>>>(ne, ctx->ecol + e0, ctx->npe, (int)ctx->n_el,
                                                                          ctx->inc_slot, ctx->ye64, nn, ctx->ncol + n0, qv);
                JFM_HIP(ctx, hipGetLastError());
            }
        }
    }
    const int64_t nn = (int64_t)nc * nc;
    k_coarse_fix<<<nblk(nn), kBlock, 0, s>>>(nc, ctx->Ac);
    JFM_HIP(ctx, hipGetLastError());
    // Cholesky inverse (rocSOLVER, column-major lower = our upper; A_c is symmetric)
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return jfm_fail(ctx, JFM_EHIP, "rocblas_create_handle failed");
        ctx->blas = h;
    }
    rocblas_handle h = (rocblas_handle)ctx->blas;
    if (rocblas_set_stream(h, s) != rocblas_status_success) return jfm_fail(ctx, JFM_EHIP, "rocblas_set_stream failed");
    int *info = (int *)(ctx->counter + 48);  // small integer scratch
    if (rocsolver_dpotrf(h, rocblas_fill_lower, nc, ctx->Ac, nc, info) != rocblas_status_success)
        return jfm_fail(ctx, JFM_EHIP, "rocsolver_dpotrf failed");
    int hinfo = 0;
    JFM_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    JFM_HIP(ctx, hipStreamSynchronize(s));
    if (hinfo != 0) return jfm_fail(ctx, JFM_ENOTSPD, "coarse matrix is not positive definite");
    if (rocsolver_dpotri(h, rocblas_fill_lower, nc, ctx->Ac, nc, info) != rocblas_status_success)
        return jfm_fail(ctx, JFM_EHIP, "rocsolver_dpotri failed");
    k_coarse_fold<<<nblk(nn), kBlock, 0, s>>>(nc, ctx->Ac, ctx->Rinv, ctx->dropped, ctx->Ahat32);
    JFM_HIP(ctx, hipGetLastError());
    JFM_HIP(ctx, hipEventRecord(e1, s));
    JFM_HIP(ctx, hipEventSynchronize(e1));
    float t = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms += t;
    ctx->coarse_ready = true;
    ctx->coarse_tag = op == JFM_OP_TANGENT ? 1 : 0;
    return JFM_OK;
}

// coarse correction after the update kernel of a PCG iteration (or after the
// init kernel with set_p): z += Q A_c^-1 Q' r, partials (r.z, r.r) to `part`
template <typename T>
hipError_t launch_coarse_correct(jfm_ctx *ctx, const T *r, T *z, T *p, double *part, const int *done, int *npart) {
    const T *cd = sizeof(T) == 8 ? (const T *)ctx->cd64 : (const T *)ctx->cd32;
    const int n_agg = (int)ctx->n_agg, nc = (int)ctx->nc;
    hipStream_t s = ctx->stream;
    k_coarse_restrict<T><<<n_agg, kBlock, 0, s>>>(n_agg, ctx->agg_ptr, ctx->agg_nodes, cd, r, ctx->wc32, done);
    k_coarse_gemv<<<nblk(nc, kBlock / 64), kBlock, 0, s>>>(nc, ctx->Ahat32, ctx->wc32, ctx->zc32, done);
    const unsigned g = std::min<unsigned>(capped_grid(ctx->N, kBlock), 512);
    *npart = (int)g;
    k_coarse_prolong<T><<<g, kBlock, 0, s>>>((int)ctx->N, ctx->node_agg, cd, ctx->fixed, ctx->zc32, r, z, p, part, done);
    return hipGetLastError();
}

template hipError_t launch_coarse_correct<double>(jfm_ctx *, const double *, double *, double *, double *, const int *, int *);
template hipError_t launch_coarse_correct<float>(jfm_ctx *, const float *, float *, float *, double *, const int *, int *);

}  // namespace jfem
