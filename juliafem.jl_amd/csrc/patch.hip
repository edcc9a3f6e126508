// patch.hip -- fused matrix-free K.u over element patches (gfx950).
//
// The two-phase operator (fem_kernels.hip) writes every element's 3 x npe
// terms to HBM and reads them back in the nodal gather.  Here the elements are
// ordered along a Morton curve of their centroids and cut into patches of
// kPatch = 480 (one workgroup each, one element per thread).  A workgroup computes its
// elements' terms with the same element product (mf_element.h) into LDS and
// then sums them per patch node, in ascending local element order, from a
// patch-local incidence list:
//  * a node whose incidences all lie in the patch (interior) is finished and
//    written to y directly;
//  * any other node writes the patch's partial sum to a slot; a second kernel
//    adds each such node's partials in ascending patch order.
// Every node is written exactly once (nodes of no element through the second
// kernel, as zero), there are no floating-point atomics and the result does
// not depend on the launch: the order of every sum is fixed by the patch
// structure.  The result equals the two-phase operator's up to the order of
// the nodal sums.
//
// Reference: the operator of ext/JuliaFEMCUDAExt.jl:179-193 / 293-361 (element
// loop and nodal assembly), elements per src/problems_elasticity.jl:203-451.
#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "ebe_element.h"
#include "jfem_device.h"
#include "mf_element.h"

namespace jfem {

// Elements per patch = threads per workgroup.  The LDS image (57.6 KB FP32,
// 115 KB FP64 for Tet10) allows two workgroups per CU, so the patch count is
// best a multiple of 2 x 256: 480 gives exactly 512 patches for 245,760
// elements and 4,096 for 1,966,080 (512 elements: 480 patches, 32 CUs with
// one, and 4% slower).
#ifndef JFEM_PATCH_ELEMS
#define JFEM_PATCH_ELEMS 480
#endif
constexpr int kPatch = JFEM_PATCH_ELEMS;

// Fused element product and patch-local nodal sums.  LDS holds the patch's
// element terms as ys[(3 i + c) * kPatch + local element].  The patch's nodes
// are taken in rounds of kPatch, one per thread; round q lists the j-th
// incidence of its thread t's node at linT[rbase[q] + j kPatch + t] (j below
// the node's count), so the lanes of a wave read consecutive entries.  Nodes
// are sorted by count within a patch, so a round's counts are close; the
// first rounds' node data is loaded before the element product.  A
// shared node's partial sum goes to its slot s (pn_out = -(s + 1)):
// bpart[3 s + c].
// Block partials of y.dotv over the interior nodes go to ppart; with ppart
// and no dotv (affine Tet10 matrix-free form), the block partials of the
// element energies x_e.K_e x_e instead, which sum to x.Kx.
// The element product is the matrix-free one (mf_element.h) from the
// patch-ordered geometric factors el, or with EBE the stored-matrix one
// (ebe_element.h) from element matrices el stored in the patch order.
template <int NPE, typename T, bool AFF, bool EBE, bool C16, bool FOLD>
__global__ __launch_bounds__(kPatch) void k_ku_patch(int n_el, int npatch, int npn, const int32_t *__restrict__ connP,
                                                    const uint16_t *__restrict__ connP16, const int32_t *__restrict__ pbase,
                                                    const T *__restrict__ el, const T *__restrict__ x,
                                                    const int32_t *__restrict__ pn_ptr, const uint16_t *__restrict__ pn_cnt,
                                                    const int32_t *__restrict__ pn_out, const int32_t *__restrict__ rstart,
                                                    const int32_t *__restrict__ rbase, const uint16_t *__restrict__ linT,
                                                    const uint8_t *__restrict__ fixed, T *__restrict__ y,
                                                    T *__restrict__ bpart, const T *__restrict__ dotv,
                                                    double *__restrict__ ppart, const int *__restrict__ done,
                                                    T rho) {
    __shared__ T ys[3 * NPE * kPatch];
    const int p = xcd_block();
    if (p >= npatch) return;  // whole workgroup (grid padded to a multiple of 8)
    const int dn = done ? *done : 0;  // tested once the prefetches below are in flight
    const int t = threadIdx.x;
    const int k = p * kPatch + t;
    // the first kPreR rounds' node data and first incidences, loaded before
    // the element product so that their latency overlaps it
    const int lb = pn_ptr[p], le = pn_ptr[p + 1], q0 = rstart[p];
    constexpr int kPre = 4, kPreR = sizeof(T) == 4 ? 3 : 1;  // FP64: registers are short
    constexpr int kExt = 20;  // round 0's incidences kPre .. kPre + kExt - 1, loaded before the barrier
    const int base0 = lb < le ? rbase[q0] : 0;
    int cnt[kPreR], oo[kPreR], mpre[kPreR][kPre];
    // every index below fits 32 bits: byte offsets from the uniform base
    // pointers keep the per-lane address arithmetic 32-bit
    auto fetch = [&](int r, int l, int &c, int &o, int (&mp)[kPre]) {
        c = load_at<uint16_t>(pn_cnt, 2u * (uint32_t)l);
        o = load_at<int32_t>(pn_out, 4u * (uint32_t)l);
        const int base = rbase[q0 + r];
#pragma unroll
        for (int j = 0; j < kPre; ++j)
            mp[j] = j < c ? (int)load_at<uint16_t>(linT, 2u * (uint32_t)(base + j * kPatch + t)) : 0;
    };
#pragma unroll
    for (int r = 0; r < kPreR; ++r) {
        cnt[r] = 0;
        oo[r] = 0;
#pragma unroll
        for (int j = 0; j < kPre; ++j) mpre[r][j] = 0;
        const int l = lb + r * kPatch + t;
        if (l < le) fetch(r, l, cnt[r], oo[r], mpre[r]);
    }
    if (dn) return;  // uniform
    // p.Kp of the PCG step as the sum of the elements' energies (affine Tet10
    // matrix-free form, when partials are asked for without dotv: the caller
    // guarantees x = 0 on fixed DOFs, so x.(masked Kx) = x.Kx)
    constexpr bool kEnergy = !EBE && AFF && NPE == 10;
    double en = 0.0;
    // local node i's id: 16-bit offsets from the patch's lowest id (C16), or
    // the 32-bit ids
    const int pb = C16 ? pbase[p] : 0;  // uniform
    auto node_of = [&](int i) -> uint32_t {
        if constexpr (C16)
            return (uint32_t)pb + load_at<uint16_t>(connP16, 2u * (uint32_t)(i * n_el + k));
        else
            return (uint32_t)load_at<int32_t>(connP, 4u * (uint32_t)(i * n_el + k));
    };
    auto x_of = [&](int i, T (&v)[3]) {
        const Trip<T> w = load_at<Trip<T>>(x, (uint32_t)(3 * sizeof(T)) * node_of(i));
        v[0] = w.a;
        v[1] = w.b;
        v[2] = w.c;
    };
    if (k < n_el) {
        auto to_lds = [&](int i, T f0, T f1, T f2) {
            ys[(3 * i) * kPatch + t] = f0;
            ys[(3 * i + 1) * kPatch + t] = f1;
            ys[(3 * i + 2) * kPatch + t] = f2;
        };
        if constexpr (EBE) {
            T f[NPE][3];
            ebe_element<NPE, T>(n_el, k, connP, el, x, f);
#pragma unroll
            for (int i = 0; i < NPE; ++i) to_lds(i, f[i][0], f[i][1], f[i][2]);
        } else if constexpr (kEnergy) {
            if (ppart && !dotv) {
                // x.Kx of the element from the values in registers (u = x - x_0,
                // y = K_e u: u.y = x.K_e x since K_e annihilates translations),
                // each node's three products in T, the nodes' sums in FP64
                T xs[NPE][3];
#pragma unroll
                for (int i = 0; i < NPE; ++i) x_of(i, xs[i]);
                mf_element_x<NPE, T, AFF, FOLD>(
                    n_el, k, el,
                    [&](int i, T (&v)[3]) {
                        v[0] = xs[i][0];
                        v[1] = xs[i][1];
                        v[2] = xs[i][2];
                    },
                    [&](int i, T f0, T f1, T f2) {
                        to_lds(i, f0, f1, f2);
                        if (i > 0)
                            en += (double)((xs[i][0] - xs[0][0]) * f0 + (xs[i][1] - xs[0][1]) * f1 +
                                           (xs[i][2] - xs[0][2]) * f2);
                    },
                    rho);
            } else {
                mf_element_x<NPE, T, AFF, FOLD>(n_el, k, el, x_of, to_lds, rho);
            }
        } else
            mf_element_x<NPE, T, AFF, FOLD>(n_el, k, el, x_of, to_lds, rho);
    }
    // Round 0 holds the patch's nodes of most incidences (up to 24 in a Tet10
    // box mesh, so one wave would otherwise load 20 of them one after the
    // other after the barrier): the rest of its incidences, and the fixed
    // flags of the prefetched rounds' finished nodes, are loaded here, once
    // the element product has freed its registers, so that their latency
    // overlaps the barrier.
    int mext[kExt];
    Trip<uint8_t> fxp[kPreR];
    {
        const int c = cnt[0];  // 0 past the patch's nodes
#pragma unroll
        for (int j = 0; j < kExt; ++j)
            mext[j] = kPre + j < c ? (int)load_at<uint16_t>(linT, 2u * (uint32_t)(base0 + (kPre + j) * kPatch + t)) : 0;
#pragma unroll
        for (int r = 0; r < kPreR; ++r)
            fxp[r] = (fixed && oo[r] >= 0 && cnt[r] > 0) ? load_at<Trip<uint8_t>>(fixed, 3u * (uint32_t)oo[r])
                                                         : Trip<uint8_t>{0, 0, 0};
    }
    __syncthreads();
    double dv = en;
    // node l of round r: the sum of its incidences, to y or to its slot
    // (fx: its fixed flags, when loaded above)
    auto node = [&](int r, int l, int c, int o, const int (&mp)[kPre], const Trip<uint8_t> *fx) {
        T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
        for (int j = 0; j < kPre; ++j)
            if (j < c) {
                const int m = mp[j];  // (3 i) * kPatch + local element
                s0 += ys[m];
                s1 += ys[m + kPatch];
                s2 += ys[m + 2 * kPatch];
            }
        int j1 = kPre;
        if (r == 0) {
#pragma unroll
            for (int j = 0; j < kExt; ++j)
                if (kPre + j < c) {
                    const int m = mext[j];
                    s0 += ys[m];
                    s1 += ys[m + kPatch];
                    s2 += ys[m + 2 * kPatch];
                }
            j1 = kPre + kExt;
        }
        if (c > j1) {
            const int base = rbase[q0 + r];
            for (int j = j1; j < c; ++j) {
                const int m = load_at<uint16_t>(linT, 2u * (uint32_t)(base + j * kPatch + t));
                s0 += ys[m];
                s1 += ys[m + kPatch];
                s2 += ys[m + 2 * kPatch];
            }
        }
        if (o >= 0) {
            const uint32_t b = 3u * (uint32_t)o;
            if (fixed) {
                const Trip<uint8_t> f = fx ? *fx : load_at<Trip<uint8_t>>(fixed, b);
                if (f.a) s0 = T(0);
                if (f.b) s1 = T(0);
                if (f.c) s2 = T(0);
            }
            store_at<Trip<T>>(y, (uint32_t)sizeof(T) * b, Trip<T>{s0, s1, s2});
            if (dotv) {
                const Trip<T> d = load_at<Trip<T>>(dotv, (uint32_t)sizeof(T) * b);
                dv += (double)s0 * (double)d.a + (double)s1 * (double)d.b + (double)s2 * (double)d.c;
            }
        } else {
            store_at<Trip<T>>(bpart, (uint32_t)sizeof(T) * 3u * (uint32_t)(-o - 1), Trip<T>{s0, s1, s2});
        }
    };
#pragma unroll
    for (int r = 0; r < kPreR; ++r) {
        const int l = lb + r * kPatch + t;
        if (l < le) node(r, l, cnt[r], oo[r], mpre[r], &fxp[r]);
    }
    for (int r = kPreR; lb + r * kPatch < le; ++r) {  // patches of more than kPreR rounds
        const int l = lb + r * kPatch + t;
        if (l < le) {
            int c, o, mp[kPre];
            fetch(r, l, c, o, mp);
            node(r, l, c, o, mp, nullptr);
        }
    }
    if (ppart) {
        double v[1] = {dv};
        block_sum_all<1>(v);
        if (t == 0) ppart[p] = v[0];
    }
}

// Nodes shared between patches (and nodes of no element): y = sum of the
// patch partials in slots bptr[b]..bptr[b+1] (ascending patch order).  The
// shared nodes are listed in the order of their first patch and each block
// takes a contiguous chunk, with the XCD mapping of the patch kernel, so that
// most partials are read from the L2 that received them.  Block partials of
// y.dotv also take in a fixed slice of the patch kernel's partials.
template <typename T>
__global__ __launch_bounds__(256) void k_ku_combine(int nb, int chunk, const int32_t *__restrict__ bptr,
                                                   const int32_t *__restrict__ bnode, const T *__restrict__ bpart,
                                                   const uint8_t *__restrict__ fixed, T *__restrict__ y,
                                                   const T *__restrict__ dotv, const double *__restrict__ ppart,
                                                   int npatch, double *__restrict__ part, const int *__restrict__ done) {
    // the first node's slot range, its first kPre partials, its mask and its
    // dotv values are all in flight before the convergence flag is tested
    // (the grid gives most threads exactly one node)
    constexpr int kPre = 4;
    const int b0 = xcd_block() * chunk, b1 = min(b0 + chunk, nb);
    auto sum_node = [&](int ra, int rb, const Trip<T> (&v)[kPre], int a, const Trip<uint8_t> &f, const Trip<T> &d,
                        double &dv) {
        T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
        for (int u = 0; u < kPre; ++u)
            if (ra + u < rb) {
                s0 += v[u].a;
                s1 += v[u].b;
                s2 += v[u].c;
            }
        for (int r = ra + kPre; r < rb; ++r) {
            const Trip<T> w = ld3(bpart, r);
            s0 += w.a;
            s1 += w.b;
            s2 += w.c;
        }
        if (f.a) s0 = T(0);
        if (f.b) s1 = T(0);
        if (f.c) s2 = T(0);
        st3(y, a, s0, s1, s2);
        if (dotv) dv += (double)s0 * (double)d.a + (double)s1 * (double)d.b + (double)s2 * (double)d.c;
    };
    auto load_node = [&](int b, int &ra, int &rb, Trip<T> (&v)[kPre], int &a, Trip<uint8_t> &f, Trip<T> &d) {
        ra = bptr[b];
        rb = bptr[b + 1];
        a = bnode[b];
#pragma unroll
        for (int u = 0; u < kPre; ++u) v[u] = ra + u < rb ? ld3(bpart, ra + u) : Trip<T>{T(0), T(0), T(0)};
        f = fixed ? ld3(fixed, a) : Trip<uint8_t>{0, 0, 0};
        d = dotv ? ld3(dotv, a) : Trip<T>{T(0), T(0), T(0)};
    };
    int ra = 0, rb = 0, a = 0;
    Trip<T> v[kPre], d{T(0), T(0), T(0)};
    Trip<uint8_t> f{0, 0, 0};
    const int bf = b0 + (int)threadIdx.x;
    if (bf < b1) load_node(bf, ra, rb, v, a, f, d);
    if (done && *done) return;
    double dv = 0.0;
    if (ppart)
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npatch; q += gridDim.x * blockDim.x) dv += ppart[q];
    if (bf < b1) sum_node(ra, rb, v, a, f, d, dv);
    for (int b = bf + blockDim.x; b < b1; b += blockDim.x) {  // chunks longer than the block
        load_node(b, ra, rb, v, a, f, d);
        sum_node(ra, rb, v, a, f, d, dv);
    }
    if (part) {
        double v[1] = {dv};
        write_partials<1>(v, part);
    }
}

// geometric factors in patch order: dst[(r * n_el) + k] = src[r * n_el + perm[k]]
template <typename T>
__global__ __launch_bounds__(256) void k_permute_gf(int64_t rows, int n_el, const int32_t *__restrict__ perm,
                                                   const T *__restrict__ src, T *__restrict__ dst) {
    const int64_t total = rows * n_el;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / n_el, k = idx - r * n_el;
        dst[idx] = src[r * n_el + perm[k]];
    }
}

// sqrt(cm) invJ in FP32 from the FP64 factors (affine: one set per element),
// cm = 0.2 mu w detJ as the affine Tet10 product scales it
__global__ __launch_bounds__(256) void k_fold_gf(int n_el, const double *__restrict__ gf, float *__restrict__ gq) {
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < 9 * (int64_t)n_el;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / n_el, k = idx - r * n_el;
        gq[idx] = (float)(sqrt(0.2 * gf[10 * (int64_t)n_el + k]) * gf[idx]);
    }
}

// ------------------------------------------------------------ host setup ---
namespace {

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

}  // namespace

bool jfem_patch_enabled() {
    static const bool use = [] {
        const char *v = std::getenv("JFEM_PATCH");
        return !(v && v[0] == '0');
    }();
    return use;
}

void jfem_patch_release(jfm_ctx *ctx) {
    for (void **p : {(void **)&ctx->connP, (void **)&ctx->gfP64, (void **)&ctx->gfP32, (void **)&ctx->gfQ32, (void **)&ctx->pn_ptr,
                     (void **)&ctx->pn_cnt, (void **)&ctx->pn_out, (void **)&ctx->rstart, (void **)&ctx->rbase,
                     (void **)&ctx->linT, (void **)&ctx->bptr, (void **)&ctx->bnode,
                     (void **)&ctx->bpart, (void **)&ctx->ppart, (void **)&ctx->eperm, (void **)&ctx->qsrc,
                     (void **)&ctx->connP16, (void **)&ctx->pbase}) {
        jfm_free(ctx, *p);
        *p = nullptr;
    }
    ctx->patch_ready = false;
    ctx->npatch = ctx->npn = ctx->nbnode = ctx->nslot = ctx->bmaxcnt = 0;
    ctx->qsrc_gen = -1;
    if (ctx->ke_perm) {  // element matrices stored in the released patch order
        jfm_free(ctx, ctx->Ke);
        jfm_free(ctx, ctx->Ke32);
        ctx->Ke = nullptr;
        ctx->Ke32 = nullptr;
        ctx->ke_perm = nullptr;
    }
}

// The patch structure on the host: element order, patch-local incidences in
// rounds and the shared-node partial lists.  Returns false if an index does
// not fit its type.
struct PatchHost {
    int npatch = 0;
    std::vector<int32_t> perm, connP, pn_ptr, pn_out, rstart, rbase, bnode, bptr, pbase;
    std::vector<uint16_t> connP16;  // empty when a patch spans more than 65,536 node ids
    std::vector<uint16_t> pn_cnt, linT;
};

static bool build_patches(const std::vector<int32_t> &conn, const std::vector<double> &X, int64_t n_el, int64_t N,
                          int npe, PatchHost &P) {
    // Morton order of the element centroids (stable in element id)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t a = 0; a < N; ++a)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], X[3 * a + c]);
            hi[c] = std::max(hi[c], X[3 * a + c]);
        }
    std::vector<uint64_t> code(n_el);
    for (int64_t e = 0; e < n_el; ++e) {
        double cen[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < npe; ++i)
            for (int c = 0; c < 3; ++c) cen[c] += X[3 * (size_t)conn[e * npe + i] + c];
        uint64_t q[3];
        for (int c = 0; c < 3; ++c) {
            const double ext = hi[c] - lo[c];
            const double f = ext > 0.0 ? (cen[c] / npe - lo[c]) / ext : 0.0;
            q[c] = (uint64_t)std::min(std::max(f, 0.0) * 2097151.0, 2097151.0);
        }
        code[e] = spread3(q[0]) << 2 | spread3(q[1]) << 1 | spread3(q[2]);
    }
    std::vector<int32_t> &perm = P.perm;
    perm.resize(n_el);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    code.clear();
    code.shrink_to_fit();
    std::vector<int32_t> val(N, 0);  // node valences
    for (int32_t v : conn) ++val[v];
    const int npatch = (int)((n_el + kPatch - 1) / kPatch);
    P.npatch = npatch;
    P.connP.assign((size_t)npe * n_el, 0);
    P.pn_ptr.assign(npatch + 1, 0);
    P.rstart.assign(npatch + 1, 0);
    P.pn_out.clear();
    P.pn_cnt.clear();
    P.rbase.clear();
    P.linT.clear();
    std::vector<std::vector<int32_t>> shared_l(N);  // per node: patch-node indices, ascending patch
    std::vector<std::pair<int64_t, int32_t>> pairs;  // (node * kPatch + local element, i)
    struct PNode { int32_t a, first, cnt; };
    std::vector<PNode> nodes;
    for (int p = 0; p < npatch; ++p) {
        const int64_t k0 = (int64_t)p * kPatch, k1 = std::min<int64_t>(k0 + kPatch, n_el);
        pairs.clear();
        for (int64_t k = k0; k < k1; ++k) {
            const int32_t e = perm[k];
            for (int i = 0; i < npe; ++i) {
                const int32_t a = conn[(size_t)e * npe + i];
                P.connP[(size_t)i * n_el + k] = a;
                pairs.emplace_back((int64_t)a * kPatch + (k - k0), i);
            }
        }
        std::sort(pairs.begin(), pairs.end());  // by node, then local element
        nodes.clear();
        for (size_t r = 0; r < pairs.size();) {
            const int32_t a = (int32_t)(pairs[r].first / kPatch);
            size_t r2 = r;
            while (r2 < pairs.size() && pairs[r2].first / kPatch == a) ++r2;
            nodes.push_back({a, (int32_t)r, (int32_t)(r2 - r)});
            r = r2;
        }
        // most incidences first, so that a round's counts are close
        std::stable_sort(nodes.begin(), nodes.end(), [](const PNode &u, const PNode &v) { return u.cnt > v.cnt; });
        const int64_t l0 = (int64_t)P.pn_out.size();
        for (size_t u = 0; u < nodes.size(); u += kPatch) {
            const int cntmax = nodes[u].cnt;
            const int64_t base = (int64_t)P.linT.size();
            if (base > INT32_MAX) return false;
            P.rbase.push_back((int32_t)base);
            P.linT.resize(base + (size_t)cntmax * kPatch, 0);
            for (size_t t = 0; t < (size_t)kPatch && u + t < nodes.size(); ++t) {
                const PNode &nd = nodes[u + t];
                for (int j = 0; j < nd.cnt; ++j) {
                    const auto &pr = pairs[nd.first + j];
                    const int kl = (int)(pr.first % kPatch), i = pr.second;
                    P.linT[base + (size_t)j * kPatch + t] = (uint16_t)(3 * i * kPatch + kl);
                }
            }
        }
        for (size_t u = 0; u < nodes.size(); ++u) {
            const PNode &nd = nodes[u];
            if (nd.cnt > 65535) return false;
            P.pn_cnt.push_back((uint16_t)nd.cnt);
            if (nd.cnt == val[nd.a]) {
                P.pn_out.push_back(nd.a);  // interior
            } else {
                P.pn_out.push_back(-1);  // shared
                shared_l[nd.a].push_back((int32_t)(l0 + (int64_t)u));
            }
        }
        P.pn_ptr[p + 1] = (int32_t)P.pn_out.size();
        P.rstart[p + 1] = (int32_t)P.rbase.size();
    }
    // shared nodes in the order of their first patch entry, then the nodes
    // of no element
    std::vector<int32_t> order;
    for (int64_t a = 0; a < N; ++a)
        if (!shared_l[a].empty()) order.push_back((int32_t)a);
    std::sort(order.begin(), order.end(), [&](int32_t u, int32_t v) { return shared_l[u][0] < shared_l[v][0]; });
    for (int64_t a = 0; a < N; ++a)
        if (val[a] == 0) order.push_back((int32_t)a);
    P.bnode.clear();
    P.bptr.assign(1, 0);
    for (int32_t a : order) {
        P.bnode.push_back(a);
        // slots of node a: consecutive, in ascending patch order
        for (size_t j = 0; j < shared_l[a].size(); ++j) P.pn_out[shared_l[a][j]] = -(P.bptr.back() + (int32_t)j + 1);
        P.bptr.push_back(P.bptr.back() + (int32_t)shared_l[a].size());
    }
    // 16-bit connectivity relative to each patch's lowest node id, when every
    // patch's ids span at most 65,536
    P.pbase.assign(npatch, 0);
    bool fits16 = true;
    for (int p = 0; p < npatch && fits16; ++p) {
        const int64_t k0 = (int64_t)p * kPatch, k1 = std::min<int64_t>(k0 + kPatch, n_el);
        int32_t lo = INT32_MAX, hi = -1;
        for (int i = 0; i < npe; ++i)
            for (int64_t k = k0; k < k1; ++k) {
                lo = std::min(lo, P.connP[(size_t)i * n_el + k]);
                hi = std::max(hi, P.connP[(size_t)i * n_el + k]);
            }
        P.pbase[p] = lo;
        fits16 = hi - lo < 65536;
    }
    if (fits16) {
        P.connP16.assign((size_t)npe * n_el, 0);
        for (int i = 0; i < npe; ++i)
            for (int64_t k = 0; k < n_el; ++k)
                P.connP16[(size_t)i * n_el + k] = (uint16_t)(P.connP[(size_t)i * n_el + k] - P.pbase[k / kPatch]);
    }
    return 3 * (npe - 1) * kPatch + kPatch <= 65536 && (int64_t)P.pn_out.size() < INT32_MAX / 3;
}

// Builds the patch structure for the uploaded mesh (once per mesh).
int jfem_patch_setup(jfm_ctx *ctx) {
    if (ctx->patch_ready) return JFM_OK;
    const int64_t n_el = ctx->n_el, N = ctx->N;
    const int npe = ctx->npe;
    if (n_el == 0) return JFM_OK;
    if ((int64_t)ctx->h_conn.size() != n_el * npe || (int64_t)ctx->h_X.size() != 3 * N)
        return jfm_fail(ctx, JFM_ESTATE, "patch setup needs the host mesh");
    PatchHost P;
    if (!build_patches(ctx->h_conn, ctx->h_X, n_el, N, npe, P))
        return jfm_fail(ctx, JFM_EINVAL, "patch structure out of range");
    const int npatch = P.npatch;
    const int64_t npn = (int64_t)P.pn_out.size(), nb = (int64_t)P.bnode.size();
    hipStream_t s = ctx->stream;
    const int64_t ngfr = (int64_t)ctx->ngf * GF;  // rows of the geometric factors
#define PALLOC(ptr, T, cnt)                                                                 \
    do {                                                                                    \
        ptr = jfm_alloc<T>(ctx, std::max<int64_t>(1, (cnt)));                               \
        if (!ptr) {                                                                         \
            jfem_patch_release(ctx);                                                        \
            return jfm_fail(ctx, JFM_ENOMEM, "patch operator allocation failed");            \
        }                                                                                   \
    } while (0)
#define PUP(ptr, vec)                                                                       \
    if (!(vec).empty())                                                                     \
        JFM_HIP(ctx, hipMemcpyAsync(ptr, (vec).data(), sizeof((vec)[0]) * (vec).size(), hipMemcpyHostToDevice, s))
    PALLOC(ctx->connP, int32_t, npe * n_el);
    PALLOC(ctx->eperm, int32_t, n_el);
    PALLOC(ctx->gfP64, double, ngfr * n_el);
    PALLOC(ctx->gfP32, float, ngfr * n_el);
    PALLOC(ctx->pn_ptr, int32_t, npatch + 1);
    PALLOC(ctx->pn_cnt, uint16_t, npn);
    PALLOC(ctx->pn_out, int32_t, npn);
    PALLOC(ctx->rstart, int32_t, npatch + 1);
    PALLOC(ctx->rbase, int32_t, (int64_t)P.rbase.size());
    PALLOC(ctx->linT, uint16_t, (int64_t)P.linT.size());
    PALLOC(ctx->bptr, int32_t, nb + 1);
    PALLOC(ctx->bnode, int32_t, nb);
    PALLOC(ctx->bpart, double, 3 * (int64_t)P.bptr.back());
    PALLOC(ctx->ppart, double, npatch);
    if (!P.connP16.empty()) {
        PALLOC(ctx->connP16, uint16_t, npe * n_el);
        PALLOC(ctx->pbase, int32_t, npatch);
    }
#undef PALLOC
    PUP(ctx->connP, P.connP);
    PUP(ctx->eperm, P.perm);
    PUP(ctx->pn_ptr, P.pn_ptr);
    PUP(ctx->pn_cnt, P.pn_cnt);
    PUP(ctx->pn_out, P.pn_out);
    PUP(ctx->rstart, P.rstart);
    PUP(ctx->rbase, P.rbase);
    PUP(ctx->linT, P.linT);
    PUP(ctx->bptr, P.bptr);
    PUP(ctx->bnode, P.bnode);
    if (!P.connP16.empty()) {
        PUP(ctx->connP16, P.connP16);
        PUP(ctx->pbase, P.pbase);
    }
#undef PUP
    k_permute_gf<double><<<1024, 256, 0, s>>>(ngfr, (int)n_el, ctx->eperm, ctx->gf64, ctx->gfP64);
    k_permute_gf<float><<<1024, 256, 0, s>>>(ngfr, (int)n_el, ctx->eperm, ctx->gf32, ctx->gfP32);
    if (ctx->affine && npe == 10 && ctx->nu_uniform) {
        // FP32 folded factors: sqrt(cm) invJ, cm = mu w detJ / 5, and lambda / mu
        ctx->gfQ32 = jfm_alloc<float>(ctx, 9 * n_el);
        if (!ctx->gfQ32) {
            jfem_patch_release(ctx);
            return jfm_fail(ctx, JFM_ENOMEM, "patch operator allocation failed");
        }
        k_fold_gf<<<1024, 256, 0, s>>>((int)n_el, ctx->gfP64, ctx->gfQ32);
        ctx->rho32 = (float)(2.0 * ctx->nu0 / (1.0 - 2.0 * ctx->nu0));
    }
    JFM_HIP(ctx, hipGetLastError());
    JFM_HIP(ctx, hipStreamSynchronize(s));  // host vectors go out of scope
    ctx->npatch = npatch;
    ctx->npn = (int)npn;
    ctx->nbnode = (int)nb;
    ctx->nslot = P.bptr.back();
    ctx->bmaxcnt = 0;
    for (int64_t i = 0; i < nb; ++i) ctx->bmaxcnt = std::max(ctx->bmaxcnt, P.bptr[i + 1] - P.bptr[i]);
    ctx->patch_ready = true;
    return JFM_OK;
}

// qsrc for the update kernel of the two-level PCG (nodes in aggregate order)
__global__ __launch_bounds__(256) void k_slot_fill(int N, int32_t *__restrict__ code) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) code[i] = -1;
}

__global__ __launch_bounds__(256) void k_slot_shared(int nb, const int32_t *__restrict__ bptr,
                                                    const int32_t *__restrict__ bnode, int32_t *__restrict__ code) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) code[bnode[b]] = bptr[b] << 5 | (bptr[b + 1] - bptr[b]);
}

__global__ __launch_bounds__(256) void k_slot_gather(int N, const int32_t *__restrict__ agg_nodes,
                                                    const int32_t *__restrict__ code, int32_t *__restrict__ qsrc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) qsrc[k] = code[agg_nodes[k]];
}

bool jfem_patch_defer_shared(const jfm_ctx *ctx, int op) {
    static const bool on = [] {
        const char *v = std::getenv("JFEM_DEFER_SHARED");
        return !(v && v[0] == '0');
    }();
    // the patch kernel's dot must be the element energies (nothing left for
    // the shared nodes), its partials final (one per patch)
    return on && op == JFM_OP_MATFREE && jfem_patch_form(ctx, op) && ctx->affine && ctx->npe == 10 &&
           ctx->npatch <= kMaxPart && ctx->qsrc && ctx->qsrc_gen == ctx->coarse_gen;
}

int jfem_patch_qsrc(jfm_ctx *ctx) {
    if (!ctx->patch_ready || !ctx->agg_nodes || ctx->n_agg == 0) return JFM_OK;
    if (ctx->qsrc && ctx->qsrc_gen == ctx->coarse_gen) return JFM_OK;
    ctx->qsrc_gen = -1;
    // slot codes hold 26 bits of slot index and 5 of count
    if (ctx->bmaxcnt > 31 || ctx->nslot >= (1 << 26)) return JFM_OK;
    const int N = (int)ctx->N;
    if (!ctx->qsrc) ctx->qsrc = jfm_alloc<int32_t>(ctx, std::max(1, N));
    int32_t *code = jfm_alloc<int32_t>(ctx, std::max(1, N));
    if (!ctx->qsrc || !code) {
        jfm_free(ctx, code);
        return jfm_fail(ctx, JFM_ENOMEM, "shared-node code allocation failed");
    }
    hipStream_t s = ctx->stream;
    const unsigned gN = (unsigned)((N + 255) / 256), gB = (unsigned)((ctx->nbnode + 255) / 256);
    k_slot_fill<<<std::max(1u, gN), 256, 0, s>>>(N, code);
    if (ctx->nbnode > 0) k_slot_shared<<<gB, 256, 0, s>>>(ctx->nbnode, ctx->bptr, ctx->bnode, code);
    k_slot_gather<<<std::max(1u, gN), 256, 0, s>>>(N, ctx->agg_nodes, code, ctx->qsrc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    jfm_free(ctx, code);
    if (e != hipSuccess) return jfm_fail(ctx, JFM_EHIP, std::string("shared-node codes: ") + hipGetErrorString(e));
    ctx->qsrc_gen = ctx->coarse_gen;
    return JFM_OK;
}

// ------------------------------------------------------------- launchers ---
template <typename T> static const T *gfP_of(jfm_ctx *ctx);
template <> const double *gfP_of<double>(jfm_ctx *ctx) { return ctx->gfP64; }
template <> const float *gfP_of<float>(jfm_ctx *ctx) { return ctx->gfP32; }
template <typename T> static const T *keP_of(jfm_ctx *ctx);
template <> const double *keP_of<double>(jfm_ctx *ctx) { return ctx->Ke; }
template <> const float *keP_of<float>(jfm_ctx *ctx) { return ctx->Ke32; }

// Whether the patch kernel forms x.(Kx) from the element energies (no dotv
// reads): the affine Tet10 matrix-free form, with dotv = x, which the PCG
// passes (its directions are zero on fixed DOFs).
// (headline at config 1: 16.10-16.14 ms against 16.26-16.27 ms with the dot
// over the nodes)
template <typename T> static bool energy_dot(const jfm_ctx *ctx, int op, const T *x, const T *dotv) {
    return op == JFM_OP_MATFREE && ctx->affine && ctx->npe == 10 && dotv && dotv == x;
}

// The block partials of the dot go to pout (ctx->ppart when null).
template <int NPE, typename T>
static void patch_kernel(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, const int *done, double *pout) {
    const dim3 grid(xcd_grid(ctx->npatch));
    const int n = (int)ctx->n_el;
    double *pp = dotv ? (pout ? pout : ctx->ppart) : nullptr;
    if (energy_dot<T>(ctx, op, x, dotv)) dotv = nullptr;
    auto launch = [&](auto kernel, const T *el, T rho) {
        kernel<<<grid, kPatch, 0, ctx->stream>>>(n, ctx->npatch, ctx->npn, ctx->connP, ctx->connP16, ctx->pbase, el, x,
                                                 ctx->pn_ptr, ctx->pn_cnt,
                                                 ctx->pn_out, ctx->rstart, ctx->rbase, ctx->linT, ctx->fixed, y,
                                                 (T *)ctx->bpart, dotv, pp, done, rho);
    };
    const bool c16 = ctx->connP16 != nullptr;
    if (op == JFM_OP_EBE) {
        launch(k_ku_patch<NPE, T, true, true, false, false>, keP_of<T>(ctx), T(0));
    } else if (ctx->affine) {
        if constexpr (sizeof(T) == 4 && NPE == 10) {
            if (ctx->gfQ32) {  // folded factors (uniform nu)
                launch(c16 ? k_ku_patch<NPE, T, true, false, true, true> : k_ku_patch<NPE, T, true, false, false, true>,
                       (const T *)ctx->gfQ32, (T)ctx->rho32);
                return;
            }
        }
        launch(c16 ? k_ku_patch<NPE, T, true, false, true, false> : k_ku_patch<NPE, T, true, false, false, false>,
               gfP_of<T>(ctx), T(0));
    } else {
        launch(c16 ? k_ku_patch<NPE, T, false, false, true, false> : k_ku_patch<NPE, T, false, false, false, false>,
               gfP_of<T>(ctx), T(0));
    }
}

template <typename T>
hipError_t launch_patch_phase(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, const int *done,
                              double *pout) {
    if (ctx->n_el == 0) return hipSuccess;
    if (!jfem_patch_form(ctx, op)) return hipErrorInvalidValue;
    switch (ctx->npe) {
    case 10: patch_kernel<10, T>(ctx, op, x, y, dotv, done, pout); break;
    case 4: patch_kernel<4, T>(ctx, op, x, y, dotv, done, pout); break;
    case 8: patch_kernel<8, T>(ctx, op, x, y, dotv, done, pout); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_patch_combine(jfm_ctx *ctx, T *y, const T *dotv, double *part, const int *done, int *npart) {
    // a multiple of 8 blocks (XCD mapping), one shared node per thread up to
    // kMaxPart blocks (544 at 1.05M DOFs: a second node per thread would
    // double the dependent-load chain of the kernel)
    const int nb = ctx->nbnode;
    const unsigned g = std::min<unsigned>(xcd_grid((unsigned)((std::max(nb, 1) + kBlock - 1) / kBlock)), kMaxPart);
    const int chunk = (int)((nb + g - 1) / g);
    if (npart) *npart = (int)g;
    k_ku_combine<T><<<g, kBlock, 0, ctx->stream>>>(nb, chunk, ctx->bptr, ctx->bnode, (const T *)ctx->bpart,
                                                   ctx->fixed, y, dotv, (ctx->n_el && part) ? ctx->ppart : nullptr,
                                                   ctx->n_el ? ctx->npatch : 0, part, done);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_patch_apply(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, double *part, const int *done,
                              int *npart) {
    const bool energy = energy_dot<T>(ctx, op, x, dotv);
    if (energy && part && ctx->npatch <= kMaxPart) {
        // the patches' energy partials are the final ones: straight to part,
        // and the shared-node kernel only sums nodes
        hipError_t e = launch_patch_phase<T>(ctx, op, x, y, dotv, done, part);
        if (e != hipSuccess) return e;
        if (npart) *npart = ctx->npatch;
        return launch_patch_combine<T>(ctx, y, nullptr, nullptr, done, nullptr);
    }
    hipError_t e = launch_patch_phase<T>(ctx, op, x, y, dotv, done, nullptr);
    if (e != hipSuccess) return e;
    // with the element energies the shared nodes add nothing to the dot
    return launch_patch_combine<T>(ctx, y, energy ? nullptr : dotv, part, done, npart);
}

template hipError_t launch_patch_phase<double>(jfm_ctx *, int, const double *, double *, const double *, const int *, double *);
template hipError_t launch_patch_phase<float>(jfm_ctx *, int, const float *, float *, const float *, const int *, double *);
template hipError_t launch_patch_combine<double>(jfm_ctx *, double *, const double *, double *, const int *, int *);
template hipError_t launch_patch_combine<float>(jfm_ctx *, float *, const float *, double *, const int *, int *);
template hipError_t launch_patch_apply<double>(jfm_ctx *, int, const double *, double *, const double *, double *, const int *, int *);
template hipError_t launch_patch_apply<float>(jfm_ctx *, int, const float *, float *, const float *, double *, const int *, int *);

}  // namespace jfem
