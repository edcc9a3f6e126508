// api.hip -- C ABI of libjfem_hip.so: context, problem upload, assembly,
// operator and solvers.  Declarations and the reference entry point each
// function replaces are in include/jfem.h.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "jfem_device.h"
#include "jfem_learned.h"

using namespace jfem;

// ================================================================ helpers ==
int jfm_fail(jfm_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int jfm_hip_fail(jfm_ctx *ctx, hipError_t e, const char *what) {
    return jfm_fail(ctx, JFM_EHIP, std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

template <typename T> T *jfm_alloc(jfm_ctx *ctx, int64_t count) {
    if (count <= 0) count = 1;
    void *p = nullptr;
    const size_t bytes = (size_t)count * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    ctx->allocs.push_back(p);
    ctx->bytes += (int64_t)bytes;
    return (T *)p;
}
template uint16_t *jfm_alloc<uint16_t>(jfm_ctx *, int64_t);

void jfm_free(jfm_ctx *ctx, void *p) {
    if (!p) return;
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    size_t sz = 0;
    (void)hipMemPtrGetInfo(p, &sz);
    ctx->bytes -= (int64_t)sz;
    (void)hipFree(p);
}

template <typename T> static void free_ptr(jfm_ctx *ctx, T *&p) {
    jfm_free(ctx, (void *)p);
    p = nullptr;
}

#define ALLOC(ctx, ptr, T, n)                                                          \
    do {                                                                               \
        (ptr) = jfm_alloc<T>((ctx), (n));                                              \
        if (!(ptr)) return jfm_fail((ctx), JFM_ENOMEM, "device allocation failed: " #ptr); \
    } while (0)

#define CHECK_LAUNCH(ctx, call)                           \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return jfm_hip_fail(ctx, e_, #call); \
    } while (0)

template <typename F> static void parallel_for(int64_t n, F f) {
    int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

static void drop_graphs(jfm_ctx *ctx);

static int sync(jfm_ctx *ctx) {
    JFM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFM_OK;
}

// ================================================================ context ==
extern "C" int jfm_create(int device, jfm_ctx **out) {
    if (!out) return JFM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return JFM_EHIP;
    if (device < 0 || device >= ndev) return JFM_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return JFM_EHIP;
    jfm_ctx *ctx = new jfm_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return JFM_EHIP;
    }
    ctx->S = jfm_alloc<PcgScalars>(ctx, 1);
    if (hipHostMalloc((void **)&ctx->S_host, sizeof(PcgScalars), hipHostMallocDefault) != hipSuccess || !ctx->S) {
        jfm_destroy(ctx);
        return JFM_ENOMEM;
    }
    std::memset(ctx->S_host, 0, sizeof(PcgScalars));
    if (int rc = jfem_noalloc_init(ctx)) {  // device address of the no-allocation entry's scratch
        jfm_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return JFM_OK;
}

extern "C" void jfm_destroy(jfm_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    drop_graphs(ctx);
    for (void *p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    if (ctx->S_host) (void)hipHostFree(ctx->S_host);
    if (ctx->stream2) {
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream2);
    }
    for (hipEvent_t ev : {ctx->ev_fork, ctx->ev_l0, ctx->ev_l1})
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *jfm_last_error(const jfm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int jfm_get_info(jfm_ctx *ctx, jfm_info *info) {
    if (!ctx || !info) return JFM_EINVAL;
    info->n_nodes = ctx->N;
    info->n_el = ctx->n_el;
    info->n_dofs = ctx->ndof;
    info->nnzb = ctx->nnzb;
    info->nnz = 9 * ctx->nnzb;
    info->n_faces = ctx->n_faces;
    info->npe = ctx->npe;
    info->affine = ctx->affine ? 1 : 0;
    info->device_bytes = ctx->bytes;
    info->n_patches = ctx->patch_ready ? ctx->npatch : 0;
    info->patch_nodes = ctx->patch_ready ? ctx->npn : 0;
    info->shared_nodes = ctx->patch_ready ? ctx->nbnode : 0;
    info->shared_slots = ctx->patch_ready ? ctx->nslot : 0;
    info->patch_conn_bytes = ctx->patch_ready ? (ctx->connP16 ? 2 : 4) : 0;
    info->patch_fp32_factor_rows = ctx->patch_ready ? (ctx->gfQ32 ? 9 : GF) : 0;
    return JFM_OK;
}

// ============================================================ mesh upload ==
static void drop_graphs(jfm_ctx *ctx) {
    for (auto &g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
}

static void release_heat(jfm_ctx *ctx) {
    free_ptr(ctx, ctx->heat_k); free_ptr(ctx, ctx->heat_fixed); free_ptr(ctx, ctx->heat_g); free_ptr(ctx, ctx->heat_Minv);
    free_ptr(ctx, ctx->hf_conn); free_ptr(ctx, ctx->hf_ptr); free_ptr(ctx, ctx->hf_inc); free_ptr(ctx, ctx->hf_h);
    free_ptr(ctx, ctx->hf_g); free_ptr(ctx, ctx->hf_text); free_ptr(ctx, ctx->hf_fe);
    ctx->hf_n = 0;
    ctx->hf_nfn = 0;
}

static void release_mesh(jfm_ctx *ctx) {
    jfem_learned_drop(ctx);  // no stream2 work may still use the buffers freed here
    free_ptr(ctx, ctx->X); free_ptr(ctx, ctx->conn); free_ptr(ctx, ctx->E); free_ptr(ctx, ctx->nu);
    free_ptr(ctx, ctx->gf64); free_ptr(ctx, ctx->gf32); free_ptr(ctx, ctx->inc_ptr); free_ptr(ctx, ctx->inc);
    free_ptr(ctx, ctx->brow); free_ptr(ctx, ctx->bcol); free_ptr(ctx, ctx->bdiag); free_ptr(ctx, ctx->blk_row);
    free_ptr(ctx, ctx->cptr); free_ptr(ctx, ctx->contrib); free_ptr(ctx, ctx->vals64); free_ptr(ctx, ctx->vals32);
    free_ptr(ctx, ctx->Ke); free_ptr(ctx, ctx->Ke32); free_ptr(ctx, ctx->ye64); free_ptr(ctx, ctx->ye32);
    free_ptr(ctx, ctx->vx64); free_ptr(ctx, ctx->vy64); free_ptr(ctx, ctx->vx32); free_ptr(ctx, ctx->vy32);
    free_ptr(ctx, ctx->fixed); free_ptr(ctx, ctx->g); free_ptr(ctx, ctx->face_conn); free_ptr(ctx, ctx->face_t);
    free_ptr(ctx, ctx->finc_ptr); free_ptr(ctx, ctx->finc); free_ptr(ctx, ctx->fe); free_ptr(ctx, ctx->f_ext);
    free_ptr(ctx, ctx->load); free_ptr(ctx, ctx->u64); free_ptr(ctx, ctx->rr64);
    free_ptr(ctx, ctx->partials); free_ptr(ctx, ctx->counter); free_ptr(ctx, ctx->inc_slot);
    drop_graphs(ctx);
    for (Workspace *w : {&ctx->ws64, &ctx->ws32}) {
        free_ptr(ctx, w->x); free_ptr(ctx, w->r); free_ptr(ctx, w->z); free_ptr(ctx, w->p);
        free_ptr(ctx, w->q); free_ptr(ctx, w->b); free_ptr(ctx, w->Minv);
        w->ok = false;
    }
    free_ptr(ctx, ctx->Mlearn64); free_ptr(ctx, ctx->Mlearn32); free_ptr(ctx, ctx->feat); free_ptr(ctx, ctx->efeat); free_ptr(ctx, ctx->lout);
    free_ptr(ctx, ctx->Muser64); free_ptr(ctx, ctx->Muser32); free_ptr(ctx, ctx->xprobe);
    free_ptr(ctx, ctx->ustate64); free_ptr(ctx, ctx->ustate32); free_ptr(ctx, ctx->Dtan);
    free_ptr(ctx, ctx->nl_u); free_ptr(ctx, ctx->nl_r); free_ptr(ctx, ctx->nl_du); free_ptr(ctx, ctx->nl_zero);
    ctx->has_state = ctx->minv_tangent = false;
    jfem_coarse_release(ctx);
    jfem_patch_release(ctx);
    release_heat(ctx);
    ctx->h_X.clear();
    ctx->has_user = false;
    ctx->has_mesh = ctx->has_bcs = ctx->assembled = ctx->has_load = false;
    ctx->learned_ready = false;
    ctx->h_conn.clear(); ctx->h_brow.clear(); ctx->h_bcol.clear();
}

// Tet10 edge mid nodes at their edge midpoints (and Hex8 parallelepipeds) make
// J constant over the element; one factor set per element then suffices.
static bool detect_affine(const std::vector<int32_t> &conn, const double *X, int64_t n_el, int npe) {
    if (npe == 4) return true;
    std::atomic<bool> ok{true};
    parallel_for(n_el, [&](int64_t a, int64_t b) {
        for (int64_t e = a; e < b && ok.load(std::memory_order_relaxed); ++e) {
            const int32_t *c = conn.data() + e * npe;
            auto P = [&](int i, int k) { return X[3 * (int64_t)c[i] + k]; };
            double h = 0;
            for (int k = 0; k < 3; ++k) h = std::max(h, std::fabs(P(1, k) - P(0, k)));
            const double tol = 1e-12 * std::max(h, 1e-300);
            if (npe == 10) {
                static const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
                for (int m = 0; m < 6; ++m)
                    for (int k = 0; k < 3; ++k)
                        if (std::fabs(P(4 + m, k) - 0.5 * (P(ed[m][0], k) + P(ed[m][1], k))) > tol) ok = false;
            } else {  // Hex8: x_i = x0 + sum of the three edge vectors it spans
                static const int s[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
                for (int i = 0; i < 8; ++i)
                    for (int k = 0; k < 3; ++k) {
                        const double v = P(0, k) + s[i][0] * (P(1, k) - P(0, k)) + s[i][1] * (P(3, k) - P(0, k)) +
                                         s[i][2] * (P(4, k) - P(0, k));
                        if (std::fabs(P(i, k) - v) > tol * 8) ok = false;
                    }
            }
        }
    });
    return ok.load();
}

extern "C" int jfm_upload_mesh(jfm_ctx *ctx, int64_t n_nodes, const double *X, int64_t n_el, int32_t npe,
                               const int64_t *conn, const double *E, const double *nu, int32_t index_base) {
    if (!ctx) return JFM_EINVAL;
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (npe != 10 && npe != 4 && npe != 8)
        return jfm_fail(ctx, JFM_EUNSUPPORTED, "unsupported element: nodes_per_element must be 10 (Tet10), 4 (Tet4) or 8 (Hex8)");
    if (n_nodes <= 0 || n_el < 0 || !X || (n_el > 0 && (!conn || !E || !nu)))
        return jfm_fail(ctx, JFM_EINVAL, "invalid mesh arguments");
    if (index_base != 0 && index_base != 1) return jfm_fail(ctx, JFM_EINVAL, "index_base must be 0 or 1");
    if (n_nodes > (int64_t)INT32_MAX / 3 || n_el * (int64_t)npe * npe > (int64_t)INT32_MAX)
        return jfm_fail(ctx, JFM_EINVAL, "mesh too large for 32-bit indices");
    release_mesh(ctx);
    ctx->N = n_nodes;
    ctx->n_el = n_el;
    ctx->ndof = 3 * n_nodes;
    ctx->npe = npe;
    ctx->ngp = npe == 10 ? 4 : (npe == 4 ? 1 : 8);
    ctx->npair = npe * (npe + 1) / 2;
    ctx->h_X.assign(X, X + 3 * n_nodes);

    // connectivity -> 0-based int32, range checked
    ctx->h_conn.resize((size_t)n_el * npe);
    std::atomic<bool> bad{false};
    parallel_for(n_el * npe, [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) {
            const int64_t v = conn[k] - index_base;
            if (v < 0 || v >= n_nodes) bad = true;
            ctx->h_conn[k] = (int32_t)v;
        }
    });
    if (bad) return jfm_fail(ctx, JFM_EINVAL, "connectivity references a node outside [base, base+n_nodes)");
    for (int64_t e = 0; e < n_el; ++e)
        if (!(E[e] > 0.0) || !(nu[e] > -1.0 && nu[e] < 0.5))
            return jfm_fail(ctx, JFM_EINVAL, "material out of range (E > 0, -1 < nu < 0.5 required)");
    ctx->nu_uniform = n_el > 0;
    for (int64_t e = 1; e < n_el && ctx->nu_uniform; ++e) ctx->nu_uniform = nu[e] == nu[0];
    ctx->nu0 = n_el > 0 ? nu[0] : 0.0;

    const std::vector<int32_t> &hc = ctx->h_conn;
    // node -> incidences (e*npe + i), ascending e  (ext/JuliaFEMCUDAExt.jl:179-193)
    std::vector<int32_t> inc_ptr(n_nodes + 1, 0), inc((size_t)n_el * npe);
    for (int64_t k = 0; k < n_el * npe; ++k) inc_ptr[hc[k] + 1]++;
    for (int64_t a = 0; a < n_nodes; ++a) inc_ptr[a + 1] += inc_ptr[a];
    {
        std::vector<int32_t> pos(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int64_t e = 0; e < n_el; ++e)
            for (int i = 0; i < npe; ++i) inc[pos[hc[e * npe + i]]++] = (int32_t)(e * npe + i);
    }

    // 3x3-block pattern = node adjacency, columns sorted (parallel over rows)
    const int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<int32_t> rowlen(n_nodes);
    std::vector<std::vector<int32_t>> parts(nt);
    {
        const int64_t chunk = (n_nodes + nt - 1) / nt;
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) {
            th.emplace_back([&, t] {
                const int64_t a0 = t * chunk, a1 = std::min<int64_t>(n_nodes, a0 + chunk);
                std::vector<int32_t> tmp;
                auto &out = parts[t];
                for (int64_t a = a0; a < a1; ++a) {
                    tmp.clear();
                    for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
                        const int e = inc[k] / npe;
                        for (int j = 0; j < npe; ++j) tmp.push_back(hc[(int64_t)e * npe + j]);
                    }
                    std::sort(tmp.begin(), tmp.end());
                    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                    rowlen[a] = (int32_t)tmp.size();
                    out.insert(out.end(), tmp.begin(), tmp.end());
                }
            });
        }
        for (auto &x : th) x.join();
    }
    ctx->h_brow.assign(n_nodes + 1, 0);
    for (int64_t a = 0; a < n_nodes; ++a) {
        if ((int64_t)ctx->h_brow[a] + rowlen[a] > INT32_MAX)
            return jfm_fail(ctx, JFM_EINVAL, "too many nonzero blocks for 32-bit offsets");
        ctx->h_brow[a + 1] = ctx->h_brow[a] + rowlen[a];
    }
    ctx->nnzb = ctx->h_brow[n_nodes];
    ctx->h_bcol.resize(ctx->nnzb);
    {
        size_t off = 0;
        for (int t = 0; t < nt; ++t) {
            std::copy(parts[t].begin(), parts[t].end(), ctx->h_bcol.begin() + off);
            off += parts[t].size();
            std::vector<int32_t>().swap(parts[t]);
        }
    }
    const std::vector<int32_t> &brow = ctx->h_brow, &bcol = ctx->h_bcol;
    const int64_t nnzb = ctx->nnzb;

    // per-block contribution lists (e, i, j), ascending e
    std::vector<int32_t> cptr(nnzb + 1, 0);
    auto block_of = [&](int64_t a, int32_t b) -> int64_t {
        const int32_t *lo = bcol.data() + brow[a], *hi = bcol.data() + brow[a + 1];
        return (std::lower_bound(lo, hi, b) - bcol.data());
    };
    parallel_for(n_nodes, [&](int64_t a0, int64_t a1) {
        for (int64_t a = a0; a < a1; ++a)
            for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
                const int e = inc[k] / npe;
                for (int j = 0; j < npe; ++j) cptr[block_of(a, hc[(int64_t)e * npe + j]) + 1]++;
            }
    });
    for (int64_t s = 0; s < nnzb; ++s) cptr[s + 1] += cptr[s];
    std::vector<int32_t> contrib(cptr[nnzb]), blk_row(nnzb), bdiag(n_nodes);
    parallel_for(n_nodes, [&](int64_t a0, int64_t a1) {
        std::vector<int32_t> cur;
        for (int64_t a = a0; a < a1; ++a) {
            cur.assign(cptr.begin() + brow[a], cptr.begin() + brow[a + 1]);
            for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
                const int e = inc[k] / npe, i = inc[k] - e * npe;
                for (int j = 0; j < npe; ++j) {
                    const int64_t s = block_of(a, hc[(int64_t)e * npe + j]);
                    contrib[cur[s - brow[a]]++] = (int32_t)(((int64_t)e * npe + i) * npe + j);
                }
            }
            for (int32_t s = brow[a]; s < brow[a + 1]; ++s) blk_row[s] = (int32_t)a;
            bdiag[a] = (int32_t)block_of(a, (int32_t)a);
        }
    });

    ctx->affine = detect_affine(hc, X, n_el, npe);
    ctx->ngf = ctx->affine ? 1 : ctx->ngp;

    // device copies
    ALLOC(ctx, ctx->X, double, 3 * n_nodes);
    ALLOC(ctx, ctx->conn, int32_t, n_el * npe);
    ALLOC(ctx, ctx->E, double, n_el);
    ALLOC(ctx, ctx->nu, double, n_el);
    ALLOC(ctx, ctx->gf64, double, (int64_t)GF * ctx->ngf * n_el);
    ALLOC(ctx, ctx->gf32, float, (int64_t)GF * ctx->ngf * n_el);
    ALLOC(ctx, ctx->inc_ptr, int32_t, n_nodes + 1);
    ALLOC(ctx, ctx->inc, int32_t, n_el * npe);
    ALLOC(ctx, ctx->brow, int32_t, n_nodes + 1);
    ALLOC(ctx, ctx->bcol, int32_t, nnzb);
    ALLOC(ctx, ctx->bdiag, int32_t, n_nodes);
    ALLOC(ctx, ctx->blk_row, int32_t, nnzb);
    ALLOC(ctx, ctx->cptr, int32_t, nnzb + 1);
    ALLOC(ctx, ctx->contrib, int32_t, (int64_t)contrib.size());
    ALLOC(ctx, ctx->ye64, double, n_el * npe * 3);
    ALLOC(ctx, ctx->ye32, float, n_el * npe * 3);
    ALLOC(ctx, ctx->fixed, uint8_t, 3 * n_nodes);
    ALLOC(ctx, ctx->g, double, 3 * n_nodes);
    ALLOC(ctx, ctx->f_ext, double, 3 * n_nodes);
    ALLOC(ctx, ctx->load, double, 3 * n_nodes);
    ALLOC(ctx, ctx->finc_ptr, int32_t, n_nodes + 1);
    ALLOC(ctx, ctx->vx64, double, 3 * n_nodes);
    ALLOC(ctx, ctx->vy64, double, 3 * n_nodes);
    ALLOC(ctx, ctx->vx32, float, 3 * n_nodes);
    ALLOC(ctx, ctx->vy32, float, 3 * n_nodes);
    ALLOC(ctx, ctx->partials, double, kPartTotal);
    ALLOC(ctx, ctx->counter, unsigned, 64);
    ALLOC(ctx, ctx->inc_slot, int32_t, n_el * npe);
    // slot of incidence (e, i) in the node-sorted list, SoA by local index
    std::vector<int32_t> slot_soa((size_t)n_el * npe);
    parallel_for(n_nodes, [&](int64_t a0, int64_t a1) {
        for (int64_t a = a0; a < a1; ++a)
            for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
                const int e = inc[k] / npe, i = inc[k] - e * npe;
                slot_soa[(size_t)i * n_el + e] = k;
            }
    });

    std::vector<int32_t> conn_soa((size_t)n_el * npe);
    parallel_for(n_el, [&](int64_t a, int64_t b) {
        for (int64_t e = a; e < b; ++e)
            for (int i = 0; i < npe; ++i) conn_soa[(size_t)i * n_el + e] = hc[(size_t)e * npe + i];
    });
    hipStream_t s = ctx->stream;
    JFM_HIP(ctx, hipMemcpyAsync(ctx->X, X, sizeof(double) * 3 * n_nodes, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->conn, conn_soa.data(), sizeof(int32_t) * n_el * npe, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->E, E, sizeof(double) * n_el, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->nu, nu, sizeof(double) * n_el, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->inc_ptr, inc_ptr.data(), sizeof(int32_t) * (n_nodes + 1), hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->inc, inc.data(), sizeof(int32_t) * n_el * npe, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->brow, brow.data(), sizeof(int32_t) * (n_nodes + 1), hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->bcol, bcol.data(), sizeof(int32_t) * nnzb, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->bdiag, bdiag.data(), sizeof(int32_t) * n_nodes, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->blk_row, blk_row.data(), sizeof(int32_t) * nnzb, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->cptr, cptr.data(), sizeof(int32_t) * (nnzb + 1), hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->contrib, contrib.data(), sizeof(int32_t) * contrib.size(), hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->inc_slot, slot_soa.data(), sizeof(int32_t) * slot_soa.size(), hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemsetAsync(ctx->fixed, 0, 3 * n_nodes, s));
    JFM_HIP(ctx, hipMemsetAsync(ctx->g, 0, sizeof(double) * 3 * n_nodes, s));
    JFM_HIP(ctx, hipMemsetAsync(ctx->f_ext, 0, sizeof(double) * 3 * n_nodes, s));
    JFM_HIP(ctx, hipMemsetAsync(ctx->load, 0, sizeof(double) * 3 * n_nodes, s));
    JFM_HIP(ctx, hipMemsetAsync(ctx->finc_ptr, 0, sizeof(int32_t) * (n_nodes + 1), s));
    CHECK_LAUNCH(ctx, launch_zero_u32(ctx, ctx->counter, 64));
    int *d_bad = (int *)(ctx->counter + 32);
    CHECK_LAUNCH(ctx, launch_geom_factors(ctx, d_bad));
    int h_bad = 0;
    JFM_HIP(ctx, hipMemcpyAsync(&h_bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    if (int rc = sync(ctx)) return rc;
    // synchronous copies above read pageable host buffers that go out of scope here
    if (h_bad) return jfm_fail(ctx, JFM_ENOTSPD, "element with detJ <= 0 at a Gauss point (inverted or degenerate)");
    ctx->has_mesh = true;
    return JFM_OK;
}

// ============================================================ BCs, loads ===
extern "C" int jfm_upload_bcs(jfm_ctx *ctx, const uint8_t *fixed, const double *u_prescribed, int64_t n_faces,
                              int32_t nodes_per_face, const int64_t *face_conn, const double *traction,
                              int32_t index_base) {
    if (!ctx) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "upload the mesh before boundary conditions");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (n_faces < 0 || (n_faces > 0 && (!face_conn || !traction)))
        return jfm_fail(ctx, JFM_EINVAL, "invalid face arguments");
    if (n_faces > 0 && nodes_per_face != 6 && nodes_per_face != 3 && nodes_per_face != 4)
        return jfm_fail(ctx, JFM_EUNSUPPORTED, "nodes_per_face must be 6 (Tri6), 3 (Tri3) or 4 (Quad4)");
    const int64_t N = ctx->N, n = ctx->ndof;
    hipStream_t s = ctx->stream;
    if (fixed) JFM_HIP(ctx, hipMemcpyAsync(ctx->fixed, fixed, n, hipMemcpyHostToDevice, s));
    else JFM_HIP(ctx, hipMemsetAsync(ctx->fixed, 0, n, s));
    ctx->g_nonzero = false;
    if (u_prescribed) {
        for (int64_t i = 0; i < n; ++i)
            if (u_prescribed[i] != 0.0 && fixed && fixed[i]) { ctx->g_nonzero = true; break; }
    }
    if (ctx->g_nonzero) {
        std::vector<double> gh(n, 0.0);
        for (int64_t i = 0; i < n; ++i) gh[i] = fixed[i] ? u_prescribed[i] : 0.0;
        JFM_HIP(ctx, hipMemcpyAsync(ctx->g, gh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (int rc = sync(ctx)) return rc;
    } else {
        JFM_HIP(ctx, hipMemsetAsync(ctx->g, 0, sizeof(double) * n, s));
    }
    free_ptr(ctx, ctx->face_conn); free_ptr(ctx, ctx->face_t); free_ptr(ctx, ctx->finc); free_ptr(ctx, ctx->fe);
    ctx->n_faces = n_faces;
    ctx->nfn = nodes_per_face;
    std::vector<int32_t> fptr(N + 1, 0), finc, fsoa;
    std::vector<double> ft;
    if (n_faces > 0) {
        const int nfn = nodes_per_face;
        fsoa.resize((size_t)n_faces * nfn);
        for (int64_t f = 0; f < n_faces; ++f)
            for (int i = 0; i < nfn; ++i) {
                const int64_t v = face_conn[f * nfn + i] - index_base;
                if (v < 0 || v >= N) return jfm_fail(ctx, JFM_EINVAL, "face connectivity out of range");
                fsoa[(size_t)i * n_faces + f] = (int32_t)v;
                fptr[v + 1]++;
            }
        for (int64_t a = 0; a < N; ++a) fptr[a + 1] += fptr[a];
        finc.resize((size_t)n_faces * nfn);
        std::vector<int32_t> pos(fptr.begin(), fptr.end() - 1);
        for (int64_t f = 0; f < n_faces; ++f)
            for (int i = 0; i < nfn; ++i) finc[pos[fsoa[(size_t)i * n_faces + f]]++] = (int32_t)(f * nfn + i);
        ft.assign(traction, traction + 3 * n_faces);
        ALLOC(ctx, ctx->face_conn, int32_t, n_faces * nfn);
        ALLOC(ctx, ctx->face_t, double, 3 * n_faces);
        ALLOC(ctx, ctx->finc, int32_t, n_faces * nfn);
        ALLOC(ctx, ctx->fe, double, n_faces * nfn * 3);
        JFM_HIP(ctx, hipMemcpyAsync(ctx->face_conn, fsoa.data(), sizeof(int32_t) * fsoa.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->face_t, ft.data(), sizeof(double) * ft.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->finc, finc.data(), sizeof(int32_t) * finc.size(), hipMemcpyHostToDevice, s));
    }
    JFM_HIP(ctx, hipMemcpyAsync(ctx->finc_ptr, fptr.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice, s));
    CHECK_LAUNCH(ctx, launch_face_loads(ctx));
    // the load of an unassembled problem is the external load
    JFM_HIP(ctx, hipMemcpyAsync(ctx->load, ctx->f_ext, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (int rc = sync(ctx)) return rc;
    ctx->has_bcs = true;
    ctx->has_load = true;
    jfem_learned_drop(ctx);
    ctx->learned_ready = false;
    ctx->has_user = false;
    for (Workspace *w : {&ctx->ws64, &ctx->ws32}) w->ok = false;  // preconditioners depend on the fixed set
    jfem_coarse_release(ctx);
    return JFM_OK;
}

// ================================================================ pattern ==
extern "C" int jfm_csr_pattern(jfm_ctx *ctx, int64_t *nnz_out, int64_t *rowptr, int64_t *colind, int32_t base) {
    if (!ctx) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (nnz_out) *nnz_out = 9 * ctx->nnzb;
    if (!rowptr && !colind) return JFM_OK;
    const std::vector<int32_t> &br = ctx->h_brow, &bc = ctx->h_bcol;
    parallel_for(ctx->N, [&](int64_t a0, int64_t a1) {
        for (int64_t a = a0; a < a1; ++a) {
            const int64_t r0 = br[a], nb = br[a + 1] - r0;
            for (int c = 0; c < 3; ++c) {
                const int64_t start = 9 * r0 + c * 3 * nb;
                if (rowptr) rowptr[3 * a + c] = start + base;
                if (colind)
                    for (int64_t k = 0; k < nb; ++k)
                        for (int d = 0; d < 3; ++d) colind[start + 3 * k + d] = 3 * (int64_t)bc[r0 + k] + d + base;
            }
        }
    });
    if (rowptr) rowptr[ctx->ndof] = 9 * ctx->nnzb + base;
    return JFM_OK;
}

// =============================================================== operator ==
template <typename T> static T *vals_of(jfm_ctx *ctx);
template <> double *vals_of<double>(jfm_ctx *ctx) { return ctx->vals64; }
template <> float *vals_of<float>(jfm_ctx *ctx) { return ctx->vals32; }

// the EBE operator's element matrices (FP64 and FP32, pair-major)
// With the patch operator, the element matrices are stored in its element
// order and the EBE product runs in the patch form (element results summed
// in LDS, shared nodes by the combine kernel).
static int ensure_element_matrices(jfm_ctx *ctx) {
    if (ctx->Ke) return JFM_OK;
    if (jfem_patch_enabled())
        if (int rc = jfem_patch_setup(ctx)) return rc;
    ALLOC(ctx, ctx->Ke, double, ctx->n_el * ke_stride(ctx->npair, 2));
    ALLOC(ctx, ctx->Ke32, float, ctx->n_el * ke_stride(ctx->npair, 4));
    ctx->ke_perm = ctx->patch_ready ? ctx->eperm : nullptr;
    CHECK_LAUNCH(ctx, launch_element_matrices(ctx));
    return JFM_OK;
}

// CSR values, each block recomputed from the geometric factors of its elements
static int ensure_assembled(jfm_ctx *ctx) {
    if (ctx->assembled) return JFM_OK;
    if (!ctx->vals64) ALLOC(ctx, ctx->vals64, double, 9 * ctx->nnzb);
    if (!ctx->vals32) ALLOC(ctx, ctx->vals32, float, 9 * ctx->nnzb);
    CHECK_LAUNCH(ctx, launch_assemble_values(ctx));
    ctx->assembled = true;
    return JFM_OK;
}

static int prepare_op(jfm_ctx *ctx, int op) {
    if (op == JFM_OP_CSR) return ensure_assembled(ctx);
    if (op == JFM_OP_EBE) return ensure_element_matrices(ctx);
    if (op == JFM_OP_MATFREE) return jfem_patch_enabled() ? jfem_patch_setup(ctx) : JFM_OK;
    if (op == kOpMass) return JFM_OK;
    if (op == kOpHeat) return ctx->heat_k ? JFM_OK : jfm_fail(ctx, JFM_ESTATE, "no heat problem uploaded");
    if (op == JFM_OP_TANGENT)
        return ctx->has_state ? JFM_OK
                              : jfm_fail(ctx, JFM_ESTATE, "tangent operator needs a state: call jfm_nl_residual first");
    return jfm_fail(ctx, JFM_EINVAL, "unknown operator variant");
}

// y = K x on free rows (fixed rows zeroed); optional partials of dot(y, dotv)
// into `part` (one per block of the last kernel, count in *npart).
template <typename T>
static hipError_t op_apply(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, double *part, const int *done,
                           int *npart) {
    if (op == JFM_OP_CSR) return launch_bsr_spmv<T>(ctx, vals_of<T>(ctx), x, y, dotv, part, done, npart);
    if (op == kOpHeat) {
        if (sizeof(T) != 8) return hipErrorInvalidValue;
        return launch_heat_apply(ctx, (const double *)x, (double *)y, (const double *)dotv, part, done, npart);
    }
    return launch_two_phase<T>(ctx, op, x, y, dotv, part, done, npart);
}

// Same without the row mask (all rows), for f_int and reactions.
static hipError_t op_apply_all_rows(jfm_ctx *ctx, int op, const double *x, double *y) {
    uint8_t *saved = ctx->fixed;
    ctx->fixed = nullptr;
    hipError_t e = op_apply<double>(ctx, op, x, y, nullptr, nullptr, nullptr, nullptr);
    ctx->fixed = saved;
    return e;
}

extern "C" int jfm_apply(jfm_ctx *ctx, int32_t op, int32_t precision, const double *x, double *y) {
    if (!ctx || !x || !y) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (precision != JFM_FP64 && precision != JFM_FP32) return jfm_fail(ctx, JFM_EINVAL, "precision must be FP64 or FP32");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = prepare_op(ctx, op)) return rc;
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    JFM_HIP(ctx, hipMemcpyAsync(ctx->vx64, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (precision == JFM_FP64) {
        CHECK_LAUNCH(ctx, op_apply<double>(ctx, op, ctx->vx64, ctx->vy64, nullptr, nullptr, nullptr, nullptr));
    } else {
        CHECK_LAUNCH(ctx, (launch_convert<double, float>(ctx, ctx->vx64, ctx->vx32, n, 1.0)));
        CHECK_LAUNCH(ctx, op_apply<float>(ctx, op, ctx->vx32, ctx->vy32, nullptr, nullptr, nullptr, nullptr));
        CHECK_LAUNCH(ctx, (launch_convert<float, double>(ctx, ctx->vy32, ctx->vy64, n, 1.0)));
    }
    JFM_HIP(ctx, hipMemcpyAsync(y, ctx->vy64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    return sync(ctx);
}

extern "C" int jfm_coarse_matrix(jfm_ctx *ctx, int32_t op, int32_t coarse_nodes, int64_t *nc, double *Ac,
                                 int32_t *node_agg) {
    if (!ctx || !nc || coarse_nodes <= 0) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (op != JFM_OP_MATFREE && op != JFM_OP_CSR && op != JFM_OP_TANGENT)
        return jfm_fail(ctx, JFM_EINVAL, "op must be MATFREE, CSR or TANGENT");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    return jfem_coarse_matrix(ctx, op, coarse_nodes, nc, Ac, node_agg);
}

extern "C" int jfm_time_apply(jfm_ctx *ctx, int32_t op, int32_t precision, int32_t reps, double *avg_ms) {
    if (!ctx || !avg_ms || reps <= 0) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (precision != JFM_FP64 && precision != JFM_FP32) return jfm_fail(ctx, JFM_EINVAL, "precision must be FP64 or FP32");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = prepare_op(ctx, op)) return rc;
    hipEvent_t t0, t1;
    JFM_HIP(ctx, hipEventCreate(&t0));
    JFM_HIP(ctx, hipEventCreate(&t1));
    const bool two = op != JFM_OP_CSR;
    const bool f64 = precision == JFM_FP64;
    // part 0: whole product, 1: SpMV or element phase, 2: gather
    auto run = [&](int part) -> hipError_t {
        if (!two) {
            return f64 ? launch_bsr_spmv<double>(ctx, ctx->vals64, ctx->vx64, ctx->vy64, nullptr, nullptr, nullptr, nullptr)
                       : launch_bsr_spmv<float>(ctx, ctx->vals32, ctx->vx32, ctx->vy32, nullptr, nullptr, nullptr, nullptr);
        }
        hipError_t e = hipSuccess;
        if (jfem_patch_form(ctx, op)) {  // patch kernel, then the shared-node kernel
            if (part != 2)
                e = f64 ? launch_patch_phase<double>(ctx, op, ctx->vx64, ctx->vy64, nullptr, nullptr)
                        : launch_patch_phase<float>(ctx, op, ctx->vx32, ctx->vy32, nullptr, nullptr);
            if (e != hipSuccess || part == 1) return e;
            return f64 ? launch_patch_combine<double>(ctx, ctx->vy64, nullptr, nullptr, nullptr, nullptr)
                       : launch_patch_combine<float>(ctx, ctx->vy32, nullptr, nullptr, nullptr, nullptr);
        }
        if (part != 2) e = f64 ? launch_elem_only<double>(ctx, op, ctx->vx64, nullptr) : launch_elem_only<float>(ctx, op, ctx->vx32, nullptr);
        if (e != hipSuccess || part == 1) return e;
        return f64 ? launch_gather<double>(ctx, ctx->ye64, ctx->fixed, ctx->vy64, nullptr, nullptr, nullptr, nullptr)
                   : launch_gather<float>(ctx, ctx->ye32, ctx->fixed, ctx->vy32, nullptr, nullptr, nullptr, nullptr);
    };
    // each part as `reps` back-to-back launches between two events, so the
    // averages are kernel durations plus launch gaps, without per-launch events
    for (int part = 0; part < (two ? 3 : 2); ++part) {
        CHECK_LAUNCH(ctx, run(part));  // warm
        JFM_HIP(ctx, hipEventRecord(t0, ctx->stream));
        for (int r = 0; r < reps; ++r) CHECK_LAUNCH(ctx, run(part));
        JFM_HIP(ctx, hipEventRecord(t1, ctx->stream));
        JFM_HIP(ctx, hipEventSynchronize(t1));
        float ms = 0.f;
        JFM_HIP(ctx, hipEventElapsedTime(&ms, t0, t1));
        avg_ms[part] = ms / reps;
    }
    if (!two) avg_ms[2] = 0.0;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return JFM_OK;
}

extern "C" int jfm_assemble(jfm_ctx *ctx, const double *u, double *vals, double *f) {
    if (!ctx) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->assembled = false;
    if (int rc = ensure_assembled(ctx)) return rc;
    const int64_t n = ctx->ndof;
    // f = f_ext - f_int(u)  (src/problems_elasticity.jl:432-433)
    if (u) {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->vx64, u, sizeof(double) * n, hipMemcpyHostToDevice, s));
        CHECK_LAUNCH(ctx, op_apply_all_rows(ctx, JFM_OP_MATFREE, ctx->vx64, ctx->vy64));
        CHECK_LAUNCH(ctx, (launch_convert<double, double>(ctx, ctx->vy64, ctx->vy64, n, -1.0)));
        std::vector<double> fi(n), fe(n);
        JFM_HIP(ctx, hipMemcpyAsync(fi.data(), ctx->vy64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipMemcpyAsync(fe.data(), ctx->f_ext, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (int rc = sync(ctx)) return rc;
        for (int64_t i = 0; i < n; ++i) fe[i] += fi[i];
        JFM_HIP(ctx, hipMemcpyAsync(ctx->load, fe.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (int rc = sync(ctx)) return rc;
    } else {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->load, ctx->f_ext, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    ctx->has_load = true;
    if (vals) JFM_HIP(ctx, hipMemcpyAsync(vals, ctx->vals64, sizeof(double) * 9 * ctx->nnzb, hipMemcpyDeviceToHost, s));
    if (f) JFM_HIP(ctx, hipMemcpyAsync(f, ctx->load, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    return sync(ctx);
}

// Finite strain (finite_strain = true, src/problems_elasticity.jl:203-451):
// Km(u), Kg(u) and f = f_ext - f_int(u).  The device CSR operator then holds
// K = Km + Kg (Km alone without geometric stiffness), the K the reference
// solves (src/solvers.jl:286-292).
extern "C" int jfm_assemble_tl(jfm_ctx *ctx, int32_t geometric, const double *u, double *vals, double *vals_kg,
                               double *f) {
    if (!ctx) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (vals_kg && !geometric) return jfm_fail(ctx, JFM_EINVAL, "vals_kg requested without geometric stiffness");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->ndof;
    const size_t vbytes = sizeof(double) * 9 * (size_t)ctx->nnzb;
    double *Ke = nullptr;   // finite-strain element blocks, element-major, freed below
    ALLOC(ctx, Ke, double, ctx->n_el * ctx->npair * 9);
    if (!ctx->vals64) ALLOC(ctx, ctx->vals64, double, 9 * ctx->nnzb);
    if (!ctx->vals32) ALLOC(ctx, ctx->vals32, float, 9 * ctx->nnzb);
    ctx->assembled = false;
    if (u) JFM_HIP(ctx, hipMemcpyAsync(ctx->vx64, u, sizeof(double) * n, hipMemcpyHostToDevice, s));
    else JFM_HIP(ctx, hipMemsetAsync(ctx->vx64, 0, sizeof(double) * n, s));
    if (geometric && vals_kg) {
        CHECK_LAUNCH(ctx, launch_tl_element_matrices(ctx, ctx->vx64, 1, Ke, nullptr));
        CHECK_LAUNCH(ctx, launch_assemble_values_ke(ctx, Ke));
        JFM_HIP(ctx, hipMemcpyAsync(vals_kg, ctx->vals64, vbytes, hipMemcpyDeviceToHost, s));
    }
    // Km and the element forces; f = f_ext - f_int (src/problems_elasticity.jl:432-433)
    CHECK_LAUNCH(ctx, launch_tl_element_matrices(ctx, ctx->vx64, 0, Ke, ctx->ye64));
    CHECK_LAUNCH(ctx, launch_assemble_values_ke(ctx, Ke));
    if (vals) JFM_HIP(ctx, hipMemcpyAsync(vals, ctx->vals64, vbytes, hipMemcpyDeviceToHost, s));
    CHECK_LAUNCH(ctx, launch_gather<double>(ctx, ctx->ye64, nullptr, ctx->vy64, nullptr, nullptr, nullptr, nullptr));
    CHECK_LAUNCH(ctx, (launch_convert<double, double>(ctx, ctx->vy64, ctx->vy64, n, -1.0)));
    CHECK_LAUNCH(ctx, launch_add(ctx, ctx->f_ext, ctx->vy64, ctx->load, n));
    if (geometric) {
        CHECK_LAUNCH(ctx, launch_tl_element_matrices(ctx, ctx->vx64, 2, Ke, nullptr));
        CHECK_LAUNCH(ctx, launch_assemble_values_ke(ctx, Ke));
    }
    if (f) JFM_HIP(ctx, hipMemcpyAsync(f, ctx->load, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (int rc = sync(ctx)) return rc;
    free_ptr(ctx, Ke);
    ctx->assembled = true;
    ctx->has_load = true;
    return JFM_OK;
}

// ================================================================ solvers ==
template <typename T> static Workspace &ws_of(jfm_ctx *ctx);
template <> Workspace &ws_of<double>(jfm_ctx *ctx) { return ctx->ws64; }
template <> Workspace &ws_of<float>(jfm_ctx *ctx) { return ctx->ws32; }

static int ensure_workspaces(jfm_ctx *ctx) {
    const int64_t n = ctx->ndof, N = ctx->N;
    if (!ctx->ws64.x) {
        Workspace &w = ctx->ws64;
        ALLOC(ctx, w.x, double, n); ALLOC(ctx, w.r, double, n); ALLOC(ctx, w.z, double, n);
        ALLOC(ctx, w.p, double, n); ALLOC(ctx, w.q, double, n); ALLOC(ctx, w.b, double, n);
        ALLOC(ctx, w.Minv, double, 9 * N);
    }
    if (!ctx->ws32.x) {
        Workspace &w = ctx->ws32;
        ALLOC(ctx, w.x, float, n); ALLOC(ctx, w.r, float, n); ALLOC(ctx, w.z, float, n);
        ALLOC(ctx, w.p, float, n); ALLOC(ctx, w.q, float, n); ALLOC(ctx, w.b, float, n);
        ALLOC(ctx, w.Minv, float, 9 * N);
    }
    if (!ctx->u64) ALLOC(ctx, ctx->u64, double, n);
    if (!ctx->rr64) ALLOC(ctx, ctx->rr64, double, n);
    if (!ctx->ws64.ok) {
        CHECK_LAUNCH(ctx, launch_block_jacobi(ctx, (double *)ctx->ws64.Minv, (float *)ctx->ws32.Minv));
        ctx->ws64.ok = ctx->ws32.ok = true;
    }
    return JFM_OK;
}

int jfem_ensure_workspaces(jfm_ctx *ctx) { return ensure_workspaces(ctx); }

int jfem_apply(jfm_ctx *ctx, int op, const double *x, double *y) {
    if (int rc = prepare_op(ctx, op)) return rc;
    CHECK_LAUNCH(ctx, op_apply<double>(ctx, op, x, y, nullptr, nullptr, nullptr, nullptr));
    return JFM_OK;
}

static int read_scalars(jfm_ctx *ctx);
static int finalize_misc(jfm_ctx *ctx, int npart, int nv, int fin);

int jfem_dot(jfm_ctx *ctx, const double *a, const double *b, int64_t n, double *out) {
    int np = 0;
    CHECK_LAUNCH(ctx, launch_dot<double>(ctx, a, b, n, &np));
    if (int rc = finalize_misc(ctx, np, 1, FIN_DOT)) return rc;
    if (int rc = read_scalars(ctx)) return rc;
    *out = ctx->S_host->tmp0;
    return JFM_OK;
}

static int read_scalars(jfm_ctx *ctx) {
    JFM_HIP(ctx, hipMemcpyAsync(ctx->S_host, ctx->S, sizeof(PcgScalars), hipMemcpyDeviceToHost, ctx->stream));
    return sync(ctx);
}

// S_host is pinned and stream-ordered with the kernels that read S; the
// caller does not write S_host again before its next read_scalars, so no
// wait is needed here.
static int write_scalars(jfm_ctx *ctx) {
    JFM_HIP(ctx, hipMemcpyAsync(ctx->S, ctx->S_host, sizeof(PcgScalars), hipMemcpyHostToDevice, ctx->stream));
    return JFM_OK;
}

// One PCG iteration of parity `slot`: the operator, then the update and the
// direction (linalg_kernels.hip), with the coarse GEMV between them for the
// two-level preconditioner (coarse.hip); every kernel returns at once when
// done[slot] is set.
template <typename T>
static hipError_t pcg_iteration(jfm_ctx *ctx, int op, int slot, const T *Minv, bool coarse) {
    Workspace &w = ws_of<T>(ctx);
    T *x = (T *)w.x, *r = (T *)w.r, *z = (T *)w.z, *p = (T *)w.p, *q = (T *)w.q;
    int np_pq = 0, np_rz = 0;
    const bool fused = coarse && jfem_coarse_fused_restriction(ctx);
    // with the shared nodes' sums left to the update kernel, K p is the patch
    // kernel alone, its per-patch energies the p.q partials
    const bool defer = fused && jfem_patch_defer_shared(ctx, op);
    hipError_t e = defer ? launch_patch_phase<T>(ctx, op, p, q, p, &ctx->S->done[slot], ctx->partials + part_pq(slot))
                         : op_apply<T>(ctx, op, p, q, p, ctx->partials + part_pq(slot), &ctx->S->done[slot], &np_pq);
    if (e != hipSuccess) return e;
    if (defer) np_pq = ctx->npatch;
    if (fused) {
        // the update's blocks, one per aggregate, form w = Q' m r for the GEMV
        int np_wz = 0;
        e = launch_pcg_update<T>(ctx, slot, np_pq, x, r, z, p, q, Minv, &np_rz, nullptr, true, defer);
        if (e != hipSuccess) return e;
        e = launch_coarse_gemv_fused(ctx, slot, &np_wz);
        if (e != hipSuccess) return e;
        return launch_pcg_direction_coarse<T>(ctx, slot, np_rz, np_wz, p, z, x);  // x += alpha p here
    }
    // with the coarse level, the update also writes m o r in aggregate order
    T *rA = coarse ? coarse_rA<T>(ctx) : nullptr;
    e = launch_pcg_update<T>(ctx, slot, np_pq, x, r, z, p, q, Minv, &np_rz, rA);
    if (e != hipSuccess) return e;
    if (coarse) {  // zc = A_c^-1 Q' r, applied inside the direction update
        int np_wz = 0;
        e = launch_coarse_product_agg<T>(ctx, rA, &ctx->S->done[slot], &np_wz);
        if (e != hipSuccess) return e;
        return launch_pcg_direction_coarse<T>(ctx, slot, np_rz, np_wz, p, z, nullptr);
    }
    return launch_pcg_direction<T>(ctx, slot, np_rz, p, z);
}

bool jfem::jfem_use_graphs() {
    static const bool use = [] {
        const char *v = std::getenv("JFEM_GRAPHS");
        return !(v && v[0] == '0');
    }();
    return use;
}

// A batch of `batch` (even) iterations captured once as a HIP graph and
// replayed; kernels past convergence return immediately.
template <typename T>
static int pcg_batch(jfm_ctx *ctx, int op, const T *Minv, int batch, bool coarse) {
    const int prec = sizeof(T) == 8 ? JFM_FP64 : JFM_FP32;
    for (auto &g : ctx->graphs)
        if (g.op == op && g.prec == prec && g.minv == (const void *)Minv && g.batch == batch &&
            g.coarse == (coarse ? ctx->coarse_gen : 0)) {
            JFM_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
            return JFM_OK;
        }
    if (!jfem_use_graphs()) {
        for (int it = 0; it < batch; ++it) CHECK_LAUNCH(ctx, pcg_iteration<T>(ctx, op, it & 1, Minv, coarse));
        return JFM_OK;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ok = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    hipError_t launch_err = hipSuccess;
    if (ok) {
        for (int it = 0; it < batch && launch_err == hipSuccess; ++it) launch_err = pcg_iteration<T>(ctx, op, it & 1, Minv, coarse);
        ok = hipStreamEndCapture(ctx->stream, &graph) == hipSuccess && launch_err == hipSuccess;
    }
    if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!ok) {  // capture unavailable: launch the same kernels directly
        if (exec) (void)hipGraphExecDestroy(exec);
        for (int it = 0; it < batch; ++it) CHECK_LAUNCH(ctx, pcg_iteration<T>(ctx, op, it & 1, Minv, coarse));
        return JFM_OK;
    }
    jfm_ctx::Graph g;
    g.op = op;
    g.prec = prec;
    g.minv = Minv;
    g.batch = batch;
    g.coarse = coarse ? ctx->coarse_gen : 0;
    g.exec = exec;
    ctx->graphs.push_back(g);
    JFM_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
    return JFM_OK;
}

// sum of the misc partials -> scalars (one-block kernel)
static int finalize_misc(jfm_ctx *ctx, int npart, int nv, int fin) {
    CHECK_LAUNCH(ctx, launch_finalize(ctx, ctx->partials + kPartMisc, npart, nv, fin));
    return JFM_OK;
}

// PCG on the workspace of precision T: solves K x = b (free DOFs) from x = 0,
// or from the current x with its true residual (warm, FP64 only).  The host
// reads the scalars once per batch of `check_every` iterations.
template <typename T>
static int pcg_run(jfm_ctx *ctx, int op, const T *Minv, double rtol, int maxit, int check_every, int *iters,
                   int *status, bool warm = false, bool coarse = false) {
    Workspace &w = ws_of<T>(ctx);
    T *x = (T *)w.x, *r = (T *)w.r, *z = (T *)w.z, *p = (T *)w.p, *q = (T *)w.q, *b = (T *)w.b;
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    std::memset(ctx->S_host, 0, sizeof(PcgScalars));
    ctx->S_host->rtol2 = rtol * rtol;
    ctx->S_host->maxit = maxit;
    if (int rc = write_scalars(ctx)) return rc;
    int np = 0;
    if (warm && sizeof(T) == 8) {
        // restart from the current x with the true residual r = b - K x
        CHECK_LAUNCH(ctx, op_apply<T>(ctx, op, x, q, nullptr, nullptr, nullptr, nullptr));
        CHECK_LAUNCH(ctx, launch_residual(ctx, (const double *)b, (const double *)q, ctx->fixed, (double *)r, n, &np));
    } else {
        JFM_HIP(ctx, hipMemsetAsync(x, 0, sizeof(T) * n, s));
        JFM_HIP(ctx, hipMemcpyAsync(r, b, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    }
    CHECK_LAUNCH(ctx, launch_dot<T>(ctx, b, b, n, &np));
    if (int rc = finalize_misc(ctx, np, 1, FIN_BB)) return rc;
    CHECK_LAUNCH(ctx, launch_pcg_init<T>(ctx, r, z, p, Minv, &np));
    if (coarse) CHECK_LAUNCH(ctx, launch_coarse_correct<T>(ctx, r, z, p, ctx->partials + kPartMisc, nullptr, &np));
    if (coarse && jfem_coarse_fused_restriction(ctx))
        if (int rc = jfem_patch_qsrc(ctx)) return rc;
    if (int rc = finalize_misc(ctx, np, 2, FIN_INIT)) return rc;
    // Batches of up to check_every iterations; once the residual falls, the
    // next batch is the power of two that covers the iterations its rate
    // predicts are left, so that few launches run past convergence.  The
    // first batch follows the start without a host read (its kernels return
    // at once if the start already converged); the start's r.r comes back
    // with the first batch's scalars (rr0).
    const int bmax = std::max(2, (check_every + 1) & ~1);
    int batch = bmax;
    double rr_prev = -1.0;
    int it_prev = 0;
    for (bool first = true; first || ctx->S_host->done[0] == DONE_RUNNING; first = false) {
        if (int rc = pcg_batch<T>(ctx, op, Minv, batch, coarse)) return rc;
        if (int rc = read_scalars(ctx)) return rc;
        if (first) rr_prev = ctx->S_host->rr0;
        const double rr = ctx->S_host->rr, tol2 = ctx->S_host->tol2;
        const int it = ctx->S_host->iter;
        batch = bmax;
        if (rr > tol2 && tol2 > 0.0 && rr < rr_prev && it > it_prev) {
            const double left = std::log(tol2 / rr) / (std::log(rr / rr_prev) / (it - it_prev));
            int nb = 2;
            while (nb < left + 1.0 && nb < bmax) nb *= 2;
            batch = std::min(nb, bmax);
        }
        rr_prev = rr;
        it_prev = it;
    }
    *iters = ctx->S_host->iter;
    *status = ctx->S_host->done[0];
    return JFM_OK;
}

// FP64 true residual r = b - K u over free DOFs; *rnorm = |r|.
static int true_residual(jfm_ctx *ctx, int op, const double *b, const double *u, double *r, double *rnorm) {
    int np = 0;
    CHECK_LAUNCH(ctx, op_apply<double>(ctx, op, u, ctx->vy64, nullptr, nullptr, nullptr, nullptr));
    CHECK_LAUNCH(ctx, launch_residual(ctx, b, ctx->vy64, ctx->fixed, r, ctx->ndof, &np));
    if (int rc = finalize_misc(ctx, np, 1, FIN_DOT)) return rc;
    if (int rc = read_scalars(ctx)) return rc;
    *rnorm = std::sqrt(ctx->S_host->tmp0);
    return JFM_OK;
}

// Learned blocks against block Jacobi for the system in ws64.b: builds the
// learned blocks from the block Jacobi now in ws*.Minv, then, with probe > 0,
// runs `probe` PCG iterations from 0 with each.  After the same number of
// iterations the iterate with the larger energy b.x_k (CG decreases
// 1/2 x'Kx - b.x = -1/2 b.x_k monotonically) has the smaller energy-norm
// error; the learned blocks are kept only if they win.
static int learned_select(jfm_ctx *ctx, int op, int probe, int check, bool *use, int *iters, double *setup_ms,
                          bool coarse) {
    double lms = 0.0;
    if (int rc = jfem_learned_prepare(ctx, &lms)) return rc;
    if (setup_ms) *setup_ms += lms;
    *use = true;
    if (probe <= 0) return JFM_OK;
    // The probe runs in FP32 (FP32 operator and blocks) on b rounded to FP32:
    // it only ranks the two preconditioners by the energy b'x_k.
    double phi[2] = {0.0, 0.0};
    const float *cand[2] = {(const float *)ctx->ws32.Minv, ctx->Mlearn32};
    const int64_t n = ctx->ndof;
    if (!ctx->xprobe) ALLOC(ctx, ctx->xprobe, double, n);
    double *xd[2] = {ctx->xprobe, ctx->vy64};  // the iterates in FP64 (vy64 is free here)
    CHECK_LAUNCH(ctx, (launch_convert<double, float>(ctx, (const double *)ctx->ws64.b, (float *)ctx->ws32.b, n, 1.0)));
    for (int c = 0; c < 2; ++c) {
        int it = 0, status = 0;
        if (int rc = pcg_run<float>(ctx, op, cand[c], 1e-30, probe, check, &it, &status, false, coarse)) return rc;
        CHECK_LAUNCH(ctx, (launch_convert<float, double>(ctx, (const float *)ctx->ws32.x, xd[c], n, 1.0)));
        if (int rc = jfem_dot(ctx, (const double *)ctx->ws64.b, xd[c], n, &phi[c])) return rc;
        *iters += it;
    }
    *use = phi[1] > phi[0];
    // the winner's probe iterate starts the solve
    if (*use) JFM_HIP(ctx, hipMemcpyAsync(ctx->xprobe, xd[1], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    return JFM_OK;
}

extern "C" int jfm_solve(jfm_ctx *ctx, const jfm_solve_opts *opts, const double *f, double *u, double *la,
                         jfm_solve_stats *stats) {
    if (!ctx || !opts || !u) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (!f && !ctx->has_load) return jfm_fail(ctx, JFM_ESTATE, "no load vector: upload BCs or pass f");
    if (!(opts->rtol > 0.0) || opts->maxit <= 0) return jfm_fail(ctx, JFM_EINVAL, "rtol > 0 and maxit > 0 required");
    if (opts->method != JFM_SOLVE_PCG && opts->method != JFM_SOLVE_MIXED_IR)
        return jfm_fail(ctx, JFM_EINVAL, "unknown solve method");
    if (opts->precond != JFM_PREC_BLOCK_JACOBI && opts->precond != JFM_PREC_LEARNED && opts->precond != JFM_PREC_USER)
        return jfm_fail(ctx, JFM_EINVAL, "unknown preconditioner");
    if (opts->precond == JFM_PREC_USER && !ctx->has_user)
        return jfm_fail(ctx, JFM_ESTATE, "no user preconditioner blocks set");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    const int op = opts->op;
    const int op_outer = opts->method == JFM_SOLVE_MIXED_IR ? opts->op_outer : op;
    if (int rc = prepare_op(ctx, op)) return rc;
    if (int rc = prepare_op(ctx, op_outer)) return rc;
    if (int rc = ensure_workspaces(ctx)) return rc;
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    const int check = std::max(2, opts->check_every > 0 ? opts->check_every : 32);
    jfm_solve_stats st{};
    hipEvent_t e0, e1;
    JFM_HIP(ctx, hipEventCreate(&e0));
    JFM_HIP(ctx, hipEventCreate(&e1));

    // right-hand side b = f - K g on free rows (elimination of prescribed
    // DOFs, src/solvers.jl:206-210); f stays in vx64 for the reactions
    if (f) JFM_HIP(ctx, hipMemcpyAsync(ctx->vx64, f, sizeof(double) * n, hipMemcpyHostToDevice, s));
    else JFM_HIP(ctx, hipMemcpyAsync(ctx->vx64, ctx->load, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    double *b = (double *)ctx->ws64.b;
    JFM_HIP(ctx, hipEventRecord(e0, s));
    double bnorm = 0.0;
    if (int rc = true_residual(ctx, op_outer, ctx->vx64, ctx->g, b, &bnorm)) return rc;
    ctx->hist_it.clear();
    ctx->hist_res.clear();
    const bool coarse = opts->coarse_nodes > 0 && bnorm > 0.0;
    // the learned blocks on a second stream beside the coarse level's
    // inversion (both read only the mesh and the block Jacobi)
    if (coarse)
        if (int rc = jfem_coarse_setup(ctx, op_outer, opts->coarse_nodes, &st.setup_ms,
                                       opts->precond == JFM_PREC_LEARNED)) {
            jfem_learned_drop(ctx);  // no learned work is left running past the call
            return rc;
        }
    const double *Minv64 = (const double *)ctx->ws64.Minv;
    const float *Minv32 = (const float *)ctx->ws32.Minv;
    if (opts->precond == JFM_PREC_USER) {
        Minv64 = ctx->Muser64;
        Minv32 = ctx->Muser32;
    }
    const double *x0 = nullptr;  // initial iterate (b = f - K g system), when a probe made one
    if (opts->precond == JFM_PREC_LEARNED && bnorm > 0.0) {
        bool use = true;
        int its = 0;
        if (int rc = learned_select(ctx, op_outer, opts->learned_probe, check, &use, &its, &st.setup_ms, coarse))
            return rc;
        st.iterations += its;
        x0 = opts->learned_probe > 0 ? ctx->xprobe : nullptr;
        if (use) {
            Minv64 = ctx->Mlearn64;
            Minv32 = ctx->Mlearn32;
        } else {
            st.fallback = 1;
        }
    }

    if (bnorm == 0.0) {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->u64, ctx->g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else if (opts->method == JFM_SOLVE_PCG) {
        int it = 0, status = 0, total = 0, restarts = 0;
        if (x0) JFM_HIP(ctx, hipMemcpyAsync(ctx->ws64.x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        if (int rc = pcg_run<double>(ctx, op, Minv64, opts->rtol, opts->maxit, check, &it, &status, x0 != nullptr,
                                     coarse))
            return rc;
        total = it;
        // The recursive residual drifts from the true one near the FP64 floor:
        // restart from x with the true residual until it meets rtol (residual
        // replacement), a bounded number of times.
        while (status == DONE_CONVERGED && restarts < 8 && total < opts->maxit) {
            double rn = 0.0;
            if (int rc = true_residual(ctx, op, b, (const double *)ctx->ws64.x, ctx->rr64, &rn)) return rc;
            ctx->hist_it.push_back(total);
            ctx->hist_res.push_back(rn / bnorm);
            if (rn <= opts->rtol * bnorm) break;
            if (int rc = pcg_run<double>(ctx, op, Minv64, opts->rtol, opts->maxit - total, check, &it, &status, true,
                                         coarse))
                return rc;
            total += it;
            ++restarts;
        }
        st.iterations += total;
        st.outer = 1 + restarts;
        // u = x + g
        CHECK_LAUNCH(ctx, launch_add(ctx, ctx->g, (const double *)ctx->ws64.x, ctx->u64, n));
    } else {
        // FP64 iterative refinement around FP32 PCG: r = f - K u in FP64,
        // K d = r / |r| in FP32 to inner_rtol, u += |r| d.
        if (x0) CHECK_LAUNCH(ctx, launch_add(ctx, ctx->g, x0, ctx->u64, n));
        else JFM_HIP(ctx, hipMemcpyAsync(ctx->u64, ctx->g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        int total = 0, outer = 0;
        const double inner = opts->inner_rtol > 0 ? opts->inner_rtol : 1e-4;
        const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
        if (opts->first_rtol > 0.0) {
            // FP64 phase for the smooth low modes that FP32 cannot resolve
            int it = 0, status = 0;
            if (x0) JFM_HIP(ctx, hipMemcpyAsync(ctx->ws64.x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            if (int rc = pcg_run<double>(ctx, op_outer, Minv64, opts->first_rtol, opts->maxit, check, &it, &status,
                                         x0 != nullptr, coarse))
                return rc;
            total += it;
            CHECK_LAUNCH(ctx, launch_add(ctx, ctx->g, (const double *)ctx->ws64.x, ctx->u64, n));
        }
        while (outer < max_outer) {
            double rn = 0.0;
            if (int rc = true_residual(ctx, op_outer, ctx->vx64, ctx->u64, ctx->rr64, &rn)) return rc;
            ctx->hist_it.push_back(total);
            ctx->hist_res.push_back(rn / bnorm);
            if (rn <= opts->rtol * bnorm || total >= opts->maxit) break;
            CHECK_LAUNCH(ctx, (launch_convert<double, float>(ctx, ctx->rr64, (float *)ctx->ws32.b, n, 1.0 / rn)));
            int it = 0, status = 0;
            int cap = opts->maxit - total;
            if (opts->inner_maxit > 0) cap = std::min(cap, opts->inner_maxit);
            // spread the remaining reduction (to 0.3 x the target) evenly over
            // the fewest inner solves that each reduce by at most inner_rtol:
            // the last solve then ends well below the target instead of just
            // above it, and no solve asks FP32 for more than inner_rtol
            double rt = inner;
            const double R = 0.3 * opts->rtol * bnorm / rn;
            if (R < 1.0 && inner < 1.0) {
                const int k = std::max(1, (int)std::ceil(std::log(R) / std::log(inner) - 1e-9));
                rt = std::pow(R, 1.0 / k);
            }
            if (int rc = pcg_run<float>(ctx, op, Minv32, rt, cap, check, &it, &status, false, coarse)) return rc;
            total += it;
            ++outer;
            CHECK_LAUNCH(ctx, launch_axpy_scaled(ctx, ctx->u64, (const float *)ctx->ws32.x, rn, n));
            if (status == DONE_BREAKDOWN) break;
        }
        st.iterations += total;
        st.outer = outer;
    }
    // final FP64 true residual with the FP64 operator
    double rn = 0.0;
    if (int rc = true_residual(ctx, op_outer, ctx->vx64, ctx->u64, ctx->rr64, &rn)) return rc;
    JFM_HIP(ctx, hipEventRecord(e1, s));
    JFM_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.relres = bnorm > 0 ? rn / bnorm : 0.0;
    st.converged = st.relres <= opts->rtol ? 1 : 0;
    st.solve_ms = ms;
    JFM_HIP(ctx, hipMemcpyAsync(u, ctx->u64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (la) {
        // la = f - K u on fixed DOFs (src/solvers.jl:212-213), 0 elsewhere
        CHECK_LAUNCH(ctx, op_apply_all_rows(ctx, op_outer, ctx->u64, ctx->vy64));
        std::vector<double> ku(n), ff(n);
        std::vector<uint8_t> fxd(n);
        JFM_HIP(ctx, hipMemcpyAsync(ku.data(), ctx->vy64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipMemcpyAsync(ff.data(), ctx->vx64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        JFM_HIP(ctx, hipMemcpyAsync(fxd.data(), ctx->fixed, n, hipMemcpyDeviceToHost, s));
        if (int rc = sync(ctx)) return rc;
        for (int64_t i = 0; i < n; ++i) la[i] = fxd[i] ? ff[i] - ku[i] : 0.0;
    }
    if (int rc = sync(ctx)) return rc;
    if (stats) *stats = st;
    if (!st.converged) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "solver did not reach rtol %.3e: relres %.3e after %d iterations", opts->rtol,
                      st.relres, st.iterations);
        return jfm_fail(ctx, JFM_ENOCONV, msg);
    }
    return JFM_OK;
}

extern "C" int jfm_solve_history(jfm_ctx *ctx, int32_t *iterations, double *relres, int32_t max_entries) {
    if (!ctx || max_entries < 0) return JFM_EINVAL;
    const int n = std::min<int>(max_entries, (int)ctx->hist_it.size());
    for (int k = 0; k < n; ++k) {
        if (iterations) iterations[k] = ctx->hist_it[k];
        if (relres) relres[k] = ctx->hist_res[k];
    }
    return n;
}

extern "C" int jfm_precond_set_blocks(jfm_ctx *ctx, const double *M) {
    if (!ctx || !M) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = ctx->N;
    std::vector<double> m(M, M + 9 * N);
    std::vector<uint8_t> fx(3 * N);
    JFM_HIP(ctx, hipMemcpyAsync(fx.data(), ctx->fixed, 3 * N, hipMemcpyDeviceToHost, ctx->stream));
    if (int rc = sync(ctx)) return rc;
    std::vector<float> m32(9 * N);
    for (int64_t a = 0; a < N; ++a)
        for (int c = 0; c < 3; ++c) {
            for (int d = 0; d < 3; ++d) {
                if (fx[3 * a + c] || fx[3 * a + d]) m[9 * a + 3 * c + d] = 0.0;
                if (!std::isfinite(m[9 * a + 3 * c + d])) return jfm_fail(ctx, JFM_EINVAL, "non-finite preconditioner block");
            }
        }
    for (int64_t k = 0; k < 9 * N; ++k) m32[k] = (float)m[k];
    if (!ctx->Muser64) ALLOC(ctx, ctx->Muser64, double, 9 * N);
    if (!ctx->Muser32) ALLOC(ctx, ctx->Muser32, float, 9 * N);
    JFM_HIP(ctx, hipMemcpyAsync(ctx->Muser64, m.data(), sizeof(double) * 9 * N, hipMemcpyHostToDevice, ctx->stream));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->Muser32, m32.data(), sizeof(float) * 9 * N, hipMemcpyHostToDevice, ctx->stream));
    if (int rc = sync(ctx)) return rc;
    ctx->has_user = true;
    return JFM_OK;
}

// ============================================================== nonlinear ==
static int nl_buffers(jfm_ctx *ctx) {
    const int64_t n = ctx->ndof, N = ctx->N;
    if (!ctx->ustate64) ALLOC(ctx, ctx->ustate64, double, n);
    if (!ctx->ustate32) ALLOC(ctx, ctx->ustate32, float, n);
    if (!ctx->Dtan) ALLOC(ctx, ctx->Dtan, double, 9 * N);
    if (!ctx->nl_u) ALLOC(ctx, ctx->nl_u, double, n);
    if (!ctx->nl_r) ALLOC(ctx, ctx->nl_r, double, n);
    if (!ctx->nl_du) ALLOC(ctx, ctx->nl_du, double, n);
    if (!ctx->nl_zero) {
        ALLOC(ctx, ctx->nl_zero, double, n);
        JFM_HIP(ctx, hipMemsetAsync(ctx->nl_zero, 0, sizeof(double) * n, ctx->stream));
    }
    return JFM_OK;
}

// Makes u (device, n_dofs) the state of the tangent operator and computes
// R = lf f_ext - f_int(u) on the free DOFs (0 on fixed ones), *rnorm = |R|
// (NAN when some Gauss point has det F <= 0).  Uses vx64/vy64 as scratch.
static int nl_residual(jfm_ctx *ctx, const double *u, double lf, double *R, double *rnorm) {
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    if (u != ctx->ustate64) JFM_HIP(ctx, hipMemcpyAsync(ctx->ustate64, u, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    CHECK_LAUNCH(ctx, (launch_convert<double, float>(ctx, ctx->ustate64, ctx->ustate32, n, 1.0)));
    ctx->has_state = true;
    int np = 0;
    CHECK_LAUNCH(ctx, launch_nh_elem<double>(ctx, false, ctx->gf64, ctx->ustate64, nullptr, ctx->ye64, nullptr));
    CHECK_LAUNCH(ctx, launch_gather<double>(ctx, ctx->ye64, ctx->fixed, ctx->vy64, nullptr, nullptr, nullptr, &np));
    CHECK_LAUNCH(ctx, (launch_convert<double, double>(ctx, ctx->f_ext, ctx->vx64, n, lf)));
    CHECK_LAUNCH(ctx, launch_residual(ctx, ctx->vx64, ctx->vy64, ctx->fixed, R, n, &np));
    if (int rc = finalize_misc(ctx, np, 1, FIN_DOT)) return rc;
    if (int rc = read_scalars(ctx)) return rc;
    *rnorm = std::sqrt(ctx->S_host->tmp0);
    return JFM_OK;
}

extern "C" int jfm_nl_residual(jfm_ctx *ctx, const double *u, double load_factor, double *R, double *rnorm) {
    if (!ctx || !u) return JFM_EINVAL;
    if (!ctx->has_mesh || !ctx->has_bcs) return jfm_fail(ctx, JFM_ESTATE, "upload the mesh and boundary conditions first");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = nl_buffers(ctx)) return rc;
    const int64_t n = ctx->ndof;
    JFM_HIP(ctx, hipMemcpyAsync(ctx->nl_u, u, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    double rn = 0.0;
    if (int rc = nl_residual(ctx, ctx->nl_u, load_factor, ctx->nl_r, &rn)) return rc;
    if (R) JFM_HIP(ctx, hipMemcpyAsync(R, ctx->nl_r, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (int rc = sync(ctx)) return rc;
    if (rnorm) *rnorm = rn;
    return JFM_OK;
}

// K_T du = r (r in nl_r, free DOFs) to |r - K_T du| <= eta |r|; du -> nl_du.
static int tangent_solve(jfm_ctx *ctx, const jfm_newton_opts *o, const double *M64, const float *M32, double eta,
                         double rnorm, int check, bool coarse, int *iters) {
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    *iters = 0;
    if (o->method == JFM_SOLVE_PCG) {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->ws64.b, ctx->nl_r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        int it = 0, status = 0;
        if (int rc = pcg_run<double>(ctx, JFM_OP_TANGENT, M64, eta, o->maxit_linear, check, &it, &status, false,
                                     coarse))
            return rc;
        *iters = it;
        JFM_HIP(ctx, hipMemcpyAsync(ctx->nl_du, ctx->ws64.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        return JFM_OK;
    }
    // FP32 PCG inside FP64 refinement on the tangent
    JFM_HIP(ctx, hipMemsetAsync(ctx->nl_du, 0, sizeof(double) * n, s));
    const double inner = o->inner_rtol > 0 ? o->inner_rtol : 1e-2;
    double rl = rnorm;
    for (int outer = 0; outer < 50; ++outer) {
        if (outer > 0)
            if (int rc = true_residual(ctx, JFM_OP_TANGENT, ctx->nl_r, ctx->nl_du, ctx->rr64, &rl)) return rc;
        if (rl <= eta * rnorm || *iters >= o->maxit_linear) break;
        const double *src = outer == 0 ? ctx->nl_r : ctx->rr64;
        CHECK_LAUNCH(ctx, (launch_convert<double, float>(ctx, src, (float *)ctx->ws32.b, n, 1.0 / rl)));
        int it = 0, status = 0;
        int cap = o->maxit_linear - *iters;
        if (o->inner_maxit > 0) cap = std::min(cap, o->inner_maxit);
        // no need to solve the correction beyond what the outer tolerance asks
        const double tol = std::max(inner, 0.5 * eta * rnorm / rl);
        if (int rc = pcg_run<float>(ctx, JFM_OP_TANGENT, M32, tol, cap, check, &it, &status, false, coarse)) return rc;
        *iters += it;
        CHECK_LAUNCH(ctx, launch_axpy_scaled(ctx, ctx->nl_du, (const float *)ctx->ws32.x, rl, n));
        if (status == DONE_BREAKDOWN) break;
    }
    return JFM_OK;
}

extern "C" int jfm_solve_nonlinear(jfm_ctx *ctx, const jfm_newton_opts *opts, double *u, jfm_newton_stats *stats) {
    if (!ctx || !opts || !u) return JFM_EINVAL;
    if (!ctx->has_mesh || !ctx->has_bcs) return jfm_fail(ctx, JFM_ESTATE, "upload the mesh and boundary conditions first");
    if (opts->load_steps < 1 || opts->max_newton < 1 || !(opts->rtol > 0.0) || opts->maxit_linear < 1)
        return jfm_fail(ctx, JFM_EINVAL, "load_steps, max_newton, maxit_linear >= 1 and rtol > 0 required");
    if (opts->method != JFM_SOLVE_PCG && opts->method != JFM_SOLVE_MIXED_IR)
        return jfm_fail(ctx, JFM_EINVAL, "unknown solve method");
    if (opts->precond != JFM_PREC_BLOCK_JACOBI && opts->precond != JFM_PREC_LEARNED)
        return jfm_fail(ctx, JFM_EINVAL, "the tangent solves take JFM_PREC_BLOCK_JACOBI or JFM_PREC_LEARNED");
    if (opts->precond == JFM_PREC_LEARNED && !ctx->has_mlp)
        return jfm_fail(ctx, JFM_ESTATE, "learned preconditioner requested but no weights are loaded");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = ensure_workspaces(ctx)) return rc;
    if (int rc = nl_buffers(ctx)) return rc;
    const int64_t n = ctx->ndof;
    hipStream_t s = ctx->stream;
    const int check = std::max(2, opts->check_every > 0 ? opts->check_every : 32);
    const double eta_max = opts->eta_max > 0.0 && opts->eta_max < 1.0 ? opts->eta_max : 0.1;
    static const bool verbose = [] {
        const char *v = std::getenv("JFEM_VERBOSE");
        return v && v[0] == '1';
    }();
    jfm_newton_stats st{};
    ctx->hist_it.clear();
    ctx->hist_res.clear();
    hipEvent_t e0, e1;
    JFM_HIP(ctx, hipEventCreate(&e0));
    JFM_HIP(ctx, hipEventCreate(&e1));
    JFM_HIP(ctx, hipEventRecord(e0, s));
    // |f_ext| on the free DOFs
    double fnorm = 0.0;
    {
        int np = 0;
        CHECK_LAUNCH(ctx, launch_residual(ctx, ctx->f_ext, ctx->nl_zero, ctx->fixed, ctx->rr64, n, &np));
        if (int rc = finalize_misc(ctx, np, 1, FIN_DOT)) return rc;
        if (int rc = read_scalars(ctx)) return rc;
        fnorm = std::sqrt(ctx->S_host->tmp0);
    }
    JFM_HIP(ctx, hipMemsetAsync(ctx->nl_u, 0, sizeof(double) * n, s));
    bool all_ok = true;
    double relres = 0.0;
    struct Restore {  // the workspaces end with the tangent block Jacobi: rebuild the linear one on next use
        jfm_ctx *c;
        ~Restore() {
            c->ws64.ok = c->ws32.ok = false;
            c->minv_tangent = false;
            jfem_learned_drop(c);
            c->learned_ready = false;
        }
    } restore{ctx};
    for (int step = 1; step <= opts->load_steps && all_ok; ++step) {
        const double lf = (double)step / opts->load_steps, lf0 = (double)(step - 1) / opts->load_steps;
        // prescribed displacements follow the load factor (g is zero on free DOFs)
        if (ctx->g_nonzero) {
            CHECK_LAUNCH(ctx, (launch_convert<double, double>(ctx, ctx->g, ctx->rr64, n, lf - lf0)));
            CHECK_LAUNCH(ctx, launch_add(ctx, ctx->nl_u, ctx->rr64, ctx->nl_u, n));
        }
        double rn = 0.0;
        if (int rc = nl_residual(ctx, ctx->nl_u, lf, ctx->nl_r, &rn)) return rc;
        if (!std::isfinite(rn))
            return jfm_fail(ctx, JFM_ENOCONV, "inadmissible state (det F <= 0) at the start of a load step");
        const double ref = lf * fnorm > 0.0 ? lf * fnorm : (rn > 0.0 ? rn : 1.0);
        bool ok = false;
        double rn_prev = rn, eta_prev = eta_max;
        bool step_learned = true;  // this load step's probe outcome
        for (int it = 0;; ++it) {
            relres = rn / ref;
            ctx->hist_it.push_back(st.linear_iterations);
            ctx->hist_res.push_back(relres);
            if (rn <= opts->rtol * ref) { ok = true; break; }
            if (it >= opts->max_newton) break;
            // tangent block Jacobi at the current state (features use its diagonal)
            CHECK_LAUNCH(ctx, launch_nh_block_jacobi(ctx, ctx->ustate64, (double *)ctx->ws64.Minv, (float *)ctx->ws32.Minv,
                                                     ctx->Dtan));
            ctx->ws64.ok = ctx->ws32.ok = true;
            ctx->minv_tangent = true;
            const bool coarse = opts->coarse_nodes > 0;
            // coarse level of the tangent, rebuilt at the first iterate of each
            // load step; later iterates keep it (Q A_c^-1 Q' stays SPD, only
            // less exact as K_T moves)
            if (coarse && it == 0)
                if (int rc = jfem_coarse_setup(ctx, JFM_OP_TANGENT, opts->coarse_nodes, nullptr)) return rc;
            const double *M64 = (const double *)ctx->ws64.Minv;
            const float *M32 = (const float *)ctx->ws32.Minv;
            if (opts->precond == JFM_PREC_LEARNED) {
                // The probe runs at the first iterate of each load step, like the
                // coarse rebuild; later iterates of the step keep its outcome.
                // Accepted blocks are recomputed from the new tangent diagonal at
                // every Newton step.
                jfem_learned_drop(ctx);
                ctx->learned_ready = false;
                bool use = false;
                int its = 0;
                if (it == 0 || step_learned) {
                    JFM_HIP(ctx, hipMemcpyAsync(ctx->ws64.b, ctx->nl_r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                    const int probe = it == 0 ? opts->learned_probe : 0;
                    if (int rc = learned_select(ctx, JFM_OP_TANGENT, probe, check, &use, &its, nullptr, coarse))
                        return rc;
                    if (it == 0) step_learned = use;
                }
                st.linear_iterations += its;
                if (use) {
                    M64 = ctx->Mlearn64;
                    M32 = ctx->Mlearn32;
                } else {
                    ++st.fallbacks;
                }
            }
            // Forcing term.  The reference uses min(0.9, |R|^0.5) on the absolute
            // residual (ext/JuliaFEMCUDAExt.jl:785-787), which is unit dependent
            // and near 0.9 for a relative one; here Eisenstat-Walker choice 2,
            // eta = 0.9 (|R_k| / |R_k-1|)^2 with their safeguard, starting each
            // load step at eta_max, capped by eta_max and never tighter than the
            // Newton tolerance needs.
            double eta = eta_max;
            if (it > 0) {
                const double ratio = rn / rn_prev;
                eta = 0.9 * ratio * ratio;
                const double safe = 0.9 * eta_prev * eta_prev;
                if (safe > 0.1) eta = std::max(eta, safe);
                eta = std::min(eta, eta_max);
            }
            eta = std::max(eta, 0.5 * opts->rtol * ref / rn);
            eta_prev = eta;
            rn_prev = rn;
            int lits = 0;
            if (int rc = tangent_solve(ctx, opts, M64, M32, eta, rn, check, coarse, &lits)) return rc;
            st.linear_iterations += lits;
            ++st.newton_iterations;
            if (verbose)
                std::fprintf(stderr, "[jfem] load step %d/%d, Newton %d: relres %.3e, eta %.2e, %d PCG iterations\n",
                             step, opts->load_steps, it + 1, relres, eta, lits);
            // u += du; halve the step while some Gauss point leaves det F > 0
            JFM_HIP(ctx, hipMemcpyAsync(ctx->rr64, ctx->nl_u, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            double rnew = 0.0;
            for (int bt = 0;; ++bt) {
                CHECK_LAUNCH(ctx, launch_add(ctx, ctx->rr64, ctx->nl_du, ctx->nl_u, n));
                if (int rc = nl_residual(ctx, ctx->nl_u, lf, ctx->nl_r, &rnew)) return rc;
                if (std::isfinite(rnew) || bt >= 12) break;
                CHECK_LAUNCH(ctx, (launch_convert<double, double>(ctx, ctx->nl_du, ctx->nl_du, n, 0.5)));
            }
            if (!std::isfinite(rnew)) break;
            rn = rnew;
        }
        all_ok = ok;
    }
    JFM_HIP(ctx, hipEventRecord(e1, s));
    JFM_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    JFM_HIP(ctx, hipMemcpyAsync(u, ctx->nl_u, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (int rc = sync(ctx)) return rc;
    st.converged = all_ok ? 1 : 0;
    st.relres = relres;
    st.solve_ms = ms;
    if (stats) *stats = st;
    if (!all_ok) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "Newton did not reach rtol %.3e: relres %.3e after %d iterations", opts->rtol,
                      relres, st.newton_iterations);
        return jfm_fail(ctx, JFM_ENOCONV, msg);
    }
    return JFM_OK;
}

// ========================================================= postprocessing ==
extern "C" int jfm_postprocess(jfm_ctx *ctx, const double *u, double *strain_gp, double *stress_gp,
                               double *strain_nodal, double *stress_nodal) {
    if (!ctx || !u) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = ensure_workspaces(ctx)) return rc;
    const int64_t n = ctx->ndof, N = ctx->N, nq = ctx->n_el * ctx->ngp * 6;
    hipStream_t s = ctx->stream;
    double *du = jfm_alloc<double>(ctx, n), *eps = jfm_alloc<double>(ctx, nq), *sig = jfm_alloc<double>(ctx, nq),
           *Mm = jfm_alloc<double>(ctx, 9 * N);
    auto cleanup = [&] {
        for (void *p : {(void *)du, (void *)eps, (void *)sig, (void *)Mm}) jfm_free(ctx, p);
    };
    if (!du || !eps || !sig || !Mm) {
        cleanup();
        return jfm_fail(ctx, JFM_ENOMEM, "device allocation failed in jfm_postprocess");
    }
    int rc = JFM_OK;
    auto fail_hip = [&](hipError_t e, const char *what) {
        rc = jfm_hip_fail(ctx, e, what);
        return rc;
    };
    hipError_t e = hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_post_gp(ctx, du, eps, sig);
    if (e != hipSuccess) fail_hip(e, "launch_post_gp");
    if (rc == JFM_OK && strain_gp) {
        e = hipMemcpyAsync(strain_gp, eps, sizeof(double) * nq, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) fail_hip(e, "strain copy");
    }
    if (rc == JFM_OK && stress_gp) {
        e = hipMemcpyAsync(stress_gp, sig, sizeof(double) * nq, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) fail_hip(e, "stress copy");
    }
    // nodal least-squares fits (src/problems_elasticity.jl:550-594): M x = b
    // per three components, PCG with the mass operator and its diagonal
    if (rc == JFM_OK && (strain_nodal || stress_nodal)) {
        e = launch_mass_diag(ctx, Mm);
        if (e != hipSuccess) fail_hip(e, "launch_mass_diag");
        std::vector<double> part(n);
        double *outs[2] = {strain_nodal, stress_nodal};
        for (int f = 0; f < 2 && rc == JFM_OK; ++f) {
            if (!outs[f]) continue;
            for (int comp0 = 0; comp0 < 6 && rc == JFM_OK; comp0 += 3) {
                e = launch_post_rhs(ctx, du, f, comp0, (double *)ctx->ws64.b);
                if (e != hipSuccess) { fail_hip(e, "launch_post_rhs"); break; }
                int it = 0, status = 0;
                if ((rc = pcg_run<double>(ctx, kOpMass, Mm, 1e-13, 2000, 32, &it, &status))) break;
                e = hipMemcpyAsync(part.data(), ctx->ws64.x, sizeof(double) * n, hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) { fail_hip(e, "nodal copy"); break; }
                for (int64_t a = 0; a < N; ++a)
                    for (int c = 0; c < 3; ++c) outs[f][6 * a + comp0 + c] = part[3 * a + c];
            }
        }
    }
    (void)hipStreamSynchronize(s);
    cleanup();
    return rc;
}

// ================================================================== heat ==
extern "C" int jfm_heat_solve(jfm_ctx *ctx, const double *k, const double *q, int64_t n_faces, int32_t nodes_per_face,
                              const int64_t *face_conn, const double *flux, const double *h, const double *T_ext,
                              const uint8_t *fixed_node, const double *T_fixed, int32_t index_base, double rtol,
                              int32_t maxit, double *T, jfm_solve_stats *stats) {
    if (!ctx || !k || !T || !(rtol > 0.0) || maxit <= 0) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (n_faces < 0 || (n_faces > 0 && !face_conn)) return jfm_fail(ctx, JFM_EINVAL, "invalid face arguments");
    if (n_faces > 0 && nodes_per_face != 6 && nodes_per_face != 3 && nodes_per_face != 4)
        return jfm_fail(ctx, JFM_EUNSUPPORTED, "nodes_per_face must be 6 (Tri6), 3 (Tri3) or 4 (Quad4)");
    if ((h == nullptr) != (T_ext == nullptr)) return jfm_fail(ctx, JFM_EINVAL, "convection needs both h and T_ext");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = ensure_workspaces(ctx)) return rc;
    const int64_t N = ctx->N, n = ctx->ndof, n_el = ctx->n_el;
    hipStream_t s = ctx->stream;
    for (int64_t e = 0; e < n_el; ++e)
        if (!(k[e] > 0.0)) return jfm_fail(ctx, JFM_EINVAL, "thermal conductivity must be positive");
    // the captured PCG graphs of a previous heat problem point at its buffers
    drop_graphs(ctx);
    release_heat(ctx);
    ALLOC(ctx, ctx->heat_k, double, n_el);
    ALLOC(ctx, ctx->heat_fixed, uint8_t, n);
    ALLOC(ctx, ctx->heat_g, double, n);
    ALLOC(ctx, ctx->heat_Minv, double, 9 * N);
    JFM_HIP(ctx, hipMemcpyAsync(ctx->heat_k, k, sizeof(double) * n_el, hipMemcpyHostToDevice, s));
    std::vector<uint8_t> fx(n, 1);
    std::vector<double> g(n, 0.0);
    for (int64_t a = 0; a < N; ++a) {
        fx[3 * a] = fixed_node ? fixed_node[a] : 0;
        if (fx[3 * a] && T_fixed) g[3 * a] = T_fixed[a];
    }
    JFM_HIP(ctx, hipMemcpyAsync(ctx->heat_fixed, fx.data(), n, hipMemcpyHostToDevice, s));
    JFM_HIP(ctx, hipMemcpyAsync(ctx->heat_g, g.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    double *src = nullptr;
    if (q) {
        src = jfm_alloc<double>(ctx, n_el);
        if (!src) return jfm_fail(ctx, JFM_ENOMEM, "device allocation failed in jfm_heat_solve");
        JFM_HIP(ctx, hipMemcpyAsync(src, q, sizeof(double) * n_el, hipMemcpyHostToDevice, s));
    }
    if (n_faces > 0) {
        const int nfn = nodes_per_face;
        std::vector<int32_t> soa((size_t)n_faces * nfn), fptr(N + 1, 0);
        for (int64_t f = 0; f < n_faces; ++f)
            for (int i = 0; i < nfn; ++i) {
                const int64_t v = face_conn[f * nfn + i] - index_base;
                if (v < 0 || v >= N) {
                    if (src) jfm_free(ctx, src);
                    return jfm_fail(ctx, JFM_EINVAL, "face connectivity out of range");
                }
                soa[(size_t)i * n_faces + f] = (int32_t)v;
                fptr[v + 1]++;
            }
        for (int64_t a = 0; a < N; ++a) fptr[a + 1] += fptr[a];
        std::vector<int32_t> inc((size_t)n_faces * nfn), pos(fptr.begin(), fptr.end() - 1);
        for (int64_t f = 0; f < n_faces; ++f)  // ascending face order per node
            for (int i = 0; i < nfn; ++i) inc[pos[soa[(size_t)i * n_faces + f]]++] = (int32_t)(f * nfn + i);
        ctx->hf_n = n_faces;
        ctx->hf_nfn = nfn;
        ALLOC(ctx, ctx->hf_conn, int32_t, n_faces * nfn);
        ALLOC(ctx, ctx->hf_ptr, int32_t, N + 1);
        ALLOC(ctx, ctx->hf_inc, int32_t, n_faces * nfn);
        ALLOC(ctx, ctx->hf_fe, double, n_faces * nfn * 3);
        JFM_HIP(ctx, hipMemcpyAsync(ctx->hf_conn, soa.data(), 4 * soa.size(), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->hf_ptr, fptr.data(), 4 * (N + 1), hipMemcpyHostToDevice, s));
        JFM_HIP(ctx, hipMemcpyAsync(ctx->hf_inc, inc.data(), 4 * inc.size(), hipMemcpyHostToDevice, s));
        for (auto [dst, v] : {std::pair<double **, const double *>{&ctx->hf_g, flux}, {&ctx->hf_h, h}, {&ctx->hf_text, T_ext}})
            if (v) {
                ALLOC(ctx, *dst, double, n_faces);
                JFM_HIP(ctx, hipMemcpyAsync(*dst, v, sizeof(double) * n_faces, hipMemcpyHostToDevice, s));
            }
    }
    // the heat row mask stands in for the elasticity one for this solve
    struct MaskSwap {
        jfm_ctx *c;
        uint8_t *saved;
        ~MaskSwap() { c->fixed = saved; }
    } swap{ctx, ctx->fixed};
    ctx->fixed = ctx->heat_fixed;
    hipEvent_t e0, e1;
    JFM_HIP(ctx, hipEventCreate(&e0));
    JFM_HIP(ctx, hipEventCreate(&e1));
    JFM_HIP(ctx, hipEventRecord(e0, s));
    jfm_solve_stats st{};
    // b = f - K g on the free temperatures
    CHECK_LAUNCH(ctx, launch_heat_load(ctx, src, ctx->vx64));
    double *b = (double *)ctx->ws64.b;
    double bnorm = 0.0;
    if (int rc = true_residual(ctx, kOpHeat, ctx->vx64, ctx->heat_g, b, &bnorm)) return rc;
    CHECK_LAUNCH(ctx, launch_heat_jacobi(ctx, ctx->heat_Minv));
    const int check = 32;
    if (bnorm > 0.0) {
        int it = 0, status = 0, total = 0;
        if (int rc = pcg_run<double>(ctx, kOpHeat, ctx->heat_Minv, rtol, maxit, check, &it, &status)) return rc;
        total = it;
        for (int restart = 0; restart < 4 && total < maxit; ++restart) {  // residual replacement, as jfm_solve
            double rn = 0.0;
            if (int rc = true_residual(ctx, kOpHeat, b, (const double *)ctx->ws64.x, ctx->rr64, &rn)) return rc;
            if (rn <= rtol * bnorm) break;
            if (int rc = pcg_run<double>(ctx, kOpHeat, ctx->heat_Minv, rtol, maxit - total, check, &it, &status, true))
                return rc;
            total += it;
            st.outer++;
        }
        st.iterations = total;
        CHECK_LAUNCH(ctx, launch_add(ctx, ctx->heat_g, (const double *)ctx->ws64.x, ctx->u64, n));
    } else {
        JFM_HIP(ctx, hipMemcpyAsync(ctx->u64, ctx->heat_g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    double rn = 0.0;
    if (int rc = true_residual(ctx, kOpHeat, ctx->vx64, ctx->u64, ctx->rr64, &rn)) return rc;
    JFM_HIP(ctx, hipEventRecord(e1, s));
    JFM_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    JFM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<double> u(n);
    JFM_HIP(ctx, hipMemcpyAsync(u.data(), ctx->u64, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (int rc = sync(ctx)) return rc;
    if (src) jfm_free(ctx, src);
    for (int64_t a = 0; a < N; ++a) T[a] = u[3 * a];
    st.relres = bnorm > 0.0 ? rn / bnorm : 0.0;
    st.converged = st.relres <= rtol ? 1 : 0;
    st.solve_ms = ms;
    if (stats) *stats = st;
    if (!st.converged) return jfm_fail(ctx, JFM_ENOCONV, "heat solve did not reach rtol");
    return JFM_OK;
}
