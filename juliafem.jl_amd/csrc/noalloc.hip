// noalloc.hip -- jfm_apply_noalloc: the linear K.u on device vectors with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, fixed-memory serving loops).
//
// The two-phase operator's element results (3 npe n_el values per K.u) and,
// in FP32, the converted input and output vectors live in `scratch`, one
// file-scope device array per GPU of kScratchBytes.  Host code never passes
// the array's name to a kernel: jfem_noalloc_init() looks up its device
// address once per device (with that device current) and every launch uses
// that address.  All calls on one GPU share the array and every call writes
// each value it later reads (each incidence slot is written by the element
// phase before the gather reads it; the FP32 vectors are written by the
// conversion), so nothing needs clearing between calls.  Calls that overlap
// on one GPU (two contexts, two streams) race on the array.
#include <mutex>

#include "jfem_internal.h"

using namespace jfem;

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

// Called by jfm_create with ctx->device current.
int jfem_noalloc_init(jfm_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return jfm_fail(ctx, JFM_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[ctx->device]) return JFM_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return jfm_hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
    g_scratch[ctx->device] = (unsigned char *)p;
    return JFM_OK;
}

extern "C" int jfm_apply_noalloc(jfm_ctx *ctx, int32_t op, int32_t precision, const double *x, double *y,
                                 void *stream) {
    if (!ctx || !x || !y) return JFM_EINVAL;
    if (!ctx->has_mesh) return jfm_fail(ctx, JFM_ESTATE, "no mesh");
    if (op != JFM_OP_MATFREE && op != JFM_OP_EBE)
        return jfm_fail(ctx, JFM_EUNSUPPORTED, "jfm_apply_noalloc supports JFM_OP_MATFREE and JFM_OP_EBE");
    if (op == JFM_OP_EBE && !ctx->Ke)
        return jfm_fail(ctx, JFM_ESTATE,
                        "JFM_OP_EBE needs the element matrices on the device: call jfm_apply or jfm_assemble with it once");
    if (precision != JFM_FP64 && precision != JFM_FP32)
        return jfm_fail(ctx, JFM_EINVAL, "precision must be FP64 or FP32");
    const int64_t n = ctx->ndof;
    const size_t slots = (size_t)3 * ctx->npe * ctx->n_el;
    const size_t need = precision == JFM_FP64 ? sizeof(double) * slots
                                              : 2 * align256(sizeof(float) * n) + sizeof(float) * slots;
    if (need > kScratchBytes)
        return jfm_fail(ctx, JFM_EINVAL,
                        "mesh needs " + std::to_string(need) + " bytes of scratch; jfm_apply_noalloc holds " +
                            std::to_string(kScratchBytes) + " (use jfm_apply)");
    JFM_HIP(ctx, hipSetDevice(ctx->device));
    unsigned char *base = g_scratch[ctx->device];
    if (!base) return jfm_fail(ctx, JFM_ESTATE, "scratch address not set up for this device");

    // launch on the caller's stream (null: the context's)
    const hipStream_t saved = ctx->stream;
    if (stream) ctx->stream = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (precision == JFM_FP64) {
        double *ye = (double *)base;
        e = launch_elem_into<double>(ctx, op, x, ye);
        if (e == hipSuccess) e = launch_gather<double>(ctx, ye, ctx->fixed, y, nullptr, nullptr, nullptr, nullptr);
    } else {
        float *x32 = (float *)base;
        float *y32 = (float *)(base + align256(sizeof(float) * n));
        float *ye = (float *)(base + 2 * align256(sizeof(float) * n));
        e = launch_convert<double, float>(ctx, x, x32, n, 1.0);
        if (e == hipSuccess) e = launch_elem_into<float>(ctx, op, x32, ye);
        if (e == hipSuccess) e = launch_gather<float>(ctx, ye, ctx->fixed, y32, nullptr, nullptr, nullptr, nullptr);
        if (e == hipSuccess) e = launch_convert<float, double>(ctx, y32, y, n, 1.0);
    }
    ctx->stream = saved;
    if (e != hipSuccess) return jfm_hip_fail(ctx, e, "jfm_apply_noalloc launch");
    return JFM_OK;
}
