// learned.hip -- learned 3x3 block preconditioner (forward path).
#include <cstring>

#include "jfem_device.h"
#include "jfem_learned.h"

using namespace jfem;

extern "C" size_t jfm_learned_blob_bytes(void) { return sizeof(LearnedBlob); }

extern "C" int jfm_learned_set_weights(jfm_ctx *ctx, const void *blob, size_t bytes) {
    if (!ctx || !blob) return JFM_EINVAL;
    if (bytes != sizeof(LearnedBlob)) return jfm_fail(ctx, JFM_EINVAL, "weight blob has the wrong size");
    const LearnedBlob *b = (const LearnedBlob *)blob;
    if (b->magic != kBlobMagic || b->n_in != kFeat || b->n_h1 != kHid || b->n_h2 != kHid || b->n_out != kOut)
        return jfm_fail(ctx, JFM_EINVAL, "weight blob header does not match the 24-64-64-6 network");
    ctx->mlp_host.assign((const unsigned char *)blob, (const unsigned char *)blob + bytes);
    ctx->has_mlp = true;
    ctx->learned_ready = false;
    return JFM_OK;
}

extern "C" int jfm_learned_get_weights(jfm_ctx *ctx, void *blob, size_t bytes) {
    if (!ctx || !blob) return JFM_EINVAL;
    if (!ctx->has_mlp) return jfm_fail(ctx, JFM_ESTATE, "no weights loaded");
    if (bytes != sizeof(LearnedBlob)) return jfm_fail(ctx, JFM_EINVAL, "weight blob has the wrong size");
    std::memcpy(blob, ctx->mlp_host.data(), bytes);
    return JFM_OK;
}

int jfem_learned_prepare(jfm_ctx *ctx, double *ms) {
    (void)ms;
    if (!ctx->has_mlp) return jfm_fail(ctx, JFM_ESTATE, "learned preconditioner requested but no weights are loaded");
    return jfm_fail(ctx, JFM_EUNSUPPORTED, "learned preconditioner forward pass not built");
}
