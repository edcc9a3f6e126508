// jfem_learned.h -- learned block preconditioner: weight blob and hooks.
//
// Per-node network 24 -> 64 -> 64 -> 6 with GELU, BF16 weights and FP32
// accumulation.  Its 6 outputs o fill a lower-triangular L,
//     L = [[e^o0, 0, 0], [o1, e^o2, 0], [o3, o4, e^o5]]   (exponents clamped to +-5),
// and the node's preconditioner block is
//     M_a = C_a (L L' + eps I) C_a',   C_a C_a' = inv(D_a),
// with D_a the node's 3x3 diagonal block of K, so M_a is SPD for any weights
// and L = I reproduces block Jacobi up to the factor (1 + eps).  Nodes with a
// fixed DOF keep the block-Jacobi block.
#pragma once

#include <cstdint>

#include "jfem_internal.h"

namespace jfem {

constexpr int kFeat = 24, kHid = 64, kOut = 6;
constexpr uint32_t kBlobMagic = 0x4C4D464Au;  // "JFML"
// trainable parameters in gradient order: W1, b1, W2, b2, W3, b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kFeat, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kOut * kHid, kNParam = kOffB3 + kOut;

// Blob layout (little endian, no padding): every field in this order.
struct LearnedBlob {
    uint32_t magic, version, n_in, n_h1, n_h2, n_out;
    uint16_t W1[kHid * kFeat];   // bf16, row-major [out][in]
    uint16_t W2[kHid * kHid];
    uint16_t W3[kOut * kHid];
    float b1[kHid], b2[kHid], b3[kOut];
    float feat_shift[kFeat], feat_scale[kFeat];  // x_norm = (x - shift) * scale
    float eps;
};

// device launchers (learned.hip)
// (on stream s)
hipError_t launch_features(jfm_ctx *ctx, double *ef, float *feat, hipStream_t s);
hipError_t launch_mlp(jfm_ctx *ctx, const float *feat, const void *blob, float *out, const double *G,
                      const double *Minv, float *gpart, int *nwaves, hipStream_t s);
hipError_t launch_mblocks(jfm_ctx *ctx, const float *out, const double *Minv, float eps, double *M64, float *M32,
                          hipStream_t s);
}  // namespace jfem

// Builds features, runs the network and forms Mlearn64/Mlearn32 for the
// current mesh and boundary conditions.  *ms receives the device time.  If
// jfem_learned_launch started that work, joins it instead (the main stream
// waits for it).
int jfem_learned_prepare(jfm_ctx *ctx, double *ms);
// Starts the same work on a second stream after the work queued so far on
// the main stream, so that it runs beside what the main stream does next
// (the coarse setup); jfem_learned_prepare joins it.  No-op when the blocks
// are current or no weights are loaded.
int jfem_learned_launch(jfm_ctx *ctx);
// Waits for a started, not yet joined, learned launch and drops it.
void jfem_learned_drop(jfm_ctx *ctx);
