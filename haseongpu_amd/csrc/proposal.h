// proposal.h - the learned proposal's network layout, its input features and
// the host hooks the evaluation loop calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/hase_ase.h"
#include "prism_walk.h"

struct hase_context;

namespace hase {

constexpr int kMlpIn = HASE_MLP_INPUTS;   // real inputs
constexpr int kMlpH = HASE_MLP_HIDDEN;    // hidden width
constexpr int kMlpHidden = 4;             // hidden layers

// Offsets of the FP32 parameter vector: W1 b1 W2 b2 W3 b3 W4 b4 w5 b5,
// every W row-major [out][in].
constexpr size_t kOffW1 = 0;
constexpr size_t kOffB1 = kOffW1 + (size_t)kMlpH * kMlpIn;
constexpr size_t kOffW2 = kOffB1 + kMlpH;
constexpr size_t kOffB2 = kOffW2 + (size_t)kMlpH * kMlpH;
constexpr size_t kOffW3 = kOffB2 + kMlpH;
constexpr size_t kOffB3 = kOffW3 + (size_t)kMlpH * kMlpH;
constexpr size_t kOffW4 = kOffB3 + kMlpH;
constexpr size_t kOffB4 = kOffW4 + (size_t)kMlpH * kMlpH;
constexpr size_t kOffW5 = kOffB4 + kMlpH;
constexpr size_t kOffB5 = kOffW5 + kMlpH;
constexpr size_t kMlpParams = kOffB5 + 1;

// BF16 operand fragments of the MFMA inference kernel: 1 KB each (64 lanes x
// 8 values); layer 1 has 4 (output blocks of 32 neurons, one K block of 16
// padded inputs), each hidden layer 4 x 8.
constexpr int kFragsL1 = 4;
constexpr int kFragsHidden = 32;
constexpr int kFrags = kFragsL1 + 3 * kFragsHidden;
// FP32 vector block: b1..b4 [4][128], w5 [128], b5
constexpr int kMlpVec = 4 * kMlpH + kMlpH + 1;

// Unnormalised network inputs of one (sample image, prism) pair.  Lengths in
// cm; the sample's z is its (possibly mirrored) plane times the layer
// thickness, the prism's z its mid-layer height.
HASE_HD inline void proposal_inputs(float sx, float sy, float sz, float cx, float cy, float level, float beta,
                                    float h, float log_sigma, float f[kMlpIn])
{
    const float cz = (level + 0.5f) * h;
    const float dx = sx - cx, dy = sy - cy, dz = sz - cz;
    f[0] = sx;
    f[1] = sy;
    f[2] = sz;
    f[3] = cx;
    f[4] = cy;
    f[5] = cz;
    f[6] = level;
    f[7] = beta;
    f[8] = sqrtf(dx * dx + dy * dy + dz * dz);
    f[9] = log_sigma;
}

// Packs FP32 parameters into the inference kernel's fragments and vector block.
void pack_mlp_bf16(const float* params, uint4* frags /*[kFrags*64]*/, float* vec /*[kMlpVec]*/);

// Cluster table of the learned proposal from the context's current beta.
void build_clusters(hase_context* ctx);

// Network rows (slot, slice, cluster) for samples [sample_base, sample_base +
// slots) into logits[(slot*R + k)*C + c] on the context stream.
void launch_proposal_inference(hase_context* ctx, uint32_t sample_base, uint32_t slots, float* logits);

}  // namespace hase
