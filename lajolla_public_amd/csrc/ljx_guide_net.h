// Host-side state of the guiding network and its training step, shared by
// the network's own entry points and the guided render loop.
#pragma once

#include "../../include/lajolla_amd.h"
#include "ljx_guide.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace ljx {
namespace guide {

struct Net {
    int device = 0;
    hipStream_t stream = nullptr;
    ljx_guide_config cfg{};
    // parameters and Adam state (FP32)
    float *master = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    int steps = 0;
    // BF16 operand fragments and FP32 biases rebuilt from the master copy
    bf16x8 *ffrag = nullptr;   // kFwdFrags * 64
    bf16x8 *bfrag = nullptr;   // kBwdFrags * 64
    float *fbias = nullptr;    // kBiasFwd
    int *fmap = nullptr, *bmap = nullptr, *biasmap = nullptr;
    // training scratch: activations and deltas of up to batch_size samples in
    // 32-sample groups, per-block partial gradients, reduced gradient
    uint16_t *acts = nullptr;
    float *partial = nullptr;
    float *grad = nullptr;
    float *red = nullptr;      // weighted loss sum, weight sum
    int batch_cap = 0;
    int partial_blocks = 0;
    int num_cu = 0;
};

Net *net_create(int device, const ljx_guide_config &cfg);
void net_destroy(Net *n);
void net_set_params(Net &n, const float *host_params);
// One forward + backward + gradient reduction over n samples on stream s;
// with update, an Adam step follows and the fragments are rebuilt.
// Returns the normalised loss (synchronises the stream) if loss_out is set.
void net_step(Net &n, const float4 *feats, const float4 *dirs, int count, bool update, hipStream_t s,
              double *loss_out);

}  // namespace guide
}  // namespace ljx
