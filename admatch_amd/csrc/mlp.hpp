// Learned step selector: GPU training / inference entry points (mlp.hip).
#pragma once

#include <cstdint>
#include <string>

namespace admatch {

// Gradient of the masked MSE loss at weights w over n samples (features
// x[n][16], targets y[n][3], mask[n][3]); grad_out[5443], loss_out[1].
int mlp_grad(const float* x, const float* y, const float* mask, int64_t n, const float* w, float* grad_out,
             float* loss_out, std::string& err);
// out[n][3] = MLP(x) at weights w.
int mlp_forward(const float* x, int64_t n, const float* w, float* out, std::string& err);
// Adam on shuffled minibatches; w_inout holds the initial weights and
// receives the trained ones; loss_hist[epochs] gets the mean batch loss.
int mlp_train(const float* x, const float* y, const float* mask, int64_t n, int32_t epochs, int32_t batch, float lr,
              uint64_t seed, float* w_inout, float* loss_hist, std::string& err);

}  // namespace admatch
