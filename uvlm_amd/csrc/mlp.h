// mlp.h -- far-field correction network: FP32 training on the device
// (forward, backward, Adam) and conversion to the BF16 inference copy.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "farfield.h"
#include "mlp_dev.h"

namespace uvlm {

// Bytes of the inference copy (BF16 weights, FP32 biases) and its views.
size_t mlp_dev_bytes();
MlpDev mlp_dev_view(const void* base);
// FP32 parameters (MlpShape layout, device) -> inference copy (device).
void launch_mlp_to_dev(hipStream_t s, const float* w, void* dev);
// BF16 inference of n samples (features [n][32] -> y [n][kMlpOut]), device arrays.
void launch_mlp_infer(hipStream_t s, MlpDev net, const float* X, int n, float* Y);

class MlpTrainer {
 public:
  ~MlpTrainer();
  // FP32 forward of n samples: Y [n][kMlpOut] (device arrays).
  int forward(hipStream_t s, const float* w, const float* X, int n, float* Y, std::string* err);
  // Loss 1/2 mean |y - V|^2 and its gradient (MlpShape layout) over n
  // samples; g and loss are device pointers.
  int grad(hipStream_t s, const float* w, const float* X, const float* V, int n, float* g,
           float* loss, std::string* err);
  // Adam on mini-batches of `batch` rows drawn without replacement per
  // epoch (seeded permutation) from the n samples; w updated in place;
  // losses (host, nullable) receives the loss of every iteration.
  int train(hipStream_t s, float* w, const float* X, const float* V, int n, int iters, int batch,
            float lr, unsigned seed, std::vector<float>* losses, std::string* err);

 private:
  int ensure(int rows, std::string* err);
  void fwd(hipStream_t s, const float* w, const float* X, int B);
  void bwd(hipStream_t s, const float* w, const float* X, const float* V, int B, float* g,
           float* loss);

  int cap_ = 0;
  float *Z1_ = nullptr, *Z2_ = nullptr, *Z3_ = nullptr, *H1_ = nullptr, *H2_ = nullptr,
        *H3_ = nullptr, *Y_ = nullptr, *dY_ = nullptr, *D3_ = nullptr, *D2_ = nullptr,
        *D1_ = nullptr, *WT_ = nullptr, *P_ = nullptr, *lrow_ = nullptr, *Xb_ = nullptr,
        *Vb_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr, *loss_ = nullptr;
  int* idx_ = nullptr;
  size_t idx_cap_ = 0;
};

}  // namespace uvlm
