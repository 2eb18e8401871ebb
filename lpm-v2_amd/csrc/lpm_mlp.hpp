// lpm_mlp.hpp -- the far-field surrogate network object (device resident).
#pragma once

#include <stdint.h>

#include <vector>

#include "lpm_internal.hpp"

namespace lpm {

int64_t mlp_param_count(int nin, int width, int nhidden, int nout);

struct Mlp {
  lpm_ctx* ctx = nullptr;
  int nin = 0, width = 0, nhidden = 0, nout = 0;
  int act_kind = 1;  // hidden-layer activation: 1 tanh, 2 ISRU
  int64_t np = 0;
  int adam_t = 0;
  int cur_n = 0;
  DevBuf p, grad, m, v;    // FP32 parameters, gradient, Adam moments
  DevBuf frag, bias;       // FP16 MFMA fragments + FP32 biases (inference)
  DevBuf act, pre, scratch, loss_part, splitk;
  DevBuf xb, tb, idx;      // training minibatch staging
  ~Mlp();
  void release();
  int init(lpm_ctx* c, int nin, int width, int nhidden, int nout, int act, uint64_t seed);
  bool mfma_ok() const;
  int pack();  // FP32 params -> FP16 fragments
  int ensure_batch(int n);
  int forward_f32(int n, const float* X);  // device X [n][nin]; output in act slot nhidden
  int loss_grad(int n, const float* X, const float* T, double* loss);
  int adam(float lr);
  // cosine: learning rate lr (1 + cos(pi epoch / epochs)) / 2 per epoch, else
  // constant; targets are multiplied by tscale as they are read
  int train(int n, const float* X, const float* T, int epochs, int batch, float lr, uint64_t seed,
            double* final_loss, std::vector<double>* history, bool cosine = false, float tscale = 1.0f);
  int scale_output(float s);                      // last layer W, b *= s
  int rms(int64_t n, const float* v, double* out);  // sqrt(mean v^2) of device values
  int forward_mfma(int n, const float* X, float* Y);
};

}  // namespace lpm
