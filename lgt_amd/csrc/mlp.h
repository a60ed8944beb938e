// Small multilayer perceptron on the GPU (learned ALS warm start).
#pragma once
#include "peps_internal.h"

namespace peps {

struct Mlp {
  Ctx* c;
  int nin, w, nout;
  Tensor W[3], b[3];            // W[l] (rows x cols), b[l] (rows)
  Tensor mW[3], vW[3], mb[3], vb[3];  // Adam moments
  int64_t t = 0;                // Adam step count

  Mlp(Ctx& ctx, int n_in, int width, int n_out, uint64_t seed);
  int64_t param_count() const;
  // Y = W3 relu(W2 relu(W1 X + b1) + b2) + b3; hidden activations optional
  Tensor forward(const Tensor& X, Tensor* H1 = nullptr, Tensor* H2 = nullptr);
  // loss 1/(2B) ||Y - T||^2 and its gradients
  double gradients(const Tensor& X, const Tensor& T, Tensor gW[3], Tensor gb[3]);
  // one Adam step (beta 0.9 / 0.999, eps 1e-8); returns the loss before it
  double train_step(const Tensor& X, const Tensor& T, double lr);
  // parameters flattened as W1, b1, W2, b2, W3, b3 (each column-major)
  void get_params(double* out);
  void set_params(const double* in);
};

}  // namespace peps
