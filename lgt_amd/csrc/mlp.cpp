// Small multilayer perceptron trained and evaluated on the GPU: the learned
// warm start of the ALS bond solve.  Layers in -> W -> W -> out with ReLU on
// the two hidden layers and a linear output; the batch runs along columns.
// Products use the engine's GEMM in the FP32 MFMA tier (FP64 storage),
// everything else the mlp_* kernels.  Loss: L = 1/(2B) sum (Y - T)^2.
#include "mlp.h"

#include <cmath>

namespace peps {

Mlp::Mlp(Ctx& ctx, int n_in, int width, int n_out, uint64_t seed) : c(&ctx), nin(n_in), w(width), nout(n_out) {
  PEPS_CHECK(n_in >= 1 && width >= 1 && n_out >= 1 && width <= 8192 && n_in <= 8192 && n_out <= 8192, PS_EARG,
             "mlp: layer sizes out of range");
  const int rows[3] = {width, width, n_out}, cols[3] = {n_in, width, width};
  for (int l = 0; l < 3; ++l) {
    W[l] = c->alloc({rows[l], cols[l]}, false);
    fill_random(*c, W[l], seed + 977u * uint64_t(l + 1));
    // uniform(-a, a) with a = sqrt(6 / fan_in): He initialisation for ReLU
    scale(*c, W[l], std::sqrt(6.0 / cols[l]));
    b[l] = c->zeros({rows[l]}, false);
    mW[l] = c->zeros({rows[l], cols[l]}, false);
    vW[l] = c->zeros({rows[l], cols[l]}, false);
    mb[l] = c->zeros({rows[l]}, false);
    vb[l] = c->zeros({rows[l]}, false);
  }
}

int64_t Mlp::param_count() const {
  int64_t n = 0;
  for (int l = 0; l < 3; ++l) n += W[l].numel() + b[l].numel();
  return n;
}

Tensor Mlp::forward(const Tensor& X, Tensor* H1, Tensor* H2) {
  PEPS_CHECK(X.rank() == 2 && X.dims[0] == nin && !X.cplx, PS_ESHAPE, "mlp: input must be (n_in x batch)");
  TierGuard fp32(*c, TIER_FP32);
  Tensor h1 = matmul(*c, W[0], X);
  mlp_bias_act(*c, h1, b[0], true);
  Tensor h2 = matmul(*c, W[1], h1);
  mlp_bias_act(*c, h2, b[1], true);
  Tensor y = matmul(*c, W[2], h2);
  mlp_bias_act(*c, y, b[2], false);
  if (H1) *H1 = h1;
  if (H2) *H2 = h2;
  return y;
}

double Mlp::gradients(const Tensor& X, const Tensor& T, Tensor gW[3], Tensor gb[3]) {
  const int64_t B = X.dims[1];
  PEPS_CHECK(T.rank() == 2 && T.dims[0] == nout && T.dims[1] == B && B >= 1, PS_ESHAPE,
             "mlp: targets must be (n_out x batch)");
  Tensor h1, h2;
  Tensor y = forward(X, &h1, &h2);
  TierGuard fp32(*c, TIER_FP32);
  Tensor d3 = c->alloc(y.dims, false);
  mlp_scaled_diff(*c, y, T, d3, 1.0);
  const double loss = 0.5 * norm2sq(*c, d3) / double(B);
  scale(*c, d3, 1.0 / double(B));  // dL/dY
  gW[2] = matmul(*c, d3, h2, false, false, false, true);  // d3 h2^T
  gb[2] = c->alloc({nout}, false);
  mlp_row_sum(*c, d3, gb[2]);
  Tensor d2 = matmul(*c, W[2], d3, false, false, true, false);  // W3^T d3
  mlp_relu_grad(*c, d2, h2);
  gW[1] = matmul(*c, d2, h1, false, false, false, true);
  gb[1] = c->alloc({w}, false);
  mlp_row_sum(*c, d2, gb[1]);
  Tensor d1 = matmul(*c, W[1], d2, false, false, true, false);
  mlp_relu_grad(*c, d1, h1);
  gW[0] = matmul(*c, d1, X, false, false, false, true);
  gb[0] = c->alloc({w}, false);
  mlp_row_sum(*c, d1, gb[0]);
  return loss;
}

double Mlp::train_step(const Tensor& X, const Tensor& T, double lr) {
  Tensor gW[3], gb[3];
  const double loss = gradients(X, T, gW, gb);
  ++t;
  for (int l = 0; l < 3; ++l) {
    mlp_adam(*c, W[l], gW[l], mW[l], vW[l], lr, 0.9, 0.999, 1e-8, t);
    mlp_adam(*c, b[l], gb[l], mb[l], vb[l], lr, 0.9, 0.999, 1e-8, t);
  }
  return loss;
}

void Mlp::get_params(double* out) {
  int64_t off = 0;
  for (int l = 0; l < 3; ++l) {
    c->download(W[l], out + off);
    off += W[l].numel();
    c->download(b[l], out + off);
    off += b[l].numel();
  }
}

void Mlp::set_params(const double* in) {
  int64_t off = 0;
  for (int l = 0; l < 3; ++l) {
    W[l] = c->upload(W[l].dims, false, in + off);
    off += W[l].numel();
    b[l] = c->upload(b[l].dims, false, in + off);
    off += b[l].numel();
  }
}

}  // namespace peps
