// mlp.hpp -- per-vertex MLP surrogate (mlp.hip).
#pragma once

#include <vector>

#include "internal.hpp"

namespace nosh {

struct Mlp {
  int nin = 0, width = 0, depth = 0, nout = 0;
  std::vector<int> dims;           // nin, width x depth, nout
  std::vector<int64_t> woff, boff; // per layer offsets into the parameter vector
  int64_t nweights = 0, nparams = 0;
  int steps = 0;
  // Parameters: weights (layer-major, each [out][in]) then biases; FP32
  // masters, Adam moments and gradients; BF16 copy of the weights.
  DBuf<float> master, grad, m1, m2;
  DBuf<__bf16> wbf;
  // Activations of the last forward pass (capacity cap rows).
  int64_t cap = 0;
  std::vector<DBuf<float>> zbuf;   // pre-activations per hidden layer [M][width]
  std::vector<DBuf<__bf16>> hbuf;  // activations per hidden layer
  DBuf<__bf16> dzb, dza;
  DBuf<float> y, dy, red;
  DBuf<double> partials, loss;

  void init(int nin, int width, int depth, int nout, uint64_t seed);
  void ensure(int64_t M);
  const float* bias(int l) const;
  void gemm(const __bf16* A, int64_t sam, int64_t sak, const __bf16* B, int64_t sbk, int64_t sbn,
            int64_t M, int64_t N, int64_t K, int epi, const float* bias, const float* zin,
            float* outf, __bf16* outb, int splits, hipStream_t s);
  // keep_pre: store the hidden pre-activations (needed by gradients()).
  void forward(const float* X, int64_t M, float* Y, hipStream_t s, bool keep_pre = true);
  // Inference only (no stored activations): the fused one-kernel network for
  // width 64, 128 or 256 and n_in <= 32, else forward() in row chunks.
  void infer(const float* X, int64_t M, float* Y, hipStream_t s);
  // Fills grad (same layout as the parameters) for the mean squared error
  // of the outputs against T; returns the loss (with read_loss false it
  // returns NaN without waiting for the stream, and last_loss reads it later).
  double gradients(const float* X, const float* T, int64_t M, hipStream_t s,
                   bool read_loss = true);
  double last_loss(hipStream_t s);
  void adam(float lr, hipStream_t s);
  void sync_bf16(hipStream_t s);
};

// Root mean square of n FP32 values on the device (deterministic: block
// partial sums in FP64 added in block order on the host).
double rms_f32(const float* x, int64_t n, hipStream_t s);
// dst[r][:] = scale * src[idx[r]][:] for r < B rows of width w (minibatch gather).
void launch_gather_rows(const float* src, const int64_t* idx, int64_t B, int w, float scale,
                        float* dst, hipStream_t s);

}  // namespace nosh
