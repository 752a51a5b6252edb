// mlp.hip -- the per-vertex MLP surrogate: hand-written BF16 MFMA kernels
// (gfx950 v_mfma_f32_32x32x16_bf16) for training and inference, FP32
// accumulation, FP32 master weights and Adam.
//
// Network: x (n_in) -> [Linear(width) + GELU] x depth -> Linear(n_out).
// Weights are stored "out x in" (row n holds the weights of output n) in
// BF16 for the forward/backward kernels and in FP32 for the optimizer.
// The hidden width x width layers run on MFMA through one generic kernel
//   C[m][n] = sum_k A(m,k) B(k,n),  A(m,k) = A[m*sam + k*sak],
//                                   B(k,n) = B[k*sbk + n*sbn],
// which covers the forward product h W^T, the backward dh = dz W and the
// weight gradient dW = dz^T h (split over K = batch rows, partials reduced
// in a fixed order). Tiles of A and B are staged through LDS with k
// contiguous, so one 16-byte LDS read feeds each MFMA operand.
// The n_in -> width and width -> n_out layers (K = 8 or N = 2) are FP32
// dot products on the vector units. GELU is the tanh form
//   gelu(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))),
// tanh from the exponential (tanh_fast).

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "kernels.hpp"
#include "mlp.hpp"

namespace nosh {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 128, kBK = 32, kPad = 8;
constexpr int kLds = kBK + kPad;  // bf16 per LDS row

// tanh(u) = 1 - 2 / (exp(2u) + 1) with the hardware exponential and
// reciprocal (relative error ~1e-6, far below the BF16 rounding the
// activations take next; exp(2u) = inf gives 1, 0 gives -1). The
// library tanhf made GELU 45 % of the fused inference time.
__device__ __forceinline__ float tanh_fast(float u) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * u) + 1.0f);
}

__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  const float u = c * (z + 0.044715f * z * z * z);
  return 0.5f * z * (1.0f + tanh_fast(u));
}

__device__ __forceinline__ float dgelu(float z) {
  const float c = 0.7978845608028654f;
  const float u = c * (z + 0.044715f * z * z * z);
  const float t = tanh_fast(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

struct GemmArgs {
  const __bf16* A;
  int64_t sam, sak;
  const __bf16* B;
  int64_t sbk, sbn;
  int64_t M, N, K;     // K: per split
  int64_t kchunk;      // K range per split (grid.z)
  int epi;             // 0: z = acc + bias, h = gelu(z); 1: dz = acc * gelu'(zin); 2: store acc
  const float* bias;   // [N]
  const float* zin;    // [M][N] pre-activations (epi 1)
  float* outf;         // [M][N] fp32 (z for epi 0, partials for epi 2 at + z*M*N)
  __bf16* outb;        // [M][N] bf16 (h for epi 0, dz for epi 1)
};

// Loads a kBM x kBK tile of A (rows m0.., k0..) into As[m][k].
__device__ __forceinline__ void load_a(const GemmArgs& g, int64_t m0, int64_t k0, int64_t kend,
                                       __bf16 (*As)[kLds]) {
  const int t = threadIdx.x;
  if (g.sak == 1) {
    const int row = t >> 2, kq = (t & 3) * 8;
    const int64_t m = m0 + row, k = k0 + kq;
    bf16x8 v;
    if (m < g.M && k + 8 <= kend) {
      v = *reinterpret_cast<const bf16x8*>(g.A + m * g.sam + k);
    } else {
      for (int q = 0; q < 8; ++q)
        v[q] = (m < g.M && k + q < kend) ? g.A[m * g.sam + k + q] : __bf16(0.0f);
    }
    *reinterpret_cast<bf16x8*>(&As[row][kq]) = v;
  } else {
    // m contiguous: 8 consecutive m for one k.
    const int kk = t >> 3, mq = (t & 7) * 8;
    const int64_t k = k0 + kk;
    for (int q = 0; q < 8; ++q) {
      const int64_t m = m0 + mq + q;
      As[mq + q][kk] = (m < g.M && k < kend) ? g.A[m * g.sam + k * g.sak] : __bf16(0.0f);
    }
  }
}

// Loads a kBK x kBN tile of B (k0.., n0..) into Bs[n][k].
__device__ __forceinline__ void load_b(const GemmArgs& g, int64_t k0, int64_t n0, int64_t kend,
                                       __bf16 (*Bs)[kLds]) {
  const int t = threadIdx.x;
  if (g.sbk == 1) {
    const int n = t >> 1, kq = (t & 1) * 16;
    const int64_t nn = n0 + n;
    for (int h = 0; h < 2; ++h) {
      const int64_t k = k0 + kq + 8 * h;
      bf16x8 v;
      if (nn < g.N && k + 8 <= kend) {
        v = *reinterpret_cast<const bf16x8*>(g.B + nn * g.sbn + k);
      } else {
        for (int q = 0; q < 8; ++q)
          v[q] = (nn < g.N && k + q < kend) ? g.B[nn * g.sbn + k + q] : __bf16(0.0f);
      }
      *reinterpret_cast<bf16x8*>(&Bs[n][kq + 8 * h]) = v;
    }
  } else {
    const int kk = t >> 3, nq = (t & 7) * 16;
    const int64_t k = k0 + kk;
    for (int q = 0; q < 16; ++q) {
      const int64_t nn = n0 + nq + q;
      Bs[nq + q][kk] = (nn < g.N && k < kend) ? g.B[k * g.sbk + nn * g.sbn] : __bf16(0.0f);
    }
  }
}

// One block (4 waves, 2 x 2) computes a kBM x kBN tile; each wave 32 x 64 as
// two 32x32 MFMA accumulators. grid: (ceil(M/kBM), ceil(N/kBN), splits).
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ __bf16 As[kBM][kLds];
  __shared__ __bf16 Bs[kBN][kLds];
  const int64_t m0 = int64_t(blockIdx.x) * kBM, n0 = int64_t(blockIdx.y) * kBN;
  const int64_t kbeg = int64_t(blockIdx.z) * g.kchunk;
  const int64_t kend = min(g.K, kbeg + g.kchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f32x16 acc[2];
  for (int t = 0; t < 2; ++t)
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
  for (int64_t k0 = kbeg; k0 < kend; k0 += kBK) {
    load_a(g, m0, k0, kend, As);
    load_b(g, k0, n0, kend, Bs);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kBK / 16; ++ks) {
      const int kk = ks * 16 + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + (lane & 31)][kk]);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const bf16x8 b =
            *reinterpret_cast<const bf16x8*>(&Bs[wn * 64 + t * 32 + (lane & 31)][kk]);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D[i][j], j = lane % 32, i = 8 (r/4) + 4 (lane/32) + r%4.
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int64_t n = n0 + wn * 64 + t * 32 + (lane & 31);
    if (n >= g.N) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (m >= g.M) continue;
      const float v = acc[t][r];
      const int64_t o = m * g.N + n;
      if (g.epi == 0) {
        const float z = v + g.bias[n];
        if (g.outf) g.outf[o] = z;
        g.outb[o] = __bf16(gelu(z));
      } else if (g.epi == 1) {
        const float dz = v * dgelu(g.zin[o]);
        g.outb[o] = __bf16(dz);
        if (g.outf) g.outf[o] = dz;
      } else {
        g.outf[int64_t(blockIdx.z) * g.M * g.N + o] = v;
      }
    }
  }
}

// out[i] = sum_s part[s * len + i] (fixed order).
__global__ void k_sum_splits(const float* __restrict__ part, int splits, int64_t len,
                             float* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += part[int64_t(z) * len + i];
  out[i] = s;
}

// First layer: z[m][n] = sum_c x[m][c] w[n][c] + b[n]; h = gelu(z).
__global__ void k_layer_in(const float* __restrict__ X, int64_t M, int nin,
                           const __bf16* __restrict__ W, const float* __restrict__ b, int N,
                           float* __restrict__ z, __bf16* __restrict__ h) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const int64_t m = i / N;
  const int n = int(i - m * N);
  float s = b[n];
  for (int c = 0; c < nin; ++c) s += X[m * nin + c] * float(W[n * nin + c]);
  if (z) z[i] = s;
  h[i] = __bf16(gelu(s));
}

// Last layer: y[m][o] = sum_k h[m][k] w[o][k] + b[o].
__global__ void k_layer_out(const __bf16* __restrict__ H, int64_t M, int K,
                            const __bf16* __restrict__ W, const float* __restrict__ b, int nout,
                            float* __restrict__ Y) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= M) return;
  for (int o = 0; o < nout; ++o) {
    float s = b[o];
    for (int k = 0; k < K; ++k) s += float(H[m * K + k]) * float(W[o * K + k]);
    Y[m * nout + o] = s;
  }
}

// Loss (mean squared error over M x nout) partials and dy = 2 (y - t) / (M nout).
__global__ void k_loss(const float* __restrict__ Y, const float* __restrict__ T, int64_t len,
                       float scale, float* __restrict__ dY, double* __restrict__ partials) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < len) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    v[0] = double(d) * double(d);
  }
  block_partials<1>(v, partials);
}

__global__ void k_finalize1(const double* __restrict__ partials, int nb, double scale,
                            double* __restrict__ out) {
  double v[1];
  finalize_sum<1>(v, partials, nb);
  if (threadIdx.x == 0) *out = v[0] * scale;
}

// dz4[m][k] = (sum_o dy[m][o] w[o][k]) gelu'(z4[m][k]).
__global__ void k_back_out(const float* __restrict__ dY, int64_t M, int nout,
                           const __bf16* __restrict__ W, int K, const float* __restrict__ z,
                           __bf16* __restrict__ dz) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= M * K) return;
  const int64_t m = i / K;
  const int k = int(i - m * K);
  float s = 0.0f;
  for (int o = 0; o < nout; ++o) s += dY[m * nout + o] * float(W[o * K + k]);
  dz[i] = __bf16(s * dgelu(z[i]));
}

// Column sums over rows of a [M][N] matrix in row chunks: part[c][n] = sum over chunk c.
template <class T>
__global__ void k_colsum_chunks(const T* __restrict__ A, int64_t M, int N, int64_t chunk,
                                float* __restrict__ part) {
  const int n = blockIdx.y * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t beg = int64_t(blockIdx.x) * chunk, end = min(M, beg + chunk);
  float s = 0.0f;
  for (int64_t m = beg; m < end; ++m) s += float(A[m * N + n]);
  part[int64_t(blockIdx.x) * N + n] = s;
}

// Small weight gradients: part[c][o][k] = sum over chunk c of dy[m][o] x[m][k] with
// x = FP32 (first layer inputs) or BF16 (last hidden layer).
template <class T>
__global__ void k_small_wgrad(const float* __restrict__ D, int no, const T* __restrict__ X,
                              int nk, int64_t M, int64_t chunk, float* __restrict__ part) {
  const int idx = blockIdx.y * blockDim.x + threadIdx.x;
  if (idx >= no * nk) return;
  const int o = idx / nk, k = idx - o * nk;
  const int64_t beg = int64_t(blockIdx.x) * chunk, end = min(M, beg + chunk);
  float s = 0.0f;
  for (int64_t m = beg; m < end; ++m) s += D[m * no + o] * float(X[m * nk + k]);
  part[int64_t(blockIdx.x) * no * nk + idx] = s;
}

__global__ void k_bf16_to_f32(const __bf16* __restrict__ a, int64_t n, float* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = float(a[i]);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = __bf16(wi);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, int64_t n, __bf16* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = __bf16(a[i]);
}

// Fused inference: the whole network for 128 rows per block (4 waves x 32
// rows), activations kept in LDS between layers, hidden width x width layers
// on MFMA with the weights read from L2. Same arithmetic and summation order
// as forward(): first layer FP32 over the inputs, hidden layers in 16-wide k
// chunks in increasing order, last layer FP32 over k in order; BF16
// activations and weights, FP32 accumulation.
constexpr int kInferRows = 128;
constexpr int kInferMaxIn = 32;
constexpr int kInferMaxLayers = 18;

struct InferArgs {
  const float* X;
  int64_t M;
  int nin, width, depth, nout;
  const __bf16* w;     // BF16 weights, layer l at woff[l], [out][in]
  const float* bias;   // FP32 biases, layer l at boff[l]
  int64_t woff[kInferMaxLayers];
  int64_t boff[kInferMaxLayers];
  float* Y;
};

// The weights of a hidden layer are staged through LDS in chunks of TC
// output tiles (32 TC rows of W) shared by the block's 4 waves, instead of
// every wave reading the whole layer from L2, double-buffered: each chunk's
// global loads are issued into registers before the MFMA loop over the
// previous chunk and stored to the other LDS buffer after it, so the L2
// latency of the staging hides behind the MFMAs (the next layer's first
// chunk is fetched during the current layer's last). The inputs of layer 0
// use the same LDS region before the first chunk.
template <int NT>
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs g) {
  constexpr int W = NT * 32;
  constexpr int LD = W + 8;            // padded row (bf16) against LDS bank conflicts
  constexpr int TC = NT < 2 ? NT : 2;  // output tiles per staged weight chunk
  constexpr int NC = NT / TC;          // chunks per layer
  constexpr int WROWS = 32 * TC;
  constexpr int BUFB = WROWS * LD * 2;  // bytes of one chunk buffer
  constexpr int VPT = WROWS * (W / 8) / 256;  // 16-byte vectors per thread and chunk
  static_assert(VPT >= 1 && WROWS * (W / 8) % 256 == 0, "chunk does not split over 256 threads");
  constexpr int WBYTES = 2 * BUFB > kInferRows * kInferMaxIn * 4 ? 2 * BUFB
                                                                 : kInferRows * kInferMaxIn * 4;
  __shared__ __bf16 act[kInferRows][LD];
  __shared__ __attribute__((aligned(16))) unsigned char wbuf[WBYTES];
  float (*xs)[kInferMaxIn] = reinterpret_cast<float (*)[kInferMaxIn]>(wbuf);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = int64_t(blockIdx.x) * kInferRows;
  const int rows = int(min<int64_t>(kInferRows, g.M - r0));
  for (int i = tid; i < kInferRows * g.nin; i += 256) {
    const int r = i / g.nin, c = i - r * g.nin;
    xs[r][c] = r < rows ? g.X[(r0 + r) * g.nin + c] : 0.0f;
  }
  __syncthreads();
  // Chunk staging: registers <- global, then LDS buffer <- registers.
  bf16x8 stage[VPT];
  auto fetch = [&](int l, int ch) {
    const __bf16* Wl = g.w + g.woff[l];
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      const int v = tid + q * 256;
      const int rr = v / (W / 8), cc = v - rr * (W / 8);
      stage[q] = *reinterpret_cast<const bf16x8*>(Wl + int64_t(ch * WROWS + rr) * W + cc * 8);
    }
  };
  auto commit = [&](int b) {
    __bf16 (*wl)[LD] = reinterpret_cast<__bf16 (*)[LD]>(wbuf + b * BUFB);
#pragma unroll
    for (int q = 0; q < VPT; ++q) {
      const int v = tid + q * 256;
      const int rr = v / (W / 8), cc = v - rr * (W / 8);
      *reinterpret_cast<bf16x8*>(&wl[rr][cc * 8]) = stage[q];
    }
  };
  if (g.depth > 1) fetch(1, 0);
  // Layer 0: n_in -> W on the vector units, one output column per thread.
  for (int n = tid; n < W; n += 256) {
    float wr[kInferMaxIn];
    for (int c = 0; c < g.nin; ++c) wr[c] = float(g.w[g.woff[0] + int64_t(n) * g.nin + c]);
    const float bn = g.bias[g.boff[0] + n];
    for (int r = 0; r < kInferRows; ++r) {
      float z = bn;
      for (int c = 0; c < g.nin; ++c) z += xs[r][c] * wr[c];
      act[r][n] = __bf16(gelu(z));
    }
  }
  __syncthreads();  // xs (the first buffers' region) is no longer read
  if (g.depth > 1) commit(0);
  __syncthreads();
  // Hidden layers: each wave owns rows 32 wave .. 32 wave + 31 (it reads and
  // writes only those rows of act between the block barriers).
  const int row = wave * 32 + (lane & 31);
  const int kh = 8 * (lane >> 5);
  int buf = 0;
  for (int l = 1; l < g.depth; ++l) {
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
      for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
      const bool more = ch + 1 < NC || l + 1 < g.depth;
      if (more) fetch(ch + 1 < NC ? l : l + 1, ch + 1 < NC ? ch + 1 : 0);
      __bf16 (*wl)[LD] = reinterpret_cast<__bf16 (*)[LD]>(wbuf + buf * BUFB);
#pragma unroll 2
      for (int kc = 0; kc < W / 16; ++kc) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&act[row][kc * 16 + kh]);
#pragma unroll
        for (int t = 0; t < TC; ++t) {
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(&wl[t * 32 + (lane & 31)][kc * 16 + kh]);
          acc[ch * TC + t] =
              __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[ch * TC + t], 0, 0, 0);
        }
      }
      if (more) commit(buf ^ 1);
      __syncthreads();  // the other buffer is complete; this one is free again
      buf ^= 1;
    }
    const float* bl = g.bias + g.boff[l];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = t * 32 + (lane & 31);
      const float bn = bl[n];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        act[wave * 32 + i][n] = __bf16(gelu(acc[t][r] + bn));
      }
    }
  }
  __syncthreads();
  // Last layer: W -> n_out, FP32 dot products over k in order (16-byte loads
  // of activations and weights).
  const int L = g.depth;
  for (int i = tid; i < kInferRows * g.nout; i += 256) {
    const int r = i / g.nout, o = i - r * g.nout;
    if (r >= rows) continue;
    const __bf16* wo = g.w + g.woff[L] + int64_t(o) * W;
    float s = g.bias[g.boff[L] + o];
    for (int k8 = 0; k8 < W / 8; ++k8) {
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(&act[r][k8 * 8]);
      const bf16x8 wv = *reinterpret_cast<const bf16x8*>(wo + k8 * 8);
#pragma unroll
      for (int e = 0; e < 8; ++e) s += float(av[e]) * float(wv[e]);
    }
    g.Y[(r0 + r) * g.nout + o] = s;
  }
}

unsigned g1(int64_t n, int b = 256) { return grid_for(n, b); }

}  // namespace

// ----------------------------------------------------------------- host
void Mlp::init(int nin_, int width_, int depth_, int nout_, uint64_t seed) {
  nin = nin_;
  width = width_;
  depth = depth_;
  nout = nout_;
  NOSH_CHECK(nin >= 1 && nin <= 64 && nout >= 1 && nout <= 16 && depth >= 1 && depth <= 16,
             kInvalid, "mlp: unsupported shape");
  NOSH_CHECK(width % 64 == 0 && width <= 2048, kInvalid, "mlp: width must be a multiple of 64");
  // Layer l maps dim[l] -> dim[l+1].
  dims.clear();
  dims.push_back(nin);
  for (int l = 0; l < depth; ++l) dims.push_back(width);
  dims.push_back(nout);
  woff.clear();
  boff.clear();
  int64_t nw = 0, nb = 0;
  for (size_t l = 0; l + 1 < dims.size(); ++l) {
    woff.push_back(nw);
    boff.push_back(nb);
    nw += int64_t(dims[l]) * dims[l + 1];
    nb += dims[l + 1];
  }
  nweights = nw;
  nparams = nw + nb;
  // He-style uniform init from a SplitMix64 stream (host), biases zero.
  std::vector<float> p(nparams, 0.0f);
  uint64_t st = seed;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (size_t l = 0; l + 1 < dims.size(); ++l) {
    const double a = std::sqrt(6.0 / dims[l]) * (l + 2 == dims.size() ? 0.1 : 1.0);
    for (int64_t i = 0; i < int64_t(dims[l]) * dims[l + 1]; ++i)
      p[woff[l] + i] = float((2.0 * uni() - 1.0) * a);
  }
  hipStream_t s = default_stream();
  master.alloc(nparams);
  master.upload(p.data(), nparams, s);
  grad.alloc(nparams);
  grad.zero(s);
  m1.alloc(nparams);
  m1.zero(s);
  m2.alloc(nparams);
  m2.zero(s);
  wbf.alloc(nweights);
  k_f32_to_bf16<<<g1(nweights), 256, 0, s>>>(master.p, nweights, wbf.p);
  steps = 0;
  NOSH_HIP(hipStreamSynchronize(s));
}

void Mlp::ensure(int64_t M) {
  if (M <= cap) return;
  cap = M;
  const int L = int(dims.size()) - 1;
  zbuf.resize(L - 1);
  hbuf.resize(L - 1);
  for (int l = 0; l < L - 1; ++l) {
    zbuf[l].alloc(M * dims[l + 1]);
    hbuf[l].alloc(M * dims[l + 1]);
  }
  dzb.alloc(M * width);
  dza.alloc(M * width);
  y.alloc(M * nout);
  dy.alloc(M * nout);
  partials.alloc(g1(M * nout) + 64);
  loss.alloc(1);
}

const float* Mlp::bias(int l) const { return master.p + nweights + boff[l]; }

void Mlp::gemm(const __bf16* A, int64_t sam, int64_t sak, const __bf16* B, int64_t sbk,
               int64_t sbn, int64_t M, int64_t N, int64_t K, int epi, const float* bias_,
               const float* zin, float* outf, __bf16* outb, int splits, hipStream_t s) {
  GemmArgs g;
  g.A = A;
  g.sam = sam;
  g.sak = sak;
  g.B = B;
  g.sbk = sbk;
  g.sbn = sbn;
  g.M = M;
  g.N = N;
  g.K = K;
  const int64_t kc = ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;
  g.kchunk = kc;
  g.epi = epi;
  g.bias = bias_;
  g.zin = zin;
  g.outf = outf;
  g.outb = outb;
  dim3 grid(unsigned((M + kBM - 1) / kBM), unsigned((N + kBN - 1) / kBN),
            unsigned((K + kc - 1) / kc));
  k_gemm_bf16<<<grid, 256, 0, s>>>(g);
}

void Mlp::forward(const float* X, int64_t M, float* Y, hipStream_t s, bool keep_pre) {
  ensure(M);
  const int L = int(dims.size()) - 1;
  k_layer_in<<<g1(M * dims[1]), 256, 0, s>>>(X, M, nin, wbf.p + woff[0], bias(0), dims[1],
                                             keep_pre ? zbuf[0].p : nullptr, hbuf[0].p);
  for (int l = 1; l < L - 1; ++l) {
    // z = h W^T: A = h [M][K], B(k,n) = W[n][k].
    gemm(hbuf[l - 1].p, dims[l], 1, wbf.p + woff[l], 1, dims[l], M, dims[l + 1], dims[l], 0,
         bias(l), nullptr, keep_pre ? zbuf[l].p : nullptr, hbuf[l].p, 1, s);
  }
  k_layer_out<<<g1(M), 256, 0, s>>>(hbuf[L - 2].p, M, dims[L - 1], wbf.p + woff[L - 1],
                                    bias(L - 1), nout, Y ? Y : y.p);
}

void Mlp::infer(const float* X, int64_t M, float* Y, hipStream_t s) {
  const int nt = width / 32;
  if (width % 32 != 0 || (nt != 2 && nt != 4 && nt != 8) || nin > kInferMaxIn ||
      depth + 1 > kInferMaxLayers) {
    constexpr int64_t kChunk = 1 << 19;
    for (int64_t r0 = 0; r0 < M; r0 += kChunk)
      forward(X + r0 * nin, std::min(kChunk, M - r0), Y + r0 * nout, s, false);
    return;
  }
  InferArgs g{};
  g.X = X;
  g.M = M;
  g.nin = nin;
  g.width = width;
  g.depth = depth;
  g.nout = nout;
  g.w = wbf.p;
  g.bias = master.p + nweights;
  for (int l = 0; l <= depth; ++l) {
    g.woff[l] = woff[l];
    g.boff[l] = boff[l];
  }
  g.Y = Y;
  const unsigned nb = unsigned((M + kInferRows - 1) / kInferRows);
  if (nt == 8)
    k_mlp_infer<8><<<nb, 256, 0, s>>>(g);
  else if (nt == 4)
    k_mlp_infer<4><<<nb, 256, 0, s>>>(g);
  else
    k_mlp_infer<2><<<nb, 256, 0, s>>>(g);
}

double Mlp::last_loss(hipStream_t s) {
  double lh = 0.0;
  NOSH_HIP(hipMemcpyAsync(&lh, loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  return lh;
}

double Mlp::gradients(const float* X, const float* T, int64_t M, hipStream_t s, bool read_loss) {
  ensure(M);
  forward(X, M, y.p, s);
  const int L = int(dims.size()) - 1;
  const int64_t len = M * nout;
  const unsigned nbl = g1(len);
  k_loss<<<nbl, 256, 0, s>>>(y.p, T, len, 1.0f / float(len), dy.p, partials.p);
  k_finalize1<<<1, 256, 0, s>>>(partials.p, int(nbl), 1.0 / double(len), loss.p);
  // Row chunks for the deterministic reductions over the batch (column sums
  // and the small first/last-layer weight gradients): 32 rows (at most 4096
  // chunks), so that a batch of 8 k rows gives 256 chunks of parallel work
  // (1024-row chunks took 240 us per reduction at 8 k rows, 1.6 of the 2 ms
  // of an Adam step).
  const int64_t chunk = std::max<int64_t>(32, (M + 4095) / 4096);
  const int64_t nchunk = (M + chunk - 1) / chunk;
  // Split-K of the hidden layers' weight-gradient GEMMs: 1024 rows a split.
  const int64_t nsplit = (M + 1023) / 1024;
  // Scratch for the partial sums: split-K GEMM partials (<= 64 x width^2),
  // per-chunk column sums (nchunk x width) and small weight gradients
  // (nchunk x width x max(n_in, n_out)).
  const int64_t need = std::max<int64_t>(64 * int64_t(width) * width,
                                         nchunk * int64_t(width) * std::max(nin, nout));
  if (red.n < size_t(need)) red.alloc(need);
  auto wg = [&](int l) { return grad.p + woff[l]; };
  auto bg = [&](int l) { return grad.p + nweights + boff[l]; };
  // Last layer: dW = dy^T h, db = sum dy.
  {
    const int K = dims[L - 1];
    dim3 grid(unsigned(nchunk), unsigned(g1(int64_t(nout) * K)));
    k_small_wgrad<__bf16><<<grid, 256, 0, s>>>(dy.p, nout, hbuf[L - 2].p, K, M, chunk, red.p);
    k_sum_splits<<<g1(int64_t(nout) * K), 256, 0, s>>>(red.p, int(nchunk), int64_t(nout) * K,
                                                        wg(L - 1));
    dim3 gb(unsigned(nchunk), 1);
    k_colsum_chunks<float><<<gb, 256, 0, s>>>(dy.p, M, nout, chunk, red.p);
    k_sum_splits<<<1, 256, 0, s>>>(red.p, int(nchunk), nout, bg(L - 1));
    k_back_out<<<g1(M * K), 256, 0, s>>>(dy.p, M, nout, wbf.p + woff[L - 1], K, zbuf[L - 2].p,
                                        dzb.p);
  }
  // Hidden layers l = L-2 .. 1 (MFMA): dW_l = dz^T h_{l-1}, dh_{l-1} = dz W_l.
  __bf16* dz = dzb.p;
  __bf16* dzn = dza.p;
  for (int l = L - 2; l >= 1; --l) {
    const int64_t N = dims[l + 1], K = dims[l];
    const int splits = int(std::min<int64_t>(nsplit, 64));
    // C[n][k] = sum_m dz[m][n] h[m][k]: A(n,m) = dz[m*N + n], B(m,k) = h[m*K + k].
    gemm(dz, 1, N, hbuf[l - 1].p, K, 1, N, K, M, 2, nullptr, nullptr, red.p, nullptr, splits,
         s);
    const int64_t kc = ((M + splits - 1) / splits + kBK - 1) / kBK * kBK;
    const int used = int((M + kc - 1) / kc);
    k_sum_splits<<<g1(N * K), 256, 0, s>>>(red.p, used, N * K, wg(l));
    dim3 gb(unsigned(nchunk), unsigned(g1(N)));
    k_colsum_chunks<__bf16><<<gb, 256, 0, s>>>(dz, M, int(N), chunk, red.p);
    k_sum_splits<<<g1(N), 256, 0, s>>>(red.p, int(nchunk), N, bg(l));
    // dz_{l-1} = (dz W_l) * gelu'(z_{l-1}): A = dz [M][N], B(n,k) = W[n][k].
    gemm(dz, N, 1, wbf.p + woff[l], K, 1, M, K, N, 1, nullptr, zbuf[l - 1].p, nullptr, dzn, 1,
         s);
    std::swap(dz, dzn);
  }
  // First layer: dW = dz^T x, db = sum dz.
  {
    const int K = nin, N = dims[1];
    dim3 grid(unsigned(nchunk), unsigned(g1(int64_t(N) * K)));
    // part[c][n][k] = sum dz[m][n] x[m][k] with dz in BF16: convert to FP32 rows first.
    k_bf16_to_f32<<<g1(M * N), 256, 0, s>>>(dz, M * N, zbuf[0].p);
    k_small_wgrad<float><<<grid, 256, 0, s>>>(zbuf[0].p, N, X, K, M, chunk, red.p);
    k_sum_splits<<<g1(int64_t(N) * K), 256, 0, s>>>(red.p, int(nchunk), int64_t(N) * K, wg(0));
    dim3 gb(unsigned(nchunk), unsigned(g1(N)));
    k_colsum_chunks<float><<<gb, 256, 0, s>>>(zbuf[0].p, M, N, chunk, red.p);
    k_sum_splits<<<g1(N), 256, 0, s>>>(red.p, int(nchunk), N, bg(0));
  }
  return read_loss ? last_loss(s) : std::numeric_limits<double>::quiet_NaN();
}

void Mlp::adam(float lr, hipStream_t s) {
  ++steps;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, float(steps)), c2 = 1.0f - std::pow(b2, float(steps));
  k_adam<<<g1(nweights), 256, 0, s>>>(master.p, grad.p, m1.p, m2.p, wbf.p, nweights, lr, b1, b2,
                                      eps, c1, c2);
  k_adam<<<g1(nparams - nweights), 256, 0, s>>>(master.p + nweights, grad.p + nweights,
                                                m1.p + nweights, m2.p + nweights, nullptr,
                                                nparams - nweights, lr, b1, b2, eps, c1, c2);
}

void Mlp::sync_bf16(hipStream_t s) {
  k_f32_to_bf16<<<g1(nweights), 256, 0, s>>>(master.p, nweights, wbf.p);
}

namespace {
__global__ void k_gather_rows(const float* __restrict__ src, const int64_t* __restrict__ idx,
                              int64_t B, int w, float scale, float* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * w) return;
  const int64_t r = i / w;
  dst[i] = src[idx[r] * w + (i - r * w)] * scale;
}

__global__ void k_sumsq_f32(const float* __restrict__ x, int64_t n, double* __restrict__ part) {
  double v[1] = {0.0};
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double t = x[i];
    v[0] += t * t;
  }
  block_partials<1>(v, part);
}
}  // namespace

double rms_f32(const float* x, int64_t n, hipStream_t s) {
  if (n <= 0) return 0.0;
  const int nb = int(std::min<int64_t>(1024, grid_for(n, 256)));
  DBuf<double> part{size_t(nb)};
  k_sumsq_f32<<<nb, 256, 0, s>>>(x, n, part.p);
  std::vector<double> h(static_cast<size_t>(nb));
  part.download(h.data(), h.size(), s);
  NOSH_HIP(hipStreamSynchronize(s));
  double ss = 0.0;
  for (double v : h) ss += v;
  return std::sqrt(ss / double(n));
}

void launch_gather_rows(const float* src, const int64_t* idx, int64_t B, int w, float scale,
                        float* dst, hipStream_t s) {
  k_gather_rows<<<grid_for(B * w, 256), 256, 0, s>>>(src, idx, B, w, scale, dst);
}

}  // namespace nosh
