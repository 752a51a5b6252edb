// batch_continuation.cpp -- natural-parameter continuation in mu of a batch
// of branches, solved together.
//
// Every instance of a batch is a branch (mesh, g, mu_0). One step moves all
// branches by the same increment h in mu and solves them with the lockstep
// batched Newton (FP32 inner MINRES, FP64 outer). Predictors: the tangent
//   psi_k + h dpsi/dmu,   J(psi_k) dpsi/dmu = -dF/dmu,   one lockstep solve;
// the secant through the last two accepted solutions,
//   psi_sec = psi_k + (h / h_prev) (psi_k - psi_{k-1})   (psi_k on the first step);
// or the tangent on the first step and after a step that a branch did not
// pass from its first guess, the secant otherwise. The chosen predictor's
// guess is written psi_pred below.
// With a surrogate, after `train_after` steps a per-vertex MLP (BF16 MFMA,
// trained on the run's own solutions) seeds Newton with
//   psi_guess = psi_pred + h^2 sigma Y(x, psi_k, h s, mu, g),   h s = psi_pred - psi_k,
// sigma the RMS of the training targets (psi_{k+1} - psi_pred) / h^2, then
// smoothed by smooth_sweeps damped Jacobi sweeps on its residual; a branch
// starts from the surrogate guess when its ||F|| is at most accept_ratio
// times psi_pred's (inexact Newton takes as many steps, and often more
// MINRES iterations, from a guess that is only slightly better), else from
// psi_pred. A branch whose Newton from the surrogate guess does not converge
// within the budget, or whose Gibbs energy leaves its linear trend, is
// solved again from psi_pred; at a branch transition, where Newton from
// psi_pred may wander, the branches a short try leaves are retried in
// lockstep by pseudo-transient continuation (shifted Newton steps,
// J + sigma c t), from psi_pred and then from psi_k; a branch none of these
// brings to convergence is handed, alone, to the single-instance FP64
// solver, and last to batched sub-steps.
//
// This is the caller the reference delegates to LOCA (natural stepping,
// executables/nosh-cont/nosh-cont.cpp:226-290,344 with examples/conf.xml:35-190);
// the Gibbs energy per branch is the observer's
// (src/observer.cpp:134-159): -sum cv |psi|^4 / sum cv.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "batch_handle.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

using namespace nosh;

namespace {

template <class F>
int bcguard(F&& f) {
  try {
    f();
    set_last_error("");
    return NOSH_OK;
  } catch (const nosh::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NOSH_ERR_INVALID;
  }
}

// Newton steps of the first try from each guess (12 measured slower at 96^3:
// more branches reach the full-length tries).
constexpr int kShortNewton = 20;
// A short try stops a branch whose ||F|| has not fallen below half its best
// value for this many steps and is not halving step after step: at the 96^3
// transitions Newton from a wrong guess oscillates between 5e-3 and 3e-2 for
// all 20 steps, while the tries that converge descend steadily, also after a
// jump up (96^3, 20 steps: mu 0.30 -> 0.35 from the tangent guess in 14
// steps via ||F|| = 3.7e3, falling 3.4x a step; mu 0.95 -> 1.0 from psi_k in
// 16 steps via ||F|| = 4).
constexpr int kStallSteps = 8;
// The pseudo-transient retries run without that stop: a 12-step window
// saved 2 s in two of thirteen 96^3 continuations (a retry stagnating near
// 1e-5 for 40 steps at mu = 0.93) but cost 98 s at 160^3, and an 8-step
// window 19-21 s in two 96^3 runs, where it cut off retries that would
// have converged and sent the branches to the serial fallback.
// Chebyshev degree of the single-instance FP64 fallback solves (the plain
// solver's best).
constexpr int kFallbackCheb = 6;
constexpr int kMaxHalvings = 3;   // step splits (2, 4, 8 sub-steps) for failed branches
// Pseudo-transient retries: initial shift sigma_0 (in units of the control
// volume times thickness, the mass matrix of F, so independent of the mesh
// size), forcing cap and Chebyshev degree (96^3 transition mu 0.70 -> 0.75:
// sigma_0 = 0.1 in 10 steps, 1 in 22, 10 did not converge in 40). Forcing
// cap over five continuations (96^3, 8 branches; 20 steps with each
// predictor, 50 steps auto and secant), total seconds: 1e-4 28.6, 1e-3
// 28.0, 1e-2 27.7 (the fastest in each of the five), 1e-1 28.0 (one
// transition in 80 instead of 68 Newton steps); over eight more (16-40
// steps, three predictors) 1e-2 43.6 s and 1e-4 165.1 s (two 30-step runs
// took 894 and 1346 lockstep Newton steps). With the cap at 1e-2:
// degree 5 27.8; sigma_0 = 0.2 27.5, 0.05 24.3, 0.03 31.0, but at 160^3
// (50 steps, auto and secant) sigma_0 = 0.1 took 122.7 s, 0.05 151.3 s and
// 0.03 193.9 s (the transition at mu = 0.94 in 36, 75 and 121 steps).
constexpr double kPtcShift0 = 0.1;
constexpr double kPtcForcing = 1e-2;
constexpr int kPtcCheb = 6;
// Relative MINRES tolerance of the tangent solve: the guess then errs by
// about kTangentRtol h |dpsi/dmu| from the solve against ~h^2 |psi''| / 2 of
// the tangent itself (96^3, 8 branches, mu 0 -> 1: 20 steps 4.73 / 4.68 /
// 7.18 s and 50 steps 11.6 / 10.4 / 9.9 s at 1e-3 / 1e-2 / 1e-1).
constexpr double kTangentRtol = 1e-2;
// Damping of the Jacobi sweeps that smooth the surrogate's guess.
constexpr double kSmoothOmega = 0.6;
constexpr int kRed = 256;       // threads of the per-lane reduction blocks
constexpr int kRedBlocks = 512; // blocks of the per-lane reduction (grid-stride)

// out = cur + r (cur - prev) over a group's interleaved [n][bp] state.
__global__ void k_bc_secant(int64_t m, double r, const double2* __restrict__ cur,
                            const double2* __restrict__ prev, double2* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double2 c = cur[i], p = prev[i];
  out[i] = make_double2(c.x + r * (c.x - p.x), c.y + r * (c.y - p.y));
}

// Feature rows (row, b) -> X[(row * B + b) * nin ...]:
// coordinates * cs (dim), Re/Im psi_k, Re/Im (psi_pred - psi_k), mu_b, g_b.
__global__ void k_bc_features(int64_t n, int bp, int B, int dim, const int32_t* __restrict__ perm,
                              const double* __restrict__ coords, double cs,
                              const double2* __restrict__ cur, const double2* __restrict__ sec,
                              const double* __restrict__ mu, const double* __restrict__ g,
                              int inst0, float* __restrict__ X) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t row = i / B;
  const int b = int(i - row * B);
  const int nin = dim + 6;
  float* x = X + i * nin;
  const int64_t v = perm[row];
  for (int d = 0; d < dim; ++d) x[d] = float(coords[3 * v + d] * cs);
  const double2 c = cur[row * bp + b], s = sec[row * bp + b];
  x[dim] = float(c.x);
  x[dim + 1] = float(c.y);
  x[dim + 2] = float(s.x - c.x);
  x[dim + 3] = float(s.y - c.y);
  x[dim + 4] = float(mu[inst0 + b]);
  x[dim + 5] = float(g[inst0 + b]);
}

// Training target (psi_new - psi_pred) / h^2 per (row, b).
__global__ void k_bc_target(int64_t n, int bp, int B, const double2* __restrict__ pnew,
                            const double2* __restrict__ sec, double ih2, float* __restrict__ T) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t row = i / B;
  const int b = int(i - row * B);
  const double2 a = pnew[row * bp + b], s = sec[row * bp + b];
  T[2 * i] = float((a.x - s.x) * ih2);
  T[2 * i + 1] = float((a.y - s.y) * ih2);
}

// out = sec + scale Y for real lanes; padding lanes copy sec.
__global__ void k_bc_apply(int64_t n, int bp, int B, const double2* __restrict__ sec,
                           const float* __restrict__ Y, double scale, double2* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * bp) return;
  const int64_t row = i / bp;
  const int b = int(i - row * bp);
  double2 s = sec[i];
  if (b < B) {
    const int64_t k = row * B + b;
    s.x += scale * double(Y[2 * k]);
    s.y += scale * double(Y[2 * k + 1]);
  }
  out[i] = s;
}

// One damped Jacobi sweep on the residual, for the real lanes:
//   out = psi - omega F(psi) / d,   d = sum alpha + c t (V + 2 g |psi|^2)
// (the diagonal of J's 2 x 2 block, averaged), psi and F(psi) interleaved
// [n][bp]. It removes the rough part of a guess's error, which K amplifies
// in the residual, and leaves the smooth part alone.
__global__ void k_bc_smooth(int64_t n, int bp, int B, int inst0, const double* __restrict__ diag,
                            const double* __restrict__ ct, const double* __restrict__ V,
                            const double* __restrict__ g, double omega,
                            const double2* __restrict__ psi, const double2* __restrict__ f,
                            double2* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * bp) return;
  const int64_t row = i / bp;
  const int b = int(i - row * bp);
  const double2 p = psi[i];
  if (b >= B) {
    out[i] = p;
    return;
  }
  const double2 r = f[i];
  const double d = diag[row] + ct[row] * (V[row] + 2.0 * g[inst0 + b] * (p.x * p.x + p.y * p.y));
  const double w = omega / d;
  out[i] = make_double2(p.x - w * r.x, p.y - w * r.y);
}

// dst[row][b] = src[row][b] for the lanes with take[b] != 0.
__global__ void k_bc_take(int64_t n, int bp, const int* __restrict__ take,
                          const double2* __restrict__ src, double2* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * bp) return;
  if (take[i & (bp - 1)]) dst[i] = src[i];
}

// out[row][b] = a + r (a - c) for the lanes with take[b] != 0.
__global__ void k_bc_lin_take(int64_t n, int bp, const int* __restrict__ take, double r,
                              const double2* __restrict__ a, const double2* __restrict__ c,
                              double2* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * bp) return;
  if (!take[i & (bp - 1)]) return;
  const double2 x = a[i], y = c[i];
  out[i] = make_double2(x.x + r * (x.x - y.x), x.y + r * (x.y - y.y));
}

// dst[row] = src[row * bp + b] (one lane of an interleaved state), and back.
__global__ void k_bc_lane_get(int64_t n, int bp, int b, const double2* __restrict__ src,
                              double2* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i * bp + b];
}
__global__ void k_bc_lane_put(int64_t n, int bp, int b, const double2* __restrict__ src,
                              double2* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i * bp + b] = src[i];
}

// For the lanes with take[b] != 0: psi <- psi + (psi - before), before <- old psi.
__global__ void k_bc_advance_take(int64_t n, int bp, const int* __restrict__ take,
                                  double2* __restrict__ psi, double2* __restrict__ before) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * bp) return;
  if (!take[i & (bp - 1)]) return;
  const double2 x = psi[i], y = before[i];
  before[i] = x;
  psi[i] = make_double2(2.0 * x.x - y.x, 2.0 * x.y - y.y);
}

// Per-lane sums over rows, deterministic (fixed grid, fixed order):
// mode 0: cv |a|^4, mode 1: |a - b|^2, mode 2: cv. partials[lane][block].
__global__ __launch_bounds__(kRed) void k_bc_lane_partials(int64_t n, int bp, int mode,
                                                           const double* __restrict__ cv,
                                                           const double2* __restrict__ a,
                                                           const double2* __restrict__ b,
                                                           double* __restrict__ partials) {
  __shared__ double lds[kRed];
  const int lane = threadIdx.x & (bp - 1);
  const int rows_per_block = kRed / bp;
  double acc = 0.0;
  for (int64_t row = int64_t(blockIdx.x) * rows_per_block + threadIdx.x / bp; row < n;
       row += int64_t(gridDim.x) * rows_per_block) {
    const int64_t i = row * bp + lane;
    if (mode == 0) {
      const double2 x = a[i];
      const double r2 = x.x * x.x + x.y * x.y;
      acc += cv[row] * r2 * r2;
    } else if (mode == 1) {
      const double2 x = a[i], y = b[i];
      const double dx = x.x - y.x, dy = x.y - y.y;
      acc += dx * dx + dy * dy;
    } else {
      acc += cv[row];
    }
  }
  lds[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < bp) {
    double s = 0.0;
    for (int r = 0; r < rows_per_block; ++r) s += lds[r * bp + threadIdx.x];
    partials[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = s;
  }
}

// One block per lane: out[lane] = sum of its nb partials (fixed order).
__global__ void k_bc_lane_final(const double* __restrict__ partials, int nb,
                                double* __restrict__ out) {
  double v[1];
  finalize_sum<1>(v, partials + int64_t(blockIdx.x) * nb, nb);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

struct GroupCont {
  DBuf<double2> cur, prev, sec, guess;
  DBuf<int> take;
  std::vector<int> take_h;
  double cv_sum = 0.0;
  int64_t xoff = 0;  // first feature row of the group
};

}  // namespace

extern "C" {

int nosh_batch_continuation_run(nosh_batch_t bt, double* psi, double span, int steps,
                                int predictor, const nosh_newton_options* opts,
                                const nosh_surrogate_options* so,
                                nosh_batch_continuation_step* record, int* steps_done,
                                nosh_surrogate_stats* stats) {
  int done = 0;
  nosh_surrogate_stats sst{};
  int rc = bcguard([&] {
    NOSH_CHECK(bt && psi && record && steps >= 1, kInvalid, "invalid batch continuation call");
    NOSH_CHECK(predictor >= 1 && predictor <= 3, kInvalid,
               "batch continuation: predictor must be 1 (tangent), 2 (secant) or 3 (tangent "
               "where the secant has no smooth history, else secant)");
    NOSH_CHECK(!so || (so->train_after >= 1 && so->batch >= 64 && so->newton_budget >= 1 &&
                       so->retrain_steps >= 0 && so->smooth_sweeps >= 0 &&
                       so->accept_ratio > 0.0),
               kInvalid, "bad surrogate options");
    nosh_newton_options o;
    if (opts)
      o = *opts;
    else
      nosh_newton_default_options(&o);
    Batch& B = bt->b;
    const int ni = B.args.ninst;
    hipStream_t s = default_stream();
    const size_t ng = B.ghost.size();
    std::vector<GroupCont> gc(ng);
    DBuf<double> part(size_t(64) * kRedBlocks), lane_out(64);
    DBuf<double> fb;  // one branch's state (reference order) for the single-instance fallback
    std::vector<double> fb_h;
    std::vector<double> lane_h(64);
    int64_t xrows = 0;
    int dim = 2;
    for (size_t k = 0; k < ng; ++k) {
      GroupHost& G = B.ghost[k];
      const int64_t m = G.mesh->n * G.dev.bp;
      gc[k].cur.alloc(m);
      gc[k].prev.alloc(m);
      gc[k].sec.alloc(m);
      gc[k].guess.alloc(m);
      gc[k].take.alloc(G.dev.bp);
      gc[k].xoff = xrows;
      xrows += G.mesh->n * G.B;
      if (G.mesh->nv == 4) dim = 3;
    }
    // Per-lane reductions over a group: lane_h[b] for b < bp.
    auto lane_sums = [&](size_t k, int mode, const double2* a, const double2* b) {
      GroupHost& G = B.ghost[k];
      const int bp = G.dev.bp;
      k_bc_lane_partials<<<kRedBlocks, kRed, 0, s>>>(G.mesh->n, bp, mode, G.mesh->cv_int.p, a, b,
                                                     part.p);
      k_bc_lane_final<<<bp, kFinalizeThreads, 0, s>>>(part.p, kRedBlocks, lane_out.p);
      NOSH_HIP(hipMemcpyAsync(lane_h.data(), lane_out.p, sizeof(double) * bp,
                              hipMemcpyDeviceToHost, s));
      NOSH_HIP(hipStreamSynchronize(s));
    };
    for (size_t k = 0; k < ng; ++k) {
      lane_sums(k, 2, nullptr, nullptr);
      gc[k].cv_sum = lane_h[0];
    }
    // Gibbs energy of every instance at the current state.
    auto energies = [&](std::vector<double>& e) {
      e.assign(ni, 0.0);
      for (size_t k = 0; k < ng; ++k) {
        GroupHost& G = B.ghost[k];
        lane_sums(k, 0, G.psi.p, nullptr);
        for (int b = 0; b < G.B; ++b) e[G.dev.inst0 + b] = -lane_h[b] / gc[k].cv_sum;
      }
    };
    auto copy_state = [&](DBuf<double2> GroupCont::*dst, bool from_psi) {
      for (size_t k = 0; k < ng; ++k) {
        GroupHost& G = B.ghost[k];
        const size_t bytes = sizeof(double2) * G.mesh->n * G.dev.bp;
        if (from_psi)
          NOSH_HIP(hipMemcpyAsync((gc[k].*dst).p, G.psi.p, bytes, hipMemcpyDeviceToDevice, s));
        else
          NOSH_HIP(hipMemcpyAsync(G.psi.p, (gc[k].*dst).p, bytes, hipMemcpyDeviceToDevice, s));
      }
    };
    // Copies src into psi for the instances with take[i] != 0.
    auto take_into_psi = [&](DBuf<double2> GroupCont::*src, const std::vector<int>& take) {
      for (size_t k = 0; k < ng; ++k) {
        GroupHost& G = B.ghost[k];
        std::vector<int>& t = gc[k].take_h;
        t.assign(G.dev.bp, 0);
        for (int b = 0; b < G.B; ++b) t[b] = take[G.dev.inst0 + b];
        gc[k].take.upload(t.data(), t.size(), s);
        const int64_t m = G.mesh->n * G.dev.bp;
        k_bc_take<<<grid_for(m, 256), 256, 0, s>>>(G.mesh->n, G.dev.bp, gc[k].take.p,
                                                   (gc[k].*src).p, G.psi.p);
      }
    };
    std::vector<double> mu(bt->mu_h), e_now, e1(ni), e0(ni);
    auto set_mu = [&]() { B.set_mu(mu.data(), s); };

    // Surrogate state.
    const int nin = dim + 6;
    std::unique_ptr<Mlp> net;
    DBuf<float> X, Y, Xs, Ts, xb, tb;
    DBuf<int64_t> idx;
    std::vector<int64_t> idx_h;
    int P = 0, pairs = 0;
    bool trained = false;
    double sigma = 1.0, cs = 1.0;
    std::mt19937_64 rng(so ? so->seed : 0);
    if (so) {
      P = so->train_after;
      net = std::make_unique<Mlp>();
      net->init(nin, so->width, so->depth, 2, so->seed);
      X.alloc(xrows * nin);
      Y.alloc(xrows * 2);
      Xs.alloc(int64_t(P) * xrows * nin);
      Ts.alloc(int64_t(P) * xrows * 2);
      xb.alloc(int64_t(so->batch) * nin);
      tb.alloc(int64_t(so->batch) * 2);
      idx.alloc(so->batch);
      idx_h.resize(so->batch);
      double cmax = 1e-300;
      for (size_t k = 0; k < ng; ++k) {
        MeshCache& m = *B.ghost[k].mesh;
        std::vector<double> ch(3 * m.n);
        m.coords.download(ch.data(), ch.size(), s);
        NOSH_HIP(hipStreamSynchronize(s));
        for (double v : ch) cmax = std::max(cmax, std::fabs(v));
      }
      cs = 1.0 / cmax;
    }
    auto predict = [&](double scale) {
      for (size_t k = 0; k < ng; ++k) {
        GroupHost& G = B.ghost[k];
        k_bc_features<<<grid_for(G.mesh->n * G.B, 256), 256, 0, s>>>(
            G.mesh->n, G.dev.bp, G.B, dim, G.mesh->perm.p, G.mesh->coords.p, cs, gc[k].cur.p,
            gc[k].sec.p, B.mu.p, B.g.p, G.dev.inst0, X.p + gc[k].xoff * nin);
      }
      net->infer(X.p, xrows, Y.p, s);
      for (size_t k = 0; k < ng; ++k) {
        GroupHost& G = B.ghost[k];
        k_bc_apply<<<grid_for(G.mesh->n * G.dev.bp, 256), 256, 0, s>>>(
            G.mesh->n, G.dev.bp, G.B, gc[k].sec.p, Y.p + gc[k].xoff * 2, scale, gc[k].guess.p);
      }
      // Jacobi sweeps on the residual of the guess (each: F at the guess,
      // then the damped update).
      std::vector<double> nf_tmp;
      for (int sw = 0; sw < so->smooth_sweeps; ++sw) {
        copy_state(&GroupCont::guess, false);
        B.compute_norms(s, nf_tmp);
        for (size_t k = 0; k < ng; ++k) {
          GroupHost& G = B.ghost[k];
          k_bc_smooth<<<grid_for(G.mesh->n * G.dev.bp, 256), 256, 0, s>>>(
              G.mesh->n, G.dev.bp, G.B, G.dev.inst0, G.diag.p, G.ct.p, G.V.p, B.g.p,
              kSmoothOmega, G.psi.p, G.f.p, gc[k].guess.p);
        }
      }
    };
    // Adam steps on random vertex minibatches of the stored pairs; the first
    // call also fixes the output scale sigma (RMS of the targets).
    auto train = [&](int nsteps) {
      if (!trained) {
        sigma = rms_f32(Ts.p, int64_t(P) * xrows * 2, s);
        if (!(sigma > 0.0)) sigma = 1.0;
      }
      // Every step's minibatch drawn up front (the same draws, in the same
      // order, as one step at a time) and uploaded once, and the loss read
      // back only after the last step: the steps then queue on the stream
      // without a host round trip each.
      std::uniform_int_distribution<int64_t> pick(0, int64_t(P) * xrows - 1);
      const size_t nbat = size_t(so->batch);
      idx_h.resize(nbat * size_t(std::max(nsteps, 1)));
      for (auto& v : idx_h) v = pick(rng);
      if (idx.n < idx_h.size()) idx.alloc(idx_h.size());
      idx.upload(idx_h.data(), idx_h.size(), s);
      for (int it = 0; it < nsteps; ++it) {
        const int64_t* bi = idx.p + size_t(it) * nbat;
        launch_gather_rows(Xs.p, bi, so->batch, nin, 1.0f, xb.p, s);
        launch_gather_rows(Ts.p, bi, so->batch, 2, float(1.0 / sigma), tb.p, s);
        net->gradients(xb.p, tb.p, so->batch, s, false);
        net->adam(so->lr, s);
      }
      return nsteps > 0 ? net->last_loss(s) : 0.0;
    };

    // Step 0: solve every branch at its mu_0.
    if (bt->stage.n != size_t(bt->total_len)) bt->stage.alloc(bt->total_len);
    bt->stage.upload(psi, bt->total_len, s);
    scatter_all(bt, bt->stage.p, s);
    std::vector<nosh_batch_instance_stats> st(ni);
    {
      const auto t0 = std::chrono::steady_clock::now();
      set_mu();
      nosh_batch_continuation_step& r = record[0];
      r = nosh_batch_continuation_step{};
      r.minres_iters = newton_batch(bt, o, st.data(), s);
      for (int i = 0; i < ni; ++i) {
        r.newton_iters = std::max(r.newton_iters, st[i].newton_iters);
        r.converged += st[i].converged;
        r.max_norm_f = std::max(r.max_norm_f, st[i].final_norm_f);
      }
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      NOSH_CHECK(r.converged == ni, kNotConverged, "batch continuation: initial solve failed");
      energies(e1);
      e0 = e1;
      copy_state(&GroupCont::cur, true);
      // psi_{k-1} := psi_0: the first secant (ratio 0) then reads defined
      // values (0 * (cur - prev) of stale memory could be 0 * NaN).
      for (size_t q = 0; q < ng; ++q)
        NOSH_HIP(hipMemcpyAsync(gc[q].prev.p, gc[q].cur.p, sizeof(double2) * gc[q].cur.n,
                                hipMemcpyDeviceToDevice, s));
      done = 1;
    }
    const double h = span / steps;
    double hprev = 0.0;
    const bool progress = std::getenv("NOSH_PROGRESS") != nullptr;
    std::vector<int> ok(ni), skip(ni), use_sur(ni), maxit(ni), tried_sec(ni);
    int skip_until = 0, losses = 0;
    bool hard_prev = false;  // a branch of the previous step missed its first try
    std::vector<double> nf_sur, nf_sec;
    for (int k = 1; k <= steps; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      nosh_batch_continuation_step& r = record[k];
      r = nosh_batch_continuation_step{};
      r.param = k * h;
      bool hard = false;
      // Predictor 3: the tangent on the first step (no history) and after a
      // step that a branch did not pass from its first guess (a transition:
      // the secant through it points across the jump), the secant otherwise.
      const bool tangent = predictor == 1 || (predictor == 3 && (k == 1 || hard_prev));
      if (tangent) {
        // Tangent guess psi_k + h dpsi/dmu, J(psi_k) dpsi/dmu = -dF/dmu at
        // mu_k, all branches in one lockstep solve.
        for (int i = 0; i < ni; ++i) mu[i] = bt->mu_h[i] + (k - 1) * h;
        set_mu();
        copy_state(&GroupCont::cur, false);
        r.minres_iters += tangent_batch(bt, o, h, kTangentRtol, s);
        copy_state(&GroupCont::sec, true);
      }
      for (int i = 0; i < ni; ++i) mu[i] = bt->mu_h[i] + k * h;
      set_mu();
      if (!tangent) {
        // Secant guess (psi_k on the first step).
        const double ratio = hprev > 0.0 ? h / hprev : 0.0;
        for (size_t q = 0; q < ng; ++q) {
          GroupHost& G = B.ghost[q];
          const int64_t m = G.mesh->n * G.dev.bp;
          k_bc_secant<<<grid_for(m, 256), 256, 0, s>>>(m, ratio, gc[q].cur.p, gc[q].prev.p,
                                                       gc[q].sec.p);
        }
      }
      // The surrogate is evaluated while it earns its cost: when it gives the
      // smaller residual for fewer than a quarter of the branches, its next
      // evaluations are skipped for 2, 4, ... (at most 16) steps; a win
      // resets the back-off.
      const bool seeded = so && trained && hprev > 0.0 && k >= skip_until;
      std::fill(ok.begin(), ok.end(), 0);
      std::fill(tried_sec.begin(), tried_sec.end(), 0);
      if (seeded) {
        // Each branch starts from the guess with the smaller ||F||: the
        // surrogate's (Newton budget so->newton_budget) or the secant's (the
        // full iteration limit).
        predict(h * h * sigma);
        copy_state(&GroupCont::guess, false);
        batch_residuals(bt, nf_sur, s);
        copy_state(&GroupCont::sec, false);
        batch_residuals(bt, nf_sec, s);
        int nsur = 0;
        if (progress) {
          double rmin = 1e300, rmax = 0.0;
          for (int i = 0; i < ni; ++i) {
            const double q = nf_sur[i] / std::max(nf_sec[i], 1e-300);
            rmin = std::min(rmin, q);
            rmax = std::max(rmax, q);
          }
          std::fprintf(stderr, "[batch continuation] step %d probe: ||F(sur)||/||F(sec)|| in "
                               "[%.3g, %.3g], ||F(sec)|| %.3g\n", k, rmin, rmax, nf_sec[0]);
        }
        for (int i = 0; i < ni; ++i) {
          use_sur[i] = nf_sur[i] < so->accept_ratio * nf_sec[i];
          maxit[i] = use_sur[i] ? so->newton_budget : std::min(o.max_iters, kShortNewton);
          nsur += use_sur[i];
        }
        if (4 * nsur < ni) {
          losses = std::min(losses + 1, 4);
          skip_until = k + 1 + (1 << losses);
        } else {
          losses = 0;
        }
        if (nsur > 0) {
          // The branches left on psi_pred take their short try in the same
          // lockstep solve (and skip it below); when no branch takes the
          // surrogate's guess, no solve runs here and all take it below.
          for (int i = 0; i < ni; ++i) tried_sec[i] = !use_sur[i];
          take_into_psi(&GroupCont::guess, use_sur);
          const int it = newton_batch(bt, o, st.data(), s, nullptr, maxit.data(), 0, nullptr,
                                      kStallSteps);
          int mx = 0;
          for (int i = 0; i < ni; ++i) mx = std::max(mx, st[i].newton_iters);
          r.newton_iters += mx;
          r.minres_iters += it;
          energies(e_now);
          for (int i = 0; i < ni; ++i) {
            if (!use_sur[i]) {
              ok[i] = st[i].converged;
              continue;
            }
            const double extrap = e1[i] + (e1[i] - e0[i]);
            ok[i] = st[i].converged &&
                    std::fabs(e_now[i] - extrap) <=
                        std::max(so->energy_tol, 10.0 * std::fabs(e1[i] - e0[i]));
            r.seeded += ok[i];
            r.fallbacks += !ok[i];
          }
          sst.steps_seeded += r.seeded;
          sst.fallbacks += r.fallbacks;
        }
      }
      // Branches not yet accepted, in lockstep: Newton from psi_pred
      // (unless a branch already started from it) and then from psi_k, with
      // at most kShortNewton steps each; then, for the branches both leave,
      // pseudo-transient continuation from psi_pred and then from
      // psi_k. Near a branch transition plain Newton wanders (the old branch
      // has ended and J is near singular along the mode that carries the
      // state to the new one), and often converges from one of the two
      // guesses but not the other. The shifted Newton steps
      // (J + sigma c t) d = -F, sigma_0 = kPtcShift0 and sigma following ||F||
      // (switched evolution relaxation), damp that mode while the shift lasts
      // and reach the new state in about 10 steps where plain Newton took
      // 20-40 (96^3, 8 branches, mu 0.70 -> 0.75: 0.62 s instead of 4.1 s for
      // full-length FP32 retries without the shift and 15.7 s for the serial
      // FP64 fallback).
      for (int attempt = 0; attempt < 4; ++attempt) {
        const bool from_sec = (attempt % 2) == 0;
        const bool ptc = attempt >= 2;
        int todo = 0;
        for (int i = 0; i < ni; ++i) {
          skip[i] = ok[i] || (attempt == 0 && tried_sec[i]);
          todo += !skip[i];
          maxit[i] = ptc ? o.max_iters : std::min(o.max_iters, kShortNewton);
        }
        if (todo == 0) continue;
        if (attempt > 0) hard = true;
        const auto ta = std::chrono::steady_clock::now();
        std::vector<int> take(ni);
        for (int i = 0; i < ni; ++i) take[i] = !skip[i];
        take_into_psi(from_sec ? &GroupCont::sec : &GroupCont::cur, take);
        nosh_newton_options oa = o;
        std::vector<double> sig0;
        if (ptc) {
          oa.forcing_max = std::min(o.forcing_max, kPtcForcing);
          oa.cheb_iters = std::max(o.cheb_iters, kPtcCheb);
          sig0.assign(ni, kPtcShift0);
        }
        const int it = newton_batch(bt, oa, st.data(), s, skip.data(), maxit.data(), 0,
                                    ptc ? sig0.data() : nullptr, ptc ? 0 : kStallSteps);
        int mx = 0, nc = 0;
        for (int i = 0; i < ni; ++i)
          if (!skip[i]) {
            mx = std::max(mx, st[i].newton_iters);
            ok[i] = st[i].converged;
            nc += ok[i];
          }
        r.newton_iters += mx;
        r.minres_iters += it;
        if (ptc) r.retries += todo;
        if (progress && (ptc || nc < todo))
          std::fprintf(stderr,
                       "[batch continuation] step %d %s from %s: %d of %d converged, newton %d "
                       "minres %d, %.3fs\n",
                       k, ptc ? "pseudo-transient retry" : "short try",
                       from_sec ? "the predictor's guess" : "psi_k", nc, todo, mx, it,
                       std::chrono::duration<double>(std::chrono::steady_clock::now() - ta)
                           .count());
      }
      // Branches the short tries did not bring to convergence, one at a time
      // with the single-instance FP64 solver: Newton-MINRES (inner solves to
      // minres_rtol, no forcing) from psi_pred, then from psi_k, then
      // the tangent continuation over the step with up to kMaxHalvings step
      // halvings (the plain solver's way across a transition). A lockstep
      // launch costs about as much for one active lane as for all of them
      // (the lanes of a row share a wave), and FP64 inner solves cross a
      // transition in fewer Newton steps than FP32 inexact ones: at 160^3
      // this beat full-length lockstep tries even with all 8 branches here.
      {
        nosh_newton_options o64 = o;
        o64.precision = 0;
        o64.forcing = 0;
        o64.cheb_iters = kFallbackCheb;
        for (int i = 0; i < ni; ++i) {
          if (ok[i]) continue;
          ++r.serial_fallbacks;
          const auto tf = std::chrono::steady_clock::now();
          size_t g = 0;
          while (g + 1 < ng && i >= B.ghost[g + 1].dev.inst0) ++g;
          GroupHost& G = B.ghost[g];
          const int b = i - G.dev.inst0;
          const int64_t n = G.mesh->n;
          if (fb.n < size_t(2 * n)) fb.alloc(2 * n);
          std::map<std::string, double> p1;
          nosh_nls_s* md = bt->models[g];
          for (size_t q = 0; q < md->p_names.size(); ++q) p1[md->p_names[q]] = md->p_nominal[q];
          p1["g"] = bt->g_h[i];
          p1["mu"] = mu[i];
          int how = 0, fnewton = 0, fminres = 0, fhalv = 0;
          for (int tr = 0; tr < 2 && !ok[i]; ++tr) {
            k_bc_lane_get<<<grid_for(n, 256), 256, 0, s>>>(
                n, G.dev.bp, b, tr == 0 ? gc[g].sec.p : gc[g].cur.p,
                reinterpret_cast<double2*>(fb.p));
            nosh_newton_stats s64{};
            newton_run(bt->models[g], fb.p, p1, o64, &s64, s);
            fnewton += s64.newton_iters;
            fminres += s64.minres_iters_total;
            if (s64.converged) {
              ok[i] = 1;
              how = tr + 1;
              st[i].final_norm_f = s64.final_norm_f;
              k_bc_lane_put<<<grid_for(n, 256), 256, 0, s>>>(
                  n, G.dev.bp, b, reinterpret_cast<const double2*>(fb.p), G.psi.p);
            }
          }
          if (!ok[i]) {
            fb_h.resize(2 * n);
            launch_gather_lane(n, G.dev.bp, b, G.mesh->perm.p,
                               reinterpret_cast<const double*>(gc[g].cur.p), fb.p, s);
            fb.download(fb_h.data(), 2 * n, s);
            NOSH_HIP(hipStreamSynchronize(s));
            nosh_continuation_options co;
            nosh_continuation_default_options(&co);
            co.parameter = "mu";
            co.start = bt->mu_h[i] + (k - 1) * h;
            co.stop = mu[i];
            co.steps = 1;
            co.predictor = 1;
            co.max_halvings = kMaxHalvings;
            co.newton = o64;
            nosh_continuation_step crec[2];
            int cdone = 0;
            std::vector<const char*> cn;
            std::vector<double> cv;
            for (auto& kv : p1)
              if (kv.first != "mu") {
                cn.push_back(kv.first.c_str());
                cv.push_back(kv.second);
              }
            const int crc = nosh_continuation_run(bt->models[g], fb_h.data(), int(cn.size()),
                                                  cn.data(), cv.data(), &co, crec, &cdone);
            for (int q = 0; q < std::min(cdone + 1, 2); ++q) {
              fnewton += crec[q].newton_iters;
              fminres += crec[q].minres_iters;
            }
            if (crc == NOSH_OK && cdone == 2) {
              ok[i] = 1;
              how = 3;
              fhalv = crec[1].halvings;
              r.halvings = std::max(r.halvings, fhalv);
              st[i].final_norm_f = crec[1].norm_f;
              fb.upload(fb_h.data(), 2 * n, s);
              launch_scatter_lane(n, G.dev.bp, b, G.mesh->perm.p, fb.p,
                                  reinterpret_cast<double*>(G.psi.p), s);
            }
          }
          if (ok[i]) st[i].converged = 1;
          r.newton_iters += fnewton;
          r.minres_iters += fminres;
          if (progress)
            std::fprintf(stderr,
                         "[batch continuation] step %d branch %d: FP64 fallback %s, newton %d "
                         "minres %d halvings %d, %.3fs\n",
                         k, i,
                         how == 1 ? "converged from the predictor's guess"
                                  : how == 2 ? "converged from psi_k"
                                             : how == 3 ? "converged by tangent steps"
                                                        : "failed",
                         fnewton, fminres, fhalv,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - tf)
                             .count());
        }
      }
      // Branches still not converged (a transition none of the guesses
      // reaches): the step is split into 2, 4, ... sub-steps for them alone
      // (every instance has its own mu), from psi_k with a secant guess
      // scaled to the sub-step, then the secant over the sub-steps.
      for (int halv = 1; halv <= kMaxHalvings; ++halv) {
        std::vector<int> sf(ni);
        int nf = 0;
        for (int i = 0; i < ni; ++i) nf += (sf[i] = !ok[i]);
        if (nf == 0) break;
        r.halvings = halv;
        const int sub = 1 << halv;
        const double hs = h / sub;
        std::vector<int> skp(ni), full(ni, o.max_iters);
        for (int i = 0; i < ni; ++i) skp[i] = !sf[i];
        bool alive = true;
        for (int q = 0; q < sub && alive; ++q) {
          for (int i = 0; i < ni; ++i)
            if (sf[i]) mu[i] = bt->mu_h[i] + (k - 1) * h + (q + 1) * hs;
          set_mu();
          for (size_t g = 0; g < ng; ++g) {
            GroupHost& G = B.ghost[g];
            std::vector<int>& t = gc[g].take_h;
            t.assign(G.dev.bp, 0);
            for (int b = 0; b < G.B; ++b) t[b] = sf[G.dev.inst0 + b];
            gc[g].take.upload(t.data(), t.size(), s);
            const int64_t m = G.mesh->n * G.dev.bp;
            if (q == 0) {
              // psi = psi_k + (hs / hprev)(psi_k - psi_{k-1}); guess buffer := psi_k.
              k_bc_lin_take<<<grid_for(m, 256), 256, 0, s>>>(
                  G.mesh->n, G.dev.bp, gc[g].take.p, hprev > 0.0 ? hs / hprev : 0.0,
                  gc[g].cur.p, gc[g].prev.p, G.psi.p);
              k_bc_take<<<grid_for(m, 256), 256, 0, s>>>(G.mesh->n, G.dev.bp, gc[g].take.p,
                                                         gc[g].cur.p, gc[g].guess.p);
            } else {
              // psi <- psi + (psi - psi_before); guess buffer := the old psi.
              k_bc_advance_take<<<grid_for(m, 256), 256, 0, s>>>(G.mesh->n, G.dev.bp,
                                                                 gc[g].take.p, G.psi.p,
                                                                 gc[g].guess.p);
            }
          }
          const int it = newton_batch(bt, o, st.data(), s, skp.data(), full.data());
          int mx = 0;
          for (int i = 0; i < ni; ++i)
            if (sf[i]) mx = std::max(mx, st[i].newton_iters);
          r.newton_iters += mx;
          r.minres_iters += it;
          alive = true;
          for (int i = 0; i < ni; ++i)
            if (sf[i] && !st[i].converged) alive = false;
        }
        if (alive)
          for (int i = 0; i < ni; ++i)
            if (sf[i]) ok[i] = 1;
        for (int i = 0; i < ni; ++i) mu[i] = bt->mu_h[i] + k * h;
        set_mu();
      }
      for (int i = 0; i < ni; ++i) {
        r.converged += ok[i];
        r.max_norm_f = std::max(r.max_norm_f, st[i].final_norm_f);
      }
      if (r.converged < ni) {
        copy_state(&GroupCont::cur, false);
        r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        throw nosh::Error(kNotConverged, "batch continuation: a branch failed to converge");
      }
      // Predictor errors (mean over branches) for the record.
      double pe = 0.0, se = 0.0;
      for (size_t q = 0; q < ng; ++q) {
        GroupHost& G = B.ghost[q];
        lane_sums(q, 1, G.psi.p, gc[q].sec.p);
        for (int b = 0; b < G.B; ++b) se += std::sqrt(lane_h[b]);
        if (seeded) {
          lane_sums(q, 1, G.psi.p, gc[q].guess.p);
          for (int b = 0; b < G.B; ++b) pe += std::sqrt(lane_h[b]);
        }
      }
      r.sec_error = se / ni;
      r.pred_error = seeded ? pe / ni : r.sec_error;
      // Training pairs from whole secant steps, kept in a ring of P steps.
      // Steps whose Newton needed more than the budget (branch transitions)
      // are not stored: their targets dwarf the smooth ones.
      if (so && hprev > 0.0 && (!trained || so->retrain_steps > 0) && r.halvings == 0 &&
          r.newton_iters <= so->newton_budget) {
        const int slot = pairs % P;
        for (size_t q = 0; q < ng; ++q) {
          GroupHost& G = B.ghost[q];
          const int64_t rows = G.mesh->n * G.B;
          k_bc_features<<<grid_for(rows, 256), 256, 0, s>>>(
              G.mesh->n, G.dev.bp, G.B, dim, G.mesh->perm.p, G.mesh->coords.p, cs, gc[q].cur.p,
              gc[q].sec.p, B.mu.p, B.g.p, G.dev.inst0,
              Xs.p + (int64_t(slot) * xrows + gc[q].xoff) * nin);
          k_bc_target<<<grid_for(rows, 256), 256, 0, s>>>(
              G.mesh->n, G.dev.bp, G.B, G.psi.p, gc[q].sec.p, 1.0 / (h * h),
              Ts.p + (int64_t(slot) * xrows + gc[q].xoff) * 2);
        }
        ++pairs;
        if (!trained && pairs == P) {
          const auto tt = std::chrono::steady_clock::now();
          sst.final_loss = train(so->train_steps);
          sst.train_seconds =
              std::chrono::duration<double>(std::chrono::steady_clock::now() - tt).count();
          sst.trained = 1;
          trained = true;
        } else if (trained) {
          const auto tt = std::chrono::steady_clock::now();
          sst.final_loss = train(so->retrain_steps);
          sst.train_seconds +=
              std::chrono::duration<double>(std::chrono::steady_clock::now() - tt).count();
        }
      }
      energies(e_now);
      e0 = e1;
      e1 = e_now;
      for (size_t q = 0; q < ng; ++q) std::swap(gc[q].prev, gc[q].cur);
      copy_state(&GroupCont::cur, true);
      hprev = h;
      hard_prev = hard;
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      done = k + 1;
      if (progress)
        std::fprintf(stderr,
                     "[batch continuation] step %d/%d dmu %.4f newton %d minres %d seeded %d "
                     "%.3fs err(sur) %.3g err(sec) %.3g\n",
                     k, steps, r.param, r.newton_iters, r.minres_iters, r.seeded, r.seconds,
                     r.pred_error, r.sec_error);
    }
  });
  if (bt && psi && done > 0) {
    hipStream_t s = default_stream();
    gather_all(bt, bt->stage.p, s);
    bt->stage.download(psi, bt->total_len, s);
    (void)hipStreamSynchronize(s);
    // The batch keeps its per-instance mu_0 for later solves.
    bt->b.set_mu(bt->mu_h.data(), s);
    (void)hipStreamSynchronize(s);
  }
  if (steps_done) *steps_done = done;
  if (stats) *stats = sst;
  return rc;
}

}  // extern "C"
