// continuation.cpp -- natural-parameter continuation over Newton-MINRES, and
// the energy functionals of the model.
//
// Stands in for the LOCA stepper the reference is driven by
// (examples/conf.xml:35-75: tangent predictor, adaptive step with a minimum
// size) in its simplest form: a fixed parameter grid, tangent / secant /
// trivial predictor, and step halving when the corrector fails.

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "handles.hpp"

using namespace nosh;

namespace {

template <class F>
int cguard(F&& f) {
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

struct ContWork {
  DBuf<double> psi_prev, t, rhs, partials, scal;
  std::unique_ptr<nosh_keo_s> dk;  // dK/d(parameter)
};

__global__ void k_axpy2(int64_t m, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) y[i] += a * x[i];
}

// y = x + a (x - z)   (secant extrapolation in place of y).
__global__ void k_secant(int64_t m, double a, const double* __restrict__ x,
                         const double* __restrict__ z, double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) y[i] = x[i] + a * (x[i] - z[i]);
}

double gibbs_dev(nosh_nls_s* nls, ContWork& w, const double* psi_int, hipStream_t s) {
  MeshCache& m = *nls->mesh->m;
  launch_cv_weighted(m, psi_int, nullptr, w.partials.p, w.scal.p, s);
  double e = 0.0;
  NOSH_HIP(hipMemcpyAsync(&e, w.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  return e;
}

// Tangent t with J(psi) t = -dF/dp(psi); returns MINRES iterations.
int tangent(nosh_nls_s* nls, ContWork& w, const std::string& param, double* psi_int,
            const std::map<std::string, double>& p, const nosh_newton_options& o,
            hipStream_t s) {
  MeshCache& m = *nls->mesh->m;
  const int64_t n = m.n;
  const double g = need(p, "g");
  nls->keo->set_parameters(p, s);
  w.dk->set_parameters(p, s);
  compute_ct(m, nls->thickness, nls->ct, p, s);
  const double* V = nls->V->sf.get_v(p, s);
  // dF/dp: dK/dp psi + (p == "g" ? ct |psi|^2 psi : ct dV/dp psi).
  launch_dfdp(m, w.dk->kv, nls->ct.p, psi_int, param == "g", nls->V->sf.dvdp_constant(param),
              nls->V->sf.dvdp_values(param, s), w.rhs.p, s);
  if (nls->jblk.n != size_t(3 * n)) {
    nls->jblk.alloc(3 * n);
    nls->pblk.alloc(3 * n);
  }
  launch_rebuild_blocks(m, nls->keo->kv, nls->ct.p, V, g, psi_int, nls->jblk.p, nls->pblk.p,
                        nls->bound.p, s);
  NOSH_HIP(hipMemsetAsync(w.t.p, 0, sizeof(double) * 2 * n, s));
  double rr = 0.0;
  if (o.precision == 1) {
    if (!nls->keo->kv.valid32) launch_keo_to_fp32(m, nls->keo->kv, s);
    if (nls->jblk32.n != size_t(3 * n)) {
      nls->jblk32.alloc(3 * n);
      nls->pblk32.alloc(3 * n);
    }
    convert_to<float>(nls->jblk.p, nls->jblk32.p, 3 * n, s);
    convert_to<float>(nls->pblk.p, nls->pblk32.p, 3 * n, s);
    JacView<float> J{nls->keo->kv.val32.p, nls->jblk32.p};
    PrecView<float> P{nls->keo->kv.val32.p, nls->pblk32.p, nls->bound.p, 0.0, o.cheb_ratio,
                      o.cheb_iters};
    return minres_solve<float>(m, J, P, nls->w32, w.rhs.p, -1.0, std::max(o.forcing_min, 1e-6),
                               o.minres_max_iters, w.t.p, s, &rr, 8);
  }
  JacView<double> J{nls->keo->kv.val.p, nls->jblk.p};
  PrecView<double> P{nls->keo->kv.val.p, nls->pblk.p, nls->bound.p, 0.0, o.cheb_ratio,
                     o.cheb_iters};
  return minres_solve<double>(m, J, P, nls->w64, w.rhs.p, -1.0, 1e-8, o.minres_max_iters, w.t.p,
                              s, &rr, 8);
}

}  // namespace

extern "C" {

int nosh_nls_gibbs_energy(nosh_nls_t nls, const double* psi, double* energy) {
  return cguard([&] {
    NOSH_CHECK(nls && psi && energy, kInvalid, "null argument");
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    ContWork w;
    w.partials.alloc(2 * grid_for(m.n, 256));
    w.scal.alloc(1);
    nls->st.ensure(2 * m.n);
    nls->st.a.upload(psi, 2 * m.n, s);
    launch_permute_in(m.n, m.perm.p, nls->st.a.p, nls->st.b.p, s);
    *energy = gibbs_dev(nls, w, nls->st.b.p, s);
  });
}

int nosh_nls_inner_product(nosh_nls_t nls, const double* phi, const double* psi, double* value) {
  return cguard([&] {
    NOSH_CHECK(nls && phi && psi && value, kInvalid, "null argument");
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    DBuf<double> partials(2 * grid_for(m.n, 256)), scal(1);
    nls->st.ensure(2 * m.n);
    nls->st.a.upload(phi, 2 * m.n, s);
    launch_permute_in(m.n, m.perm.p, nls->st.a.p, nls->st.b.p, s);
    nls->st.a.upload(psi, 2 * m.n, s);
    launch_permute_in(m.n, m.perm.p, nls->st.a.p, nls->st.c.p, s);
    launch_cv_weighted(m, nls->st.b.p, nls->st.c.p, partials.p, scal.p, s);
    NOSH_HIP(hipMemcpyAsync(value, scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_continuation_default_options(nosh_continuation_options* o) {
  return cguard([&] {
    NOSH_CHECK(o, kInvalid, "null options");
    o->parameter = "mu";
    o->start = 0.0;
    o->stop = 1.0;
    o->steps = 10;
    o->predictor = 1;
    o->max_halvings = 4;
    nosh_newton_default_options(&o->newton);
  });
}

int nosh_continuation_run(nosh_nls_t nls, double* psi, int np, const char* const* names,
                          const double* values, const nosh_continuation_options* opts,
                          nosh_continuation_step* record, int* steps_done) {
  int done = 0;
  int rc = cguard([&] {
    NOSH_CHECK(nls && psi && opts && record, kInvalid, "null argument");
    NOSH_CHECK(opts->steps >= 1, kInvalid, "steps must be >= 1");
    const std::string param = opts->parameter ? opts->parameter : "mu";
    NOSH_CHECK(param == "mu" || param == "g", kInvalid, "continuation parameter must be mu or g");
    std::map<std::string, double> p;
    for (int i = 0; i < np; ++i) p[names[i]] = values[i];
    MeshCache& m = *nls->mesh->m;
    const int64_t n = m.n;
    hipStream_t s = default_stream();
    ContWork w;
    w.psi_prev.alloc(2 * n);
    w.t.alloc(2 * n);
    w.rhs.alloc(2 * n);
    w.partials.alloc(2 * grid_for(n, 256));
    w.scal.alloc(1);
    w.dk = std::make_unique<nosh_keo_s>();
    w.dk->mesh = nls->mesh;
    w.dk->thickness = nls->thickness;
    w.dk->mvp = nls->mvp;
    w.dk->dkeo = true;
    w.dk->dparam = param;
    if (nls->x.n != size_t(2 * n)) nls->x.alloc(2 * n);
    double* x = nls->x.p;
    nls->st.ensure(2 * n);
    nls->st.a.upload(psi, 2 * n, s);
    launch_permute_in(n, m.perm.p, nls->st.a.p, x, s);
    const unsigned g2 = grid_for(2 * n, 256);
    const double h = (opts->stop - opts->start) / opts->steps;

    auto solve_at = [&](double pv, nosh_newton_stats* st) {
      p[param] = pv;
      newton_run(nls, x, p, opts->newton, st, s);
    };
    // Entry 0: the solve at `start`.
    {
      const auto t0 = std::chrono::steady_clock::now();
      nosh_newton_stats st{};
      solve_at(opts->start, &st);
      nosh_continuation_step& r = record[0];
      r = nosh_continuation_step{};
      r.param = opts->start;
      r.newton_iters = st.newton_iters;
      r.minres_iters = st.minres_iters_total;
      r.converged = st.converged;
      r.norm_f = st.final_norm_f;
      r.gibbs = gibbs_dev(nls, w, x, s);
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      NOSH_CHECK(st.converged, kNotConverged, "continuation: initial Newton solve failed");
      done = 1;
    }
    double p_cur = opts->start, h_prev = 0.0;
    bool have_prev = false;
    for (int k = 1; k <= opts->steps; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      const double target = opts->start + k * h;
      nosh_continuation_step& r = record[k];
      r = nosh_continuation_step{};
      r.param = target;
      int sub = 1, halvings = 0;
      bool ok = false;
      while (!ok) {
        const double hs = (target - p_cur) / sub;
        // Predictor from (p_cur, x).
        NOSH_HIP(hipMemcpyAsync(w.psi_prev.p, x, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice,
                                s));
        double p_base = p_cur;
        bool sub_ok = true;
        std::vector<double> keep_prev;
        for (int q = 0; q < sub && sub_ok; ++q) {
          if (opts->predictor == 1) {
            p[param] = p_base;
            r.minres_iters += tangent(nls, w, param, x, p, opts->newton, s);
            k_axpy2<<<g2, 256, 0, s>>>(2 * n, hs, w.t.p, x);
          } else if (opts->predictor == 2 && have_prev && h_prev != 0.0) {
            // x <- x + (hs / h_prev)(x - x_prev_accepted); the previous
            // accepted state is in w.rhs (saved below).
            k_secant<<<g2, 256, 0, s>>>(2 * n, hs / h_prev, x, w.rhs.p, x);
          }
          nosh_newton_stats st{};
          solve_at(p_base + hs, &st);
          r.newton_iters += st.newton_iters;
          r.minres_iters += st.minres_iters_total;
          r.norm_f = st.final_norm_f;
          sub_ok = st.converged != 0;
          if (sub_ok) p_base += hs;
        }
        if (sub_ok) {
          ok = true;
          if (opts->predictor == 2) {
            // Keep the previous accepted state for the next secant step.
            NOSH_HIP(hipMemcpyAsync(w.rhs.p, w.psi_prev.p, sizeof(double) * 2 * n,
                                    hipMemcpyDeviceToDevice, s));
            have_prev = true;
            h_prev = target - p_cur;
          }
        } else {
          NOSH_HIP(hipMemcpyAsync(x, w.psi_prev.p, sizeof(double) * 2 * n,
                                  hipMemcpyDeviceToDevice, s));
          if (halvings >= opts->max_halvings) break;
          ++halvings;
          sub *= 2;
        }
      }
      r.halvings = halvings;
      r.predictor_used = opts->predictor;
      r.converged = ok ? 1 : 0;
      r.gibbs = gibbs_dev(nls, w, x, s);
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (!ok) throw nosh::Error(kNotConverged, "continuation: corrector failed after halvings");
      p_cur = target;
      done = k + 1;
      if (std::getenv("NOSH_PROGRESS"))
        std::fprintf(stderr, "[continuation] step %d/%d %s %.4f newton %d minres %d %.3fs\n", k,
                     opts->steps, param.c_str(), target, r.newton_iters, r.minres_iters,
                     r.seconds);
    }
  });
  if (nls && psi && done > 0) {
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    launch_permute_out(m.n, m.perm.p, nls->x.p, nls->st.d.p, s);
    nls->st.d.download(psi, 2 * m.n, s);
    (void)hipStreamSynchronize(s);
  }
  if (steps_done) *steps_done = done;
  return rc;
}

}  // extern "C"


// ------------------------------------------------------------------------
// Surrogate-seeded continuation.

#include <random>

#include "mlp.hpp"

namespace {


// X[i] = (x*cs, y*cs[, z*cs], Re psi, Im psi, h Re t, h Im t, mu) (internal order).
__global__ void k_features(int64_t n, int dim, const int32_t* __restrict__ perm,
                           const double* __restrict__ coords, double cs,
                           const double2* __restrict__ psi, const double2* __restrict__ t,
                           double h, double mu, float* __restrict__ X) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nin = dim + 5;
  float* x = X + i * nin;
  const int64_t v = perm[i];
  for (int d = 0; d < dim; ++d) x[d] = float(coords[3 * v + d] * cs);
  const double2 p = psi[i], tt = t[i];
  x[dim] = float(p.x);
  x[dim + 1] = float(p.y);
  x[dim + 2] = float(h * tt.x);
  x[dim + 3] = float(h * tt.y);
  x[dim + 4] = float(mu);
}

// T[i] = (psi_new - psi_old - h t) / h^2.
__global__ void k_target(int64_t n, const double2* __restrict__ pn, const double2* __restrict__ po,
                         const double2* __restrict__ t, double h, float* __restrict__ T) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 a = pn[i], b = po[i], tt = t[i];
  const double ih2 = 1.0 / (h * h);
  T[2 * i] = float((a.x - b.x - h * tt.x) * ih2);
  T[2 * i + 1] = float((a.y - b.y - h * tt.y) * ih2);
}

// out = psi_old + h t + h^2 sigma Y.
__global__ void k_apply_guess(int64_t n, const double2* __restrict__ po,
                              const double2* __restrict__ t, double h,
                              const float* __restrict__ Y, double sigma,
                              double2* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 b = po[i], tt = t[i];
  const double s = h * h * sigma;
  out[i] = make_double2(b.x + h * tt.x + s * double(Y[2 * i]),
                        b.y + h * tt.y + s * double(Y[2 * i + 1]));
}

__global__ void k_diff_partials(int64_t m, const double* __restrict__ a,
                                const double* __restrict__ b, double* __restrict__ partials) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < m) {
    const double d = a[i] - b[i];
    v[0] = d * d;
  }
  block_partials<1>(v, partials);
}

__global__ void k_diff_final(const double* __restrict__ partials, int nb, double* out) {
  double v[1];
  finalize_sum<1>(v, partials, nb);
  if (threadIdx.x == 0) *out = sqrt(v[0]);
}

double diff_norm(int64_t m, const double* a, const double* b, DBuf<double>& part,
                 DBuf<double>& scal, hipStream_t s) {
  const unsigned nb = grid_for(m, 256);
  if (part.n < nb) part.alloc(nb);
  k_diff_partials<<<nb, 256, 0, s>>>(m, a, b, part.p);
  k_diff_final<<<1, 256, 0, s>>>(part.p, int(nb), scal.p);
  double r = 0.0;
  NOSH_HIP(hipMemcpyAsync(&r, scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  return r;
}

}  // namespace

extern "C" {

int nosh_surrogate_default_options(nosh_surrogate_options* o) {
  return cguard([&] {
    NOSH_CHECK(o, kInvalid, "null options");
    o->width = 256;
    o->depth = 4;
    o->train_after = 40;
    o->train_steps = 300;
    o->batch = 16384;
    o->lr = 1e-3f;
    o->newton_budget = 4;
    o->energy_tol = 1e-8;
    o->seed = 3;
    o->retrain_steps = 0;
    o->smooth_sweeps = 0;
    o->accept_ratio = 0.5;
  });
}

int nosh_continuation_run_surrogate(nosh_nls_t nls, double* psi, int np, const char* const* names,
                                    const double* values, const nosh_continuation_options* opts,
                                    const nosh_surrogate_options* so,
                                    nosh_continuation_step* record, int* steps_done,
                                    nosh_surrogate_stats* stats) {
  int done = 0;
  nosh_surrogate_stats sst{};
  int rc = cguard([&] {
    NOSH_CHECK(nls && psi && opts && so && record, kInvalid, "null argument");
    NOSH_CHECK(opts->steps >= 1 && so->train_after >= 1 && so->batch >= 64 &&
                   so->retrain_steps >= 0 && so->accept_ratio > 0.0,
               kInvalid, "bad continuation / surrogate options");
    const std::string param = opts->parameter ? opts->parameter : "mu";
    NOSH_CHECK(param == "mu" || param == "g", kInvalid, "continuation parameter must be mu or g");
    std::map<std::string, double> p;
    for (int i = 0; i < np; ++i) p[names[i]] = values[i];
    MeshCache& m = *nls->mesh->m;
    const int64_t n = m.n;
    const int dim = m.nv == 3 ? 2 : 3;
    const int nin = dim + 5;
    hipStream_t s = default_stream();
    ContWork w;
    w.psi_prev.alloc(2 * n);
    w.t.alloc(2 * n);
    w.rhs.alloc(2 * n);
    w.partials.alloc(2 * grid_for(n, 256) + 64);
    w.scal.alloc(1);
    w.dk = std::make_unique<nosh_keo_s>();
    w.dk->mesh = nls->mesh;
    w.dk->thickness = nls->thickness;
    w.dk->mvp = nls->mvp;
    w.dk->dkeo = true;
    w.dk->dparam = param;
    DBuf<double> tanbuf(2 * n), guess(2 * n), dpart, dscal(1);
    DBuf<float> X(n * nin), Y(n * 2);
    // Coordinate scale: 1 / max |coordinate|.
    std::vector<double> ch(3 * n);
    m.coords.download(ch.data(), 3 * n, s);
    NOSH_HIP(hipStreamSynchronize(s));
    double cmax = 1e-300;
    for (double v : ch) cmax = std::max(cmax, std::fabs(v));
    const double cs = 1.0 / cmax;
    Mlp net;
    net.init(nin, so->width, so->depth, 2, so->seed);
    const int P = so->train_after;
    DBuf<float> Xs(int64_t(P) * n * nin), Ts(int64_t(P) * n * 2);
    int pairs = 0;
    double sigma = 1.0;
    bool trained = false;

    if (nls->x.n != size_t(2 * n)) nls->x.alloc(2 * n);
    double* x = nls->x.p;
    nls->st.ensure(2 * n);
    nls->st.a.upload(psi, 2 * n, s);
    launch_permute_in(n, m.perm.p, nls->st.a.p, x, s);
    const unsigned gn = grid_for(n, 256);
    const double h = (opts->stop - opts->start) / opts->steps;
    auto newton_at = [&](double pv, int maxit, nosh_newton_stats* st) {
      p[param] = pv;
      nosh_newton_options o = opts->newton;
      o.max_iters = maxit;
      newton_run(nls, x, p, o, st, s);
    };
    {
      const auto t0 = std::chrono::steady_clock::now();
      nosh_newton_stats st{};
      newton_at(opts->start, opts->newton.max_iters, &st);
      nosh_continuation_step& r = record[0];
      r = nosh_continuation_step{};
      r.param = opts->start;
      r.newton_iters = st.newton_iters;
      r.minres_iters = st.minres_iters_total;
      r.converged = st.converged;
      r.norm_f = st.final_norm_f;
      r.gibbs = gibbs_dev(nls, w, x, s);
      r.predictor_used = 1;
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      NOSH_CHECK(st.converged, kNotConverged, "continuation: initial Newton solve failed");
      done = 1;
    }
    double p_cur = opts->start;
    std::mt19937_64 rng(so->seed);
    DBuf<int64_t> idx(so->batch);
    DBuf<float> xb(int64_t(so->batch) * nin), tb(int64_t(so->batch) * 2);
    std::vector<int64_t> idx_h(so->batch);
    // Adam steps on random vertex minibatches of the stored pairs.
    // Every step's minibatch is drawn up front (the same draws in the same
    // order as one step at a time) and uploaded once, and the loss read back
    // after the last step only, so the steps queue without host round trips.
    auto train_steps = [&](int nsteps) {
      std::uniform_int_distribution<int64_t> pick(0, int64_t(P) * n - 1);
      const size_t nbat = size_t(so->batch);
      idx_h.resize(nbat * size_t(std::max(nsteps, 1)));
      for (auto& v : idx_h) v = pick(rng);
      if (idx.n < idx_h.size()) idx.alloc(idx_h.size());
      idx.upload(idx_h.data(), idx_h.size(), s);
      for (int it = 0; it < nsteps; ++it) {
        const int64_t* bi = idx.p + size_t(it) * nbat;
        launch_gather_rows(Xs.p, bi, so->batch, nin, 1.0f, xb.p, s);
        launch_gather_rows(Ts.p, bi, so->batch, 2, float(1.0 / sigma), tb.p, s);
        net.gradients(xb.p, tb.p, so->batch, s, false);
        net.adam(so->lr, s);
      }
      return nsteps > 0 ? net.last_loss(s) : 0.0;
    };
    for (int k = 1; k <= opts->steps; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      const double target = opts->start + k * h;
      nosh_continuation_step& r = record[k];
      r = nosh_continuation_step{};
      r.param = target;
      const double hs = target - p_cur;
      // Tangent at (p_cur, x).
      p[param] = p_cur;
      r.minres_iters += tangent(nls, w, param, x, p, opts->newton, s);
      NOSH_HIP(hipMemcpyAsync(w.psi_prev.p, x, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice,
                              s));
      NOSH_HIP(hipMemcpyAsync(tanbuf.p, x, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
      k_axpy2<<<grid_for(2 * n, 256), 256, 0, s>>>(2 * n, hs, w.t.p, tanbuf.p);
      k_features<<<gn, 256, 0, s>>>(n, dim, m.perm.p, m.coords.p, cs,
                                    reinterpret_cast<const double2*>(w.psi_prev.p),
                                    reinterpret_cast<const double2*>(w.t.p), hs, p_cur, X.p);
      bool ok = false, use_guess = false, probed = false;
      nosh_newton_stats st{};
      if (trained) {
        net.infer(X.p, n, Y.p, s);
        k_apply_guess<<<gn, 256, 0, s>>>(n, reinterpret_cast<const double2*>(w.psi_prev.p),
                                         reinterpret_cast<const double2*>(w.t.p), hs, Y.p,
                                         sigma, reinterpret_cast<double2*>(guess.p));
        probed = true;
        // Newton starts from the surrogate guess only when its residual is at
        // most half the tangent guess's: a guess that is only slightly better
        // saves no Newton step.
        p[param] = target;
        const double nf_g = residual_norm(nls, guess.p, p, s);
        const double nf_t = residual_norm(nls, tanbuf.p, p, s);
        use_guess = nf_g <= so->accept_ratio * nf_t;
      }
      if (use_guess) {
        NOSH_HIP(hipMemcpyAsync(x, guess.p, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
        newton_at(target, so->newton_budget, &st);
        r.newton_iters += st.newton_iters;
        r.minres_iters += st.minres_iters_total;
        if (st.converged) {
          const double e = gibbs_dev(nls, w, x, s);
          const double e1 = record[k - 1].gibbs, e0 = k >= 2 ? record[k - 2].gibbs : e1;
          const double extrap = e1 + (e1 - e0);
          ok = std::fabs(e - extrap) <= std::max(so->energy_tol, 10.0 * std::fabs(e1 - e0));
        }
        if (ok) {
          r.predictor_used = 3;
          ++sst.steps_seeded;
        } else {
          r.fallback = 1;
          ++sst.fallbacks;
        }
      }
      int halvings = 0;
      if (!ok) {
        // Tangent predictor; if Newton fails from it, the step is split in
        // 2, 4, ... tangent sub-steps (as the plain driver does).
        int sub = 1;
        for (;;) {
          double p_base = p_cur;
          bool sub_ok = true;
          const double hsub = hs / sub;
          for (int q = 0; q < sub && sub_ok; ++q) {
            if (sub == 1) {
              NOSH_HIP(hipMemcpyAsync(x, tanbuf.p, sizeof(double) * 2 * n,
                                      hipMemcpyDeviceToDevice, s));
            } else {
              if (q == 0)
                NOSH_HIP(hipMemcpyAsync(x, w.psi_prev.p, sizeof(double) * 2 * n,
                                        hipMemcpyDeviceToDevice, s));
              p[param] = p_base;
              r.minres_iters += tangent(nls, w, param, x, p, opts->newton, s);
              k_axpy2<<<grid_for(2 * n, 256), 256, 0, s>>>(2 * n, hsub, w.t.p, x);
            }
            newton_at(p_base + hsub, opts->newton.max_iters, &st);
            r.newton_iters += st.newton_iters;
            r.minres_iters += st.minres_iters_total;
            sub_ok = st.converged != 0;
            if (sub_ok) p_base += hsub;
          }
          if (sub_ok) {
            ok = true;
            break;
          }
          if (halvings >= opts->max_halvings) break;
          ++halvings;
          sub *= 2;
        }
        r.predictor_used = 1;
      }
      r.halvings = halvings;
      r.norm_f = st.final_norm_f;
      r.converged = ok ? 1 : 0;
      if (!ok) {
        NOSH_HIP(hipMemcpyAsync(x, w.psi_prev.p, sizeof(double) * 2 * n,
                                hipMemcpyDeviceToDevice, s));
        throw nosh::Error(kNotConverged, "surrogate continuation: corrector failed after halvings");
      }
      r.gibbs = gibbs_dev(nls, w, x, s);
      r.tan_error = diff_norm(2 * n, tanbuf.p, x, dpart, dscal, s);
      r.pred_error = probed ? diff_norm(2 * n, guess.p, x, dpart, dscal, s) : r.tan_error;
      // Training pairs from the run's own solutions (whole tangent steps
      // only: sub-steps overwrite the tangent the features were built from),
      // kept in a ring of P steps.
      if ((!trained || so->retrain_steps > 0) && halvings == 0) {
        const int slot = pairs % P;
        NOSH_HIP(hipMemcpyAsync(Xs.p + int64_t(slot) * n * nin, X.p, sizeof(float) * n * nin,
                                hipMemcpyDeviceToDevice, s));
        k_target<<<gn, 256, 0, s>>>(n, reinterpret_cast<const double2*>(x),
                                    reinterpret_cast<const double2*>(w.psi_prev.p),
                                    reinterpret_cast<const double2*>(w.t.p), hs,
                                    Ts.p + int64_t(slot) * n * 2);
        ++pairs;
        if (!trained && pairs == P) {
          const auto tt = std::chrono::steady_clock::now();
          sigma = rms_f32(Ts.p, int64_t(P) * n * 2, s);
          if (!(sigma > 0.0)) sigma = 1.0;
          sst.final_loss = train_steps(so->train_steps);
          trained = true;
          sst.trained = 1;
          sst.train_seconds =
              std::chrono::duration<double>(std::chrono::steady_clock::now() - tt).count();
        } else if (trained) {
          const auto tt = std::chrono::steady_clock::now();
          sst.final_loss = train_steps(so->retrain_steps);
          sst.train_seconds +=
              std::chrono::duration<double>(std::chrono::steady_clock::now() - tt).count();
        }
      }
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      p_cur = target;
      done = k + 1;
      if (std::getenv("NOSH_PROGRESS"))
        std::fprintf(stderr,
                     "[surrogate continuation] step %d/%d %s %.4f newton %d predictor %d %.3fs\n",
                     k, opts->steps, param.c_str(), target, r.newton_iters, r.predictor_used,
                     r.seconds);
    }
  });
  if (nls && psi && done > 0) {
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    launch_permute_out(m.n, m.perm.p, nls->x.p, nls->st.d.p, s);
    nls->st.d.download(psi, 2 * m.n, s);
    (void)hipStreamSynchronize(s);
  }
  if (steps_done) *steps_done = done;
  if (stats) *stats = sst;
  return rc;
}

}  // extern "C"
