// arclength.cpp -- pseudo-arclength continuation over the Newton-MINRES
// kernels, with optional stability eigenvalues per accepted step.
//
// Stands in for LOCA's "Arc Length" stepper with the tangent predictor and
// the adaptive step size of examples/conf.xml:35-75 (driven from
// executables/nosh-cont/nosh-cont.cpp:226-290). LOCA is not part of the
// reference sources; what is restated is its documented algorithm:
// pseudo-arclength (Keller) with the bordering solve for the extended
// system, and the adaptive step rule
//   ds_new = ds (1 + a ((N - k) / (N - 1))^2),
// k the Newton steps of the last corrector, N their maximum, a the
// aggressiveness, capped at the maximum step; a failed corrector halves ds.
//
// Extended system for (psi, p) with the metric theta^2 <a, b> + alpha beta:
//   F(psi, p) = 0,
//   theta^2 <psi', psi - psi_0> + p' (p - p_0) - ds = 0.
// Bordering: J a = -F, J b = -dF/dp, dp = (-N - theta^2 <psi', a>) /
// (p' + theta^2 <psi', b>), dpsi = a + dp b; the arclength equation then
// holds exactly after every correction (it is linear). J is singular along
// the phase mode i psi at every solution; both right-hand sides are
// orthogonal to it (<i psi, F> = 0 and <i psi, dF/dp> = 0 identically), so
// MINRES returns the minimum-norm corrections.

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "eigen.hpp"

using namespace nosh;

namespace {

template <class F>
int aguard(F&& f) {
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

// y = x + a t
__global__ void k_arc_axpy(int64_t m, const double* __restrict__ x, double a,
                           const double* __restrict__ t, double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) y[i] = x[i] + a * t[i];
}

// t *= c
__global__ void k_arc_scale(int64_t m, double c, double* __restrict__ t) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) t[i] *= c;
}

// x += a + d b
__global__ void k_arc_correct(int64_t m, const double* __restrict__ a, double d,
                              const double* __restrict__ b, double* __restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) x[i] += a[i] + d * b[i];
}

// t = c (x - y)
__global__ void k_arc_diff(int64_t m, double c, const double* __restrict__ x,
                           const double* __restrict__ y, double* __restrict__ t) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) t[i] = c * (x[i] - y[i]);
}

struct Arc {
  nosh_nls_s* nls;
  MeshCache& m;
  hipStream_t s;
  int64_t n, L;
  std::string param;
  std::map<std::string, double> p;
  nosh_newton_options no;
  double border_rtol;
  std::unique_ptr<nosh_keo_s> dk;
  DBuf<double> x0, tp, tprev, xprev, a, b, fl;
  BlockWork bw;
  unsigned g2;

  Arc(nosh_nls_s* nl, hipStream_t st) : nls(nl), m(*nl->mesh->m), s(st) {
    n = m.n;
    L = 2 * n;
    g2 = grid_for(L, 256);
  }

  double dot(const double* u, const double* v) {
    double r = 0.0;
    block_dots(&u, 1, &v, 1, L, bw, &r, s);
    return r;
  }

  // F at (x, lam) into nls->f; returns ||F||.
  double residual(const double* x, double lam) {
    p[param] = lam;
    return residual_norm(nls, x, p, s);
  }

  // J and P blocks and dF/dp at (x, lam).
  void rebuild(const double* x, double lam) {
    p[param] = lam;
    const double g = need(p, "g");
    nls->keo->set_parameters(p, s);
    dk->set_parameters(p, s);
    compute_ct(m, nls->thickness, nls->ct, p, s);
    const double* V = nls->V->sf.get_v(p, s);
    if (nls->jblk.n != size_t(3 * n)) {
      nls->jblk.alloc(3 * n);
      nls->pblk.alloc(3 * n);
    }
    if (nls->bound.n < 1) nls->bound.alloc(1);
    launch_dfdp(m, dk->kv, nls->ct.p, x, param == "g", nls->V->sf.dvdp_constant(param),
                nls->V->sf.dvdp_values(param, s), fl.p, s);
    launch_rebuild_blocks(m, nls->keo->kv, nls->ct.p, V, g, x, nls->jblk.p, nls->pblk.p,
                          nls->bound.p, s);
    if (no.precision == 1) {
      if (!nls->keo->kv.valid32) launch_keo_to_fp32(m, nls->keo->kv, s);
      if (nls->jblk32.n != size_t(3 * n)) {
        nls->jblk32.alloc(3 * n);
        nls->pblk32.alloc(3 * n);
      }
      convert_to<float>(nls->jblk.p, nls->jblk32.p, 3 * n, s);
      convert_to<float>(nls->pblk.p, nls->pblk32.p, 3 * n, s);
    }
  }

  // out = J^{-1} (scale rhs) to rtol (J, P from the last rebuild).
  int solve(const double* rhs, double scale, double rtol, double* out) {
    NOSH_HIP(hipMemsetAsync(out, 0, sizeof(double) * L, s));
    double rr = 0.0;
    if (no.precision == 1) {
      JacView<float> J{nls->keo->kv.val32.p, nls->jblk32.p};
      PrecView<float> P{nls->keo->kv.val32.p, nls->pblk32.p, nls->bound.p, 0.0, no.cheb_ratio,
                        no.cheb_iters};
      return minres_solve<float>(m, J, P, nls->w32, rhs, scale, std::max(rtol, 1e-6),
                                 no.minres_max_iters, out, s, &rr, 8);
    }
    JacView<double> J{nls->keo->kv.val.p, nls->jblk.p};
    PrecView<double> P{nls->keo->kv.val.p, nls->pblk.p, nls->bound.p, 0.0, no.cheb_ratio,
                       no.cheb_iters};
    return minres_solve<double>(m, J, P, nls->w64, rhs, scale, rtol, no.minres_max_iters, out, s,
                                &rr, 8);
  }
};

}  // namespace

extern "C" {

int nosh_arclength_default_options(nosh_arclength_options* o) {
  return aguard([&] {
    NOSH_CHECK(o, kInvalid, "null options");
    o->parameter = "mu";
    o->start = 0.0;
    o->min_value = -100.0;
    o->max_value = 100.0;
    o->max_steps = 100;
    o->direction = 1;
    o->initial_step = 1e-2;
    o->min_step = 1e-7;
    o->max_step = 1e-1;
    o->aggressiveness = 2.0;
    o->predictor = 1;
    o->border_rtol = 1e-10;
    o->eigenvalues = 0;
    nosh_newton_default_options(&o->newton);
  });
}

int nosh_arclength_run(nosh_nls_t nls, double* psi, int np, const char* const* names,
                       const double* values, const nosh_arclength_options* opts,
                       nosh_arclength_step* record, double* states, int* steps_done) {
  int done = 0;
  int rc = aguard([&] {
    NOSH_CHECK(nls && psi && opts && record, kInvalid, "null argument");
    const nosh_arclength_options& o = *opts;
    NOSH_CHECK(o.max_steps >= 1, kInvalid, "max_steps must be >= 1");
    NOSH_CHECK(o.initial_step > 0 && o.min_step > 0 && o.max_step >= o.min_step, kInvalid,
               "bad step sizes");
    NOSH_CHECK(o.min_value < o.max_value, kInvalid, "min_value must be below max_value");
    NOSH_CHECK(o.predictor == 1 || o.predictor == 2, kInvalid, "predictor must be 1 or 2");
    NOSH_CHECK(o.eigenvalues >= 0 && o.eigenvalues <= 8, kInvalid, "eigenvalues must be 0..8");
    NOSH_CHECK(o.newton.max_iters >= 2, kInvalid, "newton.max_iters must be >= 2");
    const std::string param = o.parameter ? o.parameter : "mu";
    NOSH_CHECK(param == "mu" || param == "g", kInvalid, "continuation parameter must be mu or g");
    hipStream_t s = default_stream();
    Arc A(nls, s);
    A.param = param;
    for (int i = 0; i < np; ++i) A.p[names[i]] = values[i];
    A.no = o.newton;
    A.border_rtol = o.border_rtol;
    A.dk = std::make_unique<nosh_keo_s>();
    A.dk->mesh = nls->mesh;
    A.dk->thickness = nls->thickness;
    A.dk->mvp = nls->mvp;
    A.dk->dkeo = true;
    A.dk->dparam = param;
    const int64_t n = A.n, L = A.L;
    for (DBuf<double>* b : {&A.x0, &A.tp, &A.tprev, &A.xprev, &A.a, &A.b, &A.fl}) b->alloc(L);
    if (nls->x.n != size_t(L)) nls->x.alloc(L);
    double* x = nls->x.p;
    nls->st.ensure(L);
    nls->st.a.upload(psi, L, s);
    launch_permute_in(n, A.m.perm.p, nls->st.a.p, x, s);

    nosh_eigen_options eo;
    nosh_eigen_default_options(&eo);
    eo.nev = std::max(1, o.eigenvalues);
    eo.block_size = std::min(16, eo.nev + 2);
    eo.cheb_iters = o.newton.cheb_iters;
    eo.cheb_ratio = o.newton.cheb_ratio;
    auto finish_record = [&](nosh_arclength_step& r, double lam, int k,
                             std::chrono::steady_clock::time_point t0) {
      A.p[param] = lam;
      DBuf<double> part(2 * grid_for(n, 256)), sc(1);
      launch_cv_weighted(A.m, x, nullptr, part.p, sc.p, s);
      NOSH_HIP(hipMemcpyAsync(&r.gibbs, sc.p, sizeof(double), hipMemcpyDeviceToHost, s));
      NOSH_HIP(hipStreamSynchronize(s));
      if (o.eigenvalues > 0) {
        double ev[16];
        nosh_eigen_stats es{};
        lobpcg_jacobian(nls, x, A.p, eo, ev, nullptr, &es, s);
        r.num_eigs = o.eigenvalues;
        for (int j = 0; j < o.eigenvalues; ++j) {
          r.eigs[j] = ev[j];
          if (ev[j] < -1e-8 * es.scale) ++r.num_negative;
        }
      }
      if (states) {
        launch_permute_out(n, A.m.perm.p, x, nls->st.d.p, s);
        nls->st.d.download(states + int64_t(k) * L, L, s);
        NOSH_HIP(hipStreamSynchronize(s));
      }
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };

    // Entry 0: Newton at p = start.
    double lam = o.start;
    {
      const auto t0 = std::chrono::steady_clock::now();
      nosh_newton_stats st{};
      A.p[param] = lam;
      newton_run(nls, x, A.p, o.newton, &st, s);
      nosh_arclength_step& r = record[0];
      r = nosh_arclength_step{};
      r.param = lam;
      r.newton_iters = st.newton_iters;
      r.minres_iters = st.minres_iters_total;
      r.converged = st.converged;
      r.norm_f = st.final_norm_f;
      NOSH_CHECK(st.converged, kNotConverged, "arclength: initial Newton solve failed");
      finish_record(r, lam, 0, t0);
      done = 1;
    }

    double theta2 = -1.0, ldot = 0.0, ds = o.initial_step;
    const int N = o.newton.max_iters;
    // Unit tangent (A.tp, ldot) at the current point (x, lam), oriented along
    // the previous one (A.tprev, ldot_prev) or, at the start, by direction.
    auto tangent = [&](bool first, double ldot_prev) -> int {
      A.rebuild(x, lam);
      const int its = A.solve(A.fl.p, -1.0, o.border_rtol, A.tp.p);
      const double zz = A.dot(A.tp.p, A.tp.p);
      if (theta2 < 0) theta2 = zz > 0 ? 1.0 / zz : 1.0;
      const double nrm = std::sqrt(theta2 * zz + 1.0);
      k_arc_scale<<<A.g2, 256, 0, s>>>(L, 1.0 / nrm, A.tp.p);
      ldot = 1.0 / nrm;
      const double orient =
          first ? ldot * o.direction : theta2 * A.dot(A.tp.p, A.tprev.p) + ldot * ldot_prev;
      if (orient < 0) {
        k_arc_scale<<<A.g2, 256, 0, s>>>(L, -1.0, A.tp.p);
        ldot = -ldot;
      }
      return its;
    };
    record[0].minres_iters += tangent(true, 0.0);
    record[0].param_dot = ldot;
    for (int k = 1; k <= o.max_steps; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      nosh_arclength_step& r = record[k];
      r = nosh_arclength_step{};
      NOSH_HIP(hipMemcpyAsync(A.x0.p, x, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
      const double lam0 = lam;
      const double c0 = A.dot(A.tp.p, A.x0.p);
      bool ok = false, at_bound = false;
      int failures = 0, kiters = 0;
      double nf = 0.0;
      while (!ok) {
        const double lp = lam0 + ds * ldot;
        if (lp > o.max_value || lp < o.min_value) {
          // Natural-parameter step onto the bound from the tangent guess.
          const double target = lp > o.max_value ? o.max_value : o.min_value;
          const double h = ldot != 0.0 ? (target - lam0) / ldot : 0.0;
          k_arc_axpy<<<A.g2, 256, 0, s>>>(L, A.x0.p, h, A.tp.p, x);
          nosh_newton_stats st{};
          A.p[param] = target;
          newton_run(nls, x, A.p, o.newton, &st, s);
          r.newton_iters += st.newton_iters;
          r.minres_iters += st.minres_iters_total;
          kiters = st.newton_iters;
          nf = st.final_norm_f;
          if (st.converged) {
            ok = true;
            at_bound = true;
            lam = target;
            break;
          }
        } else {
          k_arc_axpy<<<A.g2, 256, 0, s>>>(L, A.x0.p, ds, A.tp.p, x);
          lam = lp;
          double nf_prev = -1.0, eta_prev = o.newton.forcing_max;
          for (int it = 0;; ++it) {
            nf = A.residual(x, lam);
            if (!(nf > o.newton.tol) && std::isfinite(nf)) {
              ok = true;
              kiters = it;
              break;
            }
            if (it >= N || !std::isfinite(nf)) break;
            double eta = o.newton.minres_rtol;
            if (o.newton.forcing == 1) {
              if (nf_prev < 0.0) {
                eta = o.newton.forcing_max;
              } else {
                eta = 0.9 * (nf / nf_prev) * (nf / nf_prev);
                const double safe = 0.9 * eta_prev * eta_prev;
                if (safe > 0.1) eta = std::max(eta, safe);
              }
              eta = std::min(eta, o.newton.forcing_max);
              eta = std::max(eta, 0.5 * o.newton.tol / nf);
              eta = std::max(eta, o.newton.forcing_min);
            }
            // nls->f holds F(x, lam) (residual_norm).
            A.rebuild(x, lam);
            r.minres_iters += A.solve(nls->f.p, -1.0, eta, A.a.p);
            // b to border_rtol whatever the forcing: near a fold (J nearly
            // singular) an inexact b lets the corrector converge to another
            // intersection of the arclength hyperplane with the branch.
            r.minres_iters += A.solve(A.fl.p, -1.0, o.border_rtol, A.b.p);
            const double* cols[3] = {A.a.p, A.b.p, x};
            double d[3];
            const double* tpp = A.tp.p;
            block_dots(&tpp, 1, cols, 3, L, A.bw, d, s);
            const double Nres = theta2 * (d[2] - c0) + ldot * (lam - lam0) - ds;
            const double den = ldot + theta2 * d[1];
            NOSH_CHECK(den != 0.0 && std::isfinite(den), kNotConverged,
                       "arclength: singular bordered system");
            const double dl = (-Nres - theta2 * d[0]) / den;
            k_arc_correct<<<A.g2, 256, 0, s>>>(L, A.a.p, dl, A.b.p, x);
            lam += dl;
            ++r.newton_iters;
            nf_prev = nf;
            eta_prev = eta;
          }
        }
        if (ok) break;
        ++failures;
        ds *= 0.5;
        NOSH_HIP(hipMemcpyAsync(x, A.x0.p, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
        lam = lam0;
        if (ds < o.min_step) break;
      }
      r.step = ds;
      r.param_dot = ldot;
      r.failures = failures;
      r.converged = ok ? 1 : 0;
      r.norm_f = nf;
      r.param = lam;
      if (!ok) throw nosh::Error(kNotConverged, "arclength: step size fell below min_step");
      finish_record(r, lam, k, t0);
      done = k + 1;
      if (std::getenv("NOSH_PROGRESS"))
        std::fprintf(stderr, "[arclength] step %d %s %.6f ds %.3g p' %+.3f newton %d minres %d %.3fs\n",
                     k, param.c_str(), lam, ds, ldot, r.newton_iters, r.minres_iters, r.seconds);
      if (at_bound) break;
      // Tangent at the accepted point: the next step's predictor.
      NOSH_HIP(hipMemcpyAsync(A.tprev.p, A.tp.p, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
      const double ldot_prev = ldot;
      if (o.predictor == 1) {
        r.minres_iters += tangent(false, ldot_prev);
      } else {
        // Secant through the last two accepted points (oriented by the motion).
        k_arc_diff<<<A.g2, 256, 0, s>>>(L, 1.0, x, A.x0.p, A.tp.p);
        const double zz = A.dot(A.tp.p, A.tp.p);
        const double dl = lam - lam0;
        const double nrm = std::sqrt(theta2 * zz + dl * dl);
        NOSH_CHECK(nrm > 0, kInvalid, "arclength: zero secant");
        k_arc_scale<<<A.g2, 256, 0, s>>>(L, 1.0 / nrm, A.tp.p);
        ldot = dl / nrm;
      }
      r.param_dot = ldot;
      r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (failures == 0) {
        const double q = double(N - kiters) / double(N - 1);
        ds = std::min(o.max_step, ds * (1.0 + o.aggressiveness * q * q));
      }
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
