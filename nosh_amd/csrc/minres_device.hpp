// minres_device.hpp -- device-side pieces shared by the single-instance and
// batched MINRES kernels: Chebyshev coefficients, 2x2 block inverse and the
// Paige-Saunders scalar recurrences on a MinresScalars record.
#pragma once

#include "solver.hpp"

namespace nosh {

// Chebyshev coefficients of step j >= 1 on [hi/ratio, hi].
struct Cheb {
  double theta, delta, c1, c2;
};
__device__ __forceinline__ Cheb cheb_coef(double hi, double ratio, int j) {
  const double lo = hi / ratio;
  Cheb c;
  c.theta = 0.5 * (hi + lo);
  c.delta = 0.5 * (hi - lo);
  const double sigma = c.theta / c.delta;
  double rho = 1.0 / sigma, rho_prev = rho;
  for (int q = 1; q <= j; ++q) {
    rho_prev = rho;
    rho = 1.0 / (2.0 * sigma - rho);
  }
  c.c1 = rho * rho_prev;
  c.c2 = 2.0 * rho / c.delta;
  return c;
}

__device__ __forceinline__ double cheb_hi(const double* bound_ratio, double upper) {
  return upper > 0.0 ? upper : 1.0 + *bound_ratio;
}

template <class T>
__device__ __forceinline__ void dinv(T a, T c, T b, T x0, T x1, T* o0, T* o1) {
  const T det = a * c - b * b;
  *o0 = (c * x0 - b * x1) / det;
  *o1 = (a * x1 - b * x0) / det;
}

__device__ inline void givens_update(MinresScalars* S, double beta2) {
  const double oldb = S->beta;
  const double beta = sqrt(fmax(beta2, 0.0));
  const double oldeps = S->epsln;
  const double delta = S->cs * S->dbar + S->sn * S->alfa;
  const double gbar = S->sn * S->dbar - S->cs * S->alfa;
  S->epsln = S->sn * beta;
  S->dbar = -S->cs * beta;
  const double gamma = fmax(hypot(gbar, beta), 1e-300);
  S->cs = gbar / gamma;
  S->sn = beta / gamma;
  const double phi = S->cs * S->phibar;
  S->phibar = S->sn * S->phibar;
  S->p_oldeps = oldeps;
  S->p_delta = delta;
  S->p_denom = 1.0 / gamma;
  S->p_phi = phi;
  S->pending = 1;
  S->oldb = oldb;
  S->beta = beta;
  S->it += 1;
  S->relres = S->phibar / S->beta1;
  if (S->relres <= S->rtol || beta == 0.0 || S->it >= S->maxit) S->done = 1;
}

__device__ inline void init_update(MinresScalars* S, double beta2) {
  S->it = 0;
  S->pending = 0;
  S->relres = 1.0;
  S->alfa = 0.0;
  if (!(beta2 > 0.0)) {
    S->beta1 = 0.0;
    S->beta = 0.0;
    S->relres = 0.0;
    S->done = 1;
    return;
  }
  const double b1 = sqrt(beta2);
  S->beta1 = b1;
  S->beta = b1;
  S->oldb = 0.0;
  S->dbar = 0.0;
  S->epsln = 0.0;
  S->phibar = b1;
  S->cs = -1.0;
  S->sn = 0.0;
  S->done = 0;
}

}  // namespace nosh
