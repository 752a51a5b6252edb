// solver.hpp -- MINRES / Chebyshev interfaces (solver.hip).
#pragma once

#include "kernels.hpp"

namespace nosh {

// Device-resident MINRES scalars (Paige-Saunders recurrences).
struct MinresScalars {
  double beta1, beta, oldb, alfa, dbar, epsln, phibar, cs, sn;
  double p_oldeps, p_delta, p_denom, p_phi;  // pending w/x update
  // Batched solver: phi of the last applied w update whose x update was
  // deferred to the next one (0: none).
  double x_phi;
  double relres, rtol;
  int it, maxit, done, pending;
};

// Work vectors of one MINRES solve (precision T) and its scalars.
template <class T>
struct MinresWork {
  int64_t n = -1;
  DBuf<V2<T>> Y, Z2, v, u, R0, R1, D, W0, W1, W2;
  DBuf<double2> x;
  DBuf<double> partials;
  DBuf<MinresScalars> S;
  MinresScalars* Sh = nullptr;  // pinned host mirror
  void ensure(int64_t n);
  ~MinresWork();
};

// Operator views in the internal vertex order.
template <class T>
struct JacView {
  const V2<T>* val;  // off-diagonal values (K's)
  const T* blk;      // 3 n: Jacobian 2x2 blocks including diag(K)
};

template <class T>
struct PrecView {
  const V2<T>* val;
  const T* blk;         // 3 n: keo_regularized 2x2 blocks including diag(K)
  const double* bound;  // device: max_k sum|alpha| / lambda_min(P_k)
  double upper;         // > 0: fixed upper spectral bound
  double ratio;
  int iters;            // Chebyshev iterations, 0 = identity
};

template <class T>
void prec_apply(const MeshCache& m, const PrecView<T>& P, MinresWork<T>& w, const V2<T>* r,
                V2<T>* z, hipStream_t s);

// Solves J x = b_scale * b (x from 0) to rtol in the M^{-1} norm. b, x and
// psi_update are FP64 (internal order); when psi_update != nullptr the
// solution is added to it. Returns the iteration count.
template <class T>
int minres_solve(const MeshCache& m, const JacView<T>& J, const PrecView<T>& P,
                 MinresWork<T>& w, const double* b, double b_scale, double rtol, int maxit,
                 double* psi_update, hipStream_t s, double* relres_out, int poll);

template <class T>
void convert_to(const double* a, T* b, int64_t m, hipStream_t s);

}  // namespace nosh
