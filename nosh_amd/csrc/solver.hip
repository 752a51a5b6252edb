// solver.hip -- the preconditioner (Chebyshev iteration with 2x2 block-Jacobi
// scaling, standing in for the MueLu AMG cycle of keo_regularized::apply,
// src/keo_regularized.cpp:88-165) and preconditioned MINRES (Paige-Saunders),
// the in-repo replacement of the Belos solver (the reference example selects
// Pseudo Block GMRES at examples/conf.xml:104 with its MINRES line :105
// commented out; J is symmetric and the preconditioner SPD, so MINRES solves
// the same systems with a short recurrence, tolerance and limit as :117-123).
//
// MINRES runs as a fixed sequence of fused kernels per iteration with every
// scalar of the Lanczos/Givens recurrences kept on the device:
//   A  one Jacobian product v = y/beta, u = J v - (beta/beta_old) r1, the
//      partial sums of alpha = <v,u>, and the previous iteration's update of
//      the direction vectors w and the solution x;
//   B  r_new = u - (alpha/beta) r2 and the first Chebyshev step on r_new;
//   C  the remaining Chebyshev steps, each one product with P; the last one
//      also sums beta^2 = <r_new, z> and advances the Givens rotation.
// Reductions are deterministic: partial sums per block, then a one-block
// kernel adds them in block order and evaluates the scalar recurrences. The host only
// polls the `done` flag every few iterations, so iterations issued after
// convergence are no-ops. Buffer rotations (r1/r2, w) are indexed by the
// device iteration counter, so every iteration launches identical kernels.
// Vectors are FP64 or FP32 (template T); inner products, norms, recurrence
// scalars and the solution accumulator are always FP64.

#include "kernels.hpp"
#include "minres_device.hpp"
#include "solver.hpp"

namespace nosh {
namespace {

constexpr int kBlock = 256;

// mode: 0 standalone preconditioner, 1 MINRES initialisation, 2 MINRES iteration.
template <class T>
struct ChebArgs {
  int64_t n;
  SellView sv;  // split off-diagonal layout
  const V2<T>* pval;   // P off-diagonal values
  const T* pblk;       // P diagonal blocks (3 n)
  const double* bound; // max ratio from the rebuild
  double upper, ratio;
  int iters;           // Chebyshev iterations (0 = identity)
  V2<T>* R0;           // r buffers (mode 0/1: the input, R0 == R1)
  V2<T>* R1;
  V2<T>* D;            // Chebyshev direction
  const V2<T>* u;      // mode 2: J v - (beta/oldb) r1
  MinresScalars* S;
  double* partials;
};

// Chebyshev step 0 (z = D^{-1} r / theta), producing r first in mode 2.
template <class T>
__global__ __launch_bounds__(kBlock) void k_cheb_first(ChebArgs<T> a, int mode, V2<T>* zout) {
  MinresScalars* S = a.S;
  if (mode != 0 && S->done) return;
  const int64_t row = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int it = mode == 2 ? S->it : 0;
  V2<T>* r = (mode == 2) ? ((it & 1) ? a.R1 : a.R0) : a.R0;
  double s2[1] = {0.0};
  if (row < a.n) {
    V2<T> rv;
    if (mode == 2) {
      const V2<T>* r2 = (it & 1) ? a.R0 : a.R1;
      const T c = T(S->alfa / S->beta);
      const V2<T> uu = a.u[row], r2v = r2[row];
      rv.x = uu.x - c * r2v.x;
      rv.y = uu.y - c * r2v.y;
      r[row] = rv;
    } else {
      rv = r[row];
    }
    V2<T> z;
    if (a.iters == 0) {
      z = rv;
    } else {
      const double hi = cheb_hi(a.bound, a.upper);
      const T th = T(0.5 * (hi + hi / a.ratio));
      T o0, o1;
      dinv<T>(a.pblk[row], a.pblk[a.n + row], a.pblk[2 * a.n + row], rv.x, rv.y, &o0, &o1);
      z.x = o0 / th;
      z.y = o1 / th;
      if (a.iters > 1) a.D[row] = z;
    }
    zout[row] = z;
    s2[0] = double(rv.x) * double(z.x) + double(rv.y) * double(z.y);
  }
  if (mode == 0 || a.iters > 1) return;
  block_partials<1>(s2, a.partials);
}

// Chebyshev step j >= 1: z_out = z_in + d, d = c1 d + c2 D^{-1}(r - P z_in).
template <class T>
__global__ __launch_bounds__(kBlock) void k_cheb_step(ChebArgs<T> a, int mode, int j,
                                                     const V2<T>* __restrict__ zin,
                                                     V2<T>* __restrict__ zout) {
  MinresScalars* S = a.S;
  if (mode != 0 && S->done) return;
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  const int it = mode == 2 ? S->it : 0;
  const V2<T>* r = (mode == 2) ? ((it & 1) ? a.R1 : a.R0) : a.R0;
  const bool last = (j == a.iters - 1);
  double s2[1] = {0.0};
  if (row < a.n) {
    const Cheb c = cheb_coef(cheb_hi(a.bound, a.upper), a.ratio, j);
    const V2<T> acc = sell_row<T>(row, a.sv, a.pval, zin);
    const V2<T> zi = zin[row];
    const T b0 = a.pblk[row], b1 = a.pblk[a.n + row], b2 = a.pblk[2 * a.n + row];
    const V2<T> rv = r[row];
    const T res0 = rv.x - (acc.x + b0 * zi.x + b2 * zi.y);
    const T res1 = rv.y - (acc.y + b2 * zi.x + b1 * zi.y);
    T u0, u1;
    dinv<T>(b0, b1, b2, res0, res1, &u0, &u1);
    V2<T> d = a.D[row];
    d.x = T(c.c1) * d.x + T(c.c2) * u0;
    d.y = T(c.c1) * d.y + T(c.c2) * u1;
    V2<T> z;
    z.x = zi.x + d.x;
    z.y = zi.y + d.y;
    if (!last) a.D[row] = d;
    zout[row] = z;
    s2[0] = double(rv.x) * double(z.x) + double(rv.y) * double(z.y);
  }
  if (mode == 0 || !last) return;
  block_partials<1>(s2, a.partials);
}

// beta^2 = sum of the partials of <r, z>: MINRES initialisation (mode 1) or
// the Givens step of the current iteration (mode 2).
__global__ void k_fin_beta(MinresScalars* S, const double* __restrict__ partials, int nb,
                           int mode) {
  if (mode == 2 && S->done) return;
  double v[1];
  finalize_sum<1>(v, partials, nb);
  if (threadIdx.x == 0) {
    if (mode == 1)
      init_update(S, v[0]);
    else
      givens_update(S, v[0]);
  }
}

__global__ void k_fin_alfa(MinresScalars* S, const double* __restrict__ partials, int nb) {
  if (S->done) return;
  double v[1];
  finalize_sum<1>(v, partials, nb);
  if (threadIdx.x == 0) {
    S->alfa = v[0];
    S->pending = 0;
  }
}

template <class T>
struct MinresArgs {
  int64_t n;
  SellView sv;  // split off-diagonal layout
  const V2<T>* jval;
  const T* jblk;
  const V2<T>* y;  // preconditioned residual of the previous iteration
  V2<T>* v;
  V2<T>* u;
  const V2<T>* R0;
  const V2<T>* R1;
  V2<T>* W0;
  V2<T>* W1;
  V2<T>* W2;
  double2* x;
  MinresScalars* S;
  double* partials;
};

template <class T>
__device__ __forceinline__ V2<T>* wbuf(const MinresArgs<T>& a, int k) {
  const int q = ((k % 3) + 3) % 3;
  return q == 0 ? a.W0 : (q == 1 ? a.W1 : a.W2);
}

// Pending update of iteration k: w_k = (v_k - oldeps w_{k-2} - delta w_{k-1}) / gamma,
// x += phi w_k.
template <class T>
__device__ __forceinline__ void w_update(const MinresArgs<T>& a, const MinresScalars* S, int k,
                                         int64_t row, V2<T> vk) {
  V2<T>* wk = wbuf(a, k);
  const V2<T> w1 = wbuf(a, k - 2)[row], w2 = wbuf(a, k - 1)[row];
  const T oe = T(S->p_oldeps), de = T(S->p_delta), dn = T(S->p_denom);
  V2<T> w;
  w.x = (vk.x - oe * w1.x - de * w2.x) * dn;
  w.y = (vk.y - oe * w1.y - de * w2.y) * dn;
  wk[row] = w;
  const double phi = S->p_phi;
  double2 xv = a.x[row];
  xv.x += phi * double(w.x);
  xv.y += phi * double(w.y);
  a.x[row] = xv;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_minres_a(MinresArgs<T> a) {
  MinresScalars* S = a.S;
  if (S->done) return;
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  const int it = S->it;
  const double beta = S->beta;
  double s[1] = {0.0};
  if (row < a.n) {
    const V2<T> vold = a.v[row];
    if (S->pending) w_update<T>(a, S, it, row, vold);
    const V2<T> acc = sell_row<T>(row, a.sv, a.jval, a.y);
    const V2<T> yi = a.y[row];
    const T b0 = a.jblk[row], b1 = a.jblk[a.n + row], b2 = a.jblk[2 * a.n + row];
    const T ib = T(1.0 / beta);
    V2<T> uu;
    uu.x = (acc.x + b0 * yi.x + b2 * yi.y) * ib;
    uu.y = (acc.y + b2 * yi.x + b1 * yi.y) * ib;
    if (it >= 1) {
      const V2<T>* r1 = (it & 1) ? a.R1 : a.R0;  // r1 of iteration it+1: R[it % 2]
      const T c = T(beta / S->oldb);
      const V2<T> rv = r1[row];
      uu.x -= c * rv.x;
      uu.y -= c * rv.y;
    }
    V2<T> vn;
    vn.x = yi.x * ib;
    vn.y = yi.y * ib;
    a.v[row] = vn;
    a.u[row] = uu;
    s[0] = double(vn.x) * double(uu.x) + double(vn.y) * double(uu.y);
  }
  block_partials<1>(s, a.partials);
}

// Applies the last pending w/x update, then psi += x when psi != nullptr.
template <class T>
__global__ __launch_bounds__(kBlock) void k_minres_flush(MinresArgs<T> a, double2* psi) {
  const MinresScalars* S = a.S;
  const int64_t row = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (row >= a.n) return;
  if (S->pending) w_update<T>(a, S, S->it, row, a.v[row]);
  if (psi) {
    const double2 d = a.x[row];
    double2 p = psi[row];
    p.x += d.x;
    p.y += d.y;
    psi[row] = p;
  }
}

// r0 = r1 = scale * b (converted to T), x = 0, w_{-1} = w_0 = 0.
template <class T>
__global__ void k_minres_setup(int64_t n, const double2* __restrict__ b, double scale,
                               V2<T>* R0, V2<T>* R1, V2<T>* W0, V2<T>* W2, double2* x) {
  const int64_t row = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (row >= n) return;
  const double2 bv = b[row];
  V2<T> r;
  r.x = T(scale * bv.x);
  r.y = T(scale * bv.y);
  R0[row] = r;
  R1[row] = r;
  V2<T> z;
  z.x = T(0);
  z.y = T(0);
  W0[row] = z;
  W2[row] = z;
  x[row] = make_double2(0.0, 0.0);
}

__global__ void k_scalars_reset(MinresScalars* S, double rtol, int maxit) {
  S->done = 0;
  S->it = 0;
  S->pending = 0;
  S->rtol = rtol;
  S->maxit = maxit;
  S->relres = 1.0;
}

template <class T>
__global__ void k_to_t(int64_t m, const double* __restrict__ a, T* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < m) b[i] = T(a[i]);
}

}  // namespace

template <class T>
void MinresWork<T>::ensure(int64_t n_) {
  if (n == n_) return;
  n = n_;
  for (auto* b : {&Y, &Z2, &v, &u, &R0, &R1, &D, &W0, &W1, &W2}) b->alloc(n);
  x.alloc(n);
  const int64_t nb = grid_for(n, kBlock);
  partials.alloc(4 * nb + 64);
  if (!S.p) S.alloc(1);
  if (!Sh) NOSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&Sh), sizeof(MinresScalars)));
  NOSH_HIP(hipStreamSynchronize(default_stream()));
}

template <class T>
MinresWork<T>::~MinresWork() {
  if (Sh) (void)hipHostFree(Sh);
}

template <class T>
static ChebArgs<T> cheb_args(const MeshCache& m, const PrecView<T>& P, MinresWork<T>& w,
                             V2<T>* R0, V2<T>* R1) {
  ChebArgs<T> a;
  a.n = m.n;
  a.sv = sell_view(m);
  a.pval = P.val;
  a.pblk = P.blk;
  a.bound = P.bound;
  a.upper = P.upper;
  a.ratio = P.ratio;
  a.iters = P.iters;
  a.R0 = R0;
  a.R1 = R1;
  a.D = w.D.p;
  a.u = w.u.p;
  a.S = w.S.p;
  a.partials = w.partials.p;
  return a;
}

// Launches the Chebyshev chain so that the final z lands in `out`.
template <class T>
static void cheb_chain(const MeshCache& m, ChebArgs<T> a, int mode, V2<T>* out, V2<T>* tmp,
                       hipStream_t s) {
  const unsigned nb = grid_for(m.n, kBlock);
  const int steps = a.iters < 1 ? 1 : a.iters;
  V2<T>* zb[2] = {out, tmp};
  int cur = (steps - 1) % 2;
  k_cheb_first<T><<<nb, kBlock, 0, s>>>(a, mode, zb[cur]);
  for (int j = 1; j < steps; ++j) {
    k_cheb_step<T><<<nb, kBlock, 0, s>>>(a, mode, j, zb[cur], zb[1 - cur]);
    cur = 1 - cur;
  }
  if (mode != 0) k_fin_beta<<<1, kFinalizeThreads, 0, s>>>(a.S, a.partials, int(nb), mode);
}

template <class T>
void prec_apply(const MeshCache& m, const PrecView<T>& P, MinresWork<T>& w, const V2<T>* r,
                V2<T>* z, hipStream_t s) {
  ChebArgs<T> a = cheb_args(m, P, w, const_cast<V2<T>*>(r), const_cast<V2<T>*>(r));
  cheb_chain<T>(m, a, 0, z, w.Z2.p, s);
}

template <class T>
int minres_solve(const MeshCache& m, const JacView<T>& J, const PrecView<T>& P,
                 MinresWork<T>& w, const double* b, double b_scale, double rtol, int maxit,
                 double* psi_update, hipStream_t s, double* relres_out, int poll) {
  w.ensure(m.n);
  const unsigned nb = grid_for(m.n, kBlock);
  k_scalars_reset<<<1, 1, 0, s>>>(w.S.p, rtol, maxit);
  k_minres_setup<T><<<nb, kBlock, 0, s>>>(m.n, reinterpret_cast<const double2*>(b), b_scale,
                                          w.R0.p, w.R1.p, w.W0.p, w.W2.p, w.x.p);
  ChebArgs<T> ca = cheb_args(m, P, w, w.R0.p, w.R1.p);
  // y = M^{-1} b and beta1.
  cheb_chain<T>(m, ca, 1, w.Y.p, w.Z2.p, s);
  MinresArgs<T> ma;
  ma.n = m.n;
  ma.sv = sell_view(m);
  ma.jval = J.val;
  ma.jblk = J.blk;
  ma.y = w.Y.p;
  ma.v = w.v.p;
  ma.u = w.u.p;
  ma.R0 = w.R0.p;
  ma.R1 = w.R1.p;
  ma.W0 = w.W0.p;
  ma.W1 = w.W1.p;
  ma.W2 = w.W2.p;
  ma.x = w.x.p;
  ma.S = w.S.p;
  ma.partials = w.partials.p;
  int issued = 0;
  for (;;) {
    for (int q = 0; q < poll && issued < maxit; ++q, ++issued) {
      k_minres_a<T><<<nb, kBlock, 0, s>>>(ma);
      k_fin_alfa<<<1, kFinalizeThreads, 0, s>>>(w.S.p, w.partials.p, int(nb));
      cheb_chain<T>(m, ca, 2, w.Y.p, w.Z2.p, s);
    }
    NOSH_HIP(hipMemcpyAsync(w.Sh, w.S.p, sizeof(MinresScalars), hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipStreamSynchronize(s));
    if (w.Sh->done || issued >= maxit) break;
  }
  k_minres_flush<T><<<nb, kBlock, 0, s>>>(ma, reinterpret_cast<double2*>(psi_update));
  NOSH_HIP(hipMemcpyAsync(w.Sh, w.S.p, sizeof(MinresScalars), hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  if (relres_out) *relres_out = w.Sh->relres;
  return w.Sh->it;
}

template <class T>
void convert_to(const double* a, T* b, int64_t m, hipStream_t s) {
  k_to_t<T><<<grid_for(m, kBlock), kBlock, 0, s>>>(m, a, b);
}

template struct MinresWork<double>;
template struct MinresWork<float>;
template void prec_apply<double>(const MeshCache&, const PrecView<double>&, MinresWork<double>&,
                                 const double2*, double2*, hipStream_t);
template void prec_apply<float>(const MeshCache&, const PrecView<float>&, MinresWork<float>&,
                                const float2*, float2*, hipStream_t);
template int minres_solve<double>(const MeshCache&, const JacView<double>&,
                                  const PrecView<double>&, MinresWork<double>&, const double*,
                                  double, double, int, double*, hipStream_t, double*, int);
template int minres_solve<float>(const MeshCache&, const JacView<float>&, const PrecView<float>&,
                                 MinresWork<float>&, const double*, double, double, int, double*,
                                 hipStream_t, double*, int);
template void convert_to<float>(const double*, float*, int64_t, hipStream_t);

}  // namespace nosh
