// operators.hip -- KEO assembly and the fused operator applies of the NLS
// model: K x, F(psi), dF/dp, the Jacobian and keo_regularized blocks.
//
// Every operator is "off-diagonal sliced-ELLPACK product + per-vertex 2x2
// block", one kernel each, over interleaved complex vectors in the internal
// (RCM) vertex order. References (nosh tree):
//   keo::refill_               src/parameter_matrix_keo.cpp:74-184
//   keo::build_alpha_cache_    src/parameter_matrix_keo.cpp:186-232
//   DkeoDP::refill_            src/parameter_matrix_dkeo_dp.cpp:60-155
//   explicit_values cache      src/vector_field_explicit_values.cpp:13-50
//   nls::compute_f_            src/model_evaluator_nls.cpp:527-628
//   nls::computeDFDP_          src/model_evaluator_nls.cpp:630-695
//   jacobian_operator          src/jacobian_operator.cpp:38-200
//   keo_regularized::rebuild   src/keo_regularized.cpp:181-264

#include <hip/hip_ext.h>

#include "batched.hpp"
#include "kernels.hpp"

namespace nosh {
namespace {

constexpr int kBlock = 256;

__global__ void k_atilde(const double* __restrict__ X, const int32_t* __restrict__ edges,
                         const double* __restrict__ A, int64_t E, double* __restrict__ at) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int64_t a = edges[2 * e], b = edges[2 * e + 1];
  const double mx = 0.5 * (A[3 * a] + A[3 * b]);
  const double my = 0.5 * (A[3 * a + 1] + A[3 * b + 1]);
  const double mz = 0.5 * (A[3 * a + 2] + A[3 * b + 2]);
  at[e] = mx * (X[3 * a] - X[3 * b]) + my * (X[3 * a + 1] - X[3 * b + 1]) +
          mz * (X[3 * a + 2] - X[3 * b + 2]);
}

__global__ void k_alpha(const double* __restrict__ covol, const double* __restrict__ len,
                        const int32_t* __restrict__ edges, const double* __restrict__ t,
                        int64_t E, double* __restrict__ alpha) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double al = covol[e] / len[e];
  alpha[e] = al * 0.5 * (t[edges[2 * e]] + t[edges[2 * e + 1]]);
}

// Per slot: K_ij = -alpha e^{-i a} (row = lower endpoint) or -alpha e^{+i a}
// (row = higher endpoint), a = mu * atilde. dK/dmu_ij = -i a~ K_ij (lower) /
// +i a~ K_ij (higher), i.e. the 4x4 blocks of keo::refill_ and
// DkeoDP::refill_ in complex form.
__global__ void k_keo_fill(const int32_t* __restrict__ sedge, const double* __restrict__ alpha,
                           const double* __restrict__ atilde, double mu, int64_t nnz,
                           double2* __restrict__ val, double2* __restrict__ dval, int dmode) {
  // Slots 0 .. nnz-1 of the upper part, then the zero entry at nnz that the
  // lower part's padding points to.
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j > nnz) return;
  const int32_t code = j < nnz ? sedge[j] : -1;
  if (code < 0) {
    val[j] = make_double2(0.0, 0.0);
    if (dval) dval[j] = make_double2(0.0, 0.0);
    return;
  }
  const int32_t e = code >> 1;
  const bool hi = code & 1;
  const double al = alpha[e];
  const double a = mu * atilde[e];
  double sn, cs;
  sincos(a, &sn, &cs);
  // lower row: (v0, -v1) = (-al cos a, al sin a); higher row: (v0, v1).
  val[j] = make_double2(-cs * al, hi ? -sn * al : sn * al);
  if (dval) {
    // DkeoDP: v0 = dAdP sin a * al, v1 = -dAdP cos a * al.
    const double dadp = dmode == 1 ? atilde[e] : 0.0;
    const double v0 = dadp * sn * al, v1 = -dadp * cs * al;
    dval[j] = make_double2(v0, hi ? v1 : -v1);
  }
}

// Row sums of alpha (the diagonal of K) and of |alpha|, in slot order.
__global__ void k_keo_diag(const int32_t* __restrict__ slice_ptr,
                           const int32_t* __restrict__ sedge, const double* __restrict__ alpha,
                           int64_t n, double* __restrict__ diag, double* __restrict__ rowabs) {
  const int64_t row = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int64_t s = row >> 6;
  double d = 0.0, da = 0.0;
  for (int32_t j = slice_ptr[s] + int32_t(row & 63); j < slice_ptr[s + 1]; j += kSliceH) {
    const int32_t code = sedge[j];
    if (code < 0) continue;
    const double al = alpha[code >> 1];
    d += al;
    da += fabs(al);
  }
  diag[row] = d;
  rowabs[row] = da;
}

__global__ void k_to_fp32(const double2* __restrict__ a, int64_t nnz, float2* __restrict__ b) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  const double2 v = a[j];
  b[j] = make_float2(float(v.x), float(v.y));
}

__global__ void k_permute_in(int64_t n, const int32_t* __restrict__ perm,
                             const double2* __restrict__ xr, double2* __restrict__ xi) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) xi[i] = xr[perm[i]];
}

__global__ void k_permute_out(int64_t n, const int32_t* __restrict__ perm,
                              const double2* __restrict__ xi, double2* __restrict__ xr) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) xr[perm[i]] = xi[i];
}

__global__ void k_gather(int64_t n, const int32_t* __restrict__ perm,
                         const double* __restrict__ vr, double* __restrict__ vi) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) vi[i] = vr[perm[i]];
}

__global__ __launch_bounds__(kBlock) void k_keo_apply(
    int64_t n, SellView S,
    const double2* __restrict__ val, const double* __restrict__ diag,
    const double2* __restrict__ x, double2* __restrict__ y) {
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  if (row >= n) return;
  const double2 acc = sell_row<double>(row, S, val, x);
  const double2 xi = x[row];
  const double d = diag[row];
  y[row] = make_double2(acc.x + d * xi.x, acc.y + d * xi.y);
}

// f = K x + ct (V + g |x|^2) x  and  sum |f|^2 (when nrm2 != nullptr).
__global__ __launch_bounds__(kBlock) void k_compute_f(
    int64_t n, SellView S,
    const double2* __restrict__ val, const double* __restrict__ diag,
    const double* __restrict__ ct, const double* __restrict__ V, double g,
    const double2* __restrict__ x, double2* __restrict__ f, double* __restrict__ partials) {
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  double s2[1] = {0.0};
  if (row < n) {
    const double2 acc = sell_row<double>(row, S, val, x);
    const double2 xi = x[row];
    const double a = ct[row] * (V[row] + g * (xi.x * xi.x + xi.y * xi.y));
    const double d = diag[row];
    const double2 r = make_double2(acc.x + d * xi.x + a * xi.x, acc.y + d * xi.y + a * xi.y);
    f[row] = r;
    s2[0] = r.x * r.x + r.y * r.y;
  }
  if (partials) block_partials<1>(s2, partials);
}

__global__ void k_finalize_store(const double* __restrict__ partials, int nb,
                                 double* __restrict__ out) {
  double v[1];
  finalize_sum<1>(v, partials, nb);
  if (threadIdx.x == 0) *out = v[0];
}

// dF/dp column: dK x + ct |x|^2 x (parameter "g") or ct dV/dp x.
__global__ __launch_bounds__(kBlock) void k_dfdp(
    int64_t n, SellView S,
    const double2* __restrict__ dval, const double* __restrict__ ct, int is_g, double dvdp,
    const double* __restrict__ dvdp_vec, const double2* __restrict__ x, double2* __restrict__ y) {
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  if (row >= n) return;
  const double2 acc = sell_row<double>(row, S, dval, x);
  const double2 xi = x[row];
  const double dv = dvdp_vec ? dvdp + dvdp_vec[row] : dvdp;
  const double a = is_g ? ct[row] * (xi.x * xi.x + xi.y * xi.y) : ct[row] * dv;
  y[row] = make_double2(acc.x + a * xi.x, acc.y + a * xi.y);
}

// Jacobian and keo_regularized 2x2 blocks, each stored as three arrays
// (xx, yy, xy) of n values, both including the diagonal of K.
//   J: ct (V + 2 g |psi|^2) I + g ct [[Re^2-Im^2, 2 Re Im], [2 Re Im, Im^2-Re^2]]
//      (jacobian_operator::rebuild_diags_, src/jacobian_operator.cpp:184-197)
//   P: 2 g ct |psi|^2 I + the same g-part, only when g > 0
//      (keo_regularized::rebuild, src/keo_regularized.cpp:200-259)
// bound: max over vertices of sum_j |alpha_ij| / lambda_min(P_i), via an
// atomic max on the bit pattern (nonnegative doubles order like integers).
__global__ __launch_bounds__(kBlock) void k_rebuild_blocks(
    int64_t n, const double* __restrict__ diag, const double* __restrict__ rowabs,
    const double* __restrict__ ct, const double* __restrict__ V, double g,
    const double2* __restrict__ psi, double* __restrict__ jblk, double* __restrict__ pblk,
    unsigned long long* __restrict__ bound) {
  const int64_t row = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  double ratio = 0.0;
  if (row < n) {
    const double2 p = psi[row];
    const double c = ct[row];
    const double d = diag[row];
    const double r2 = p.x * p.x + p.y * p.y;
    const double al = c * (V[row] + g * 2.0 * r2);
    const double rx2 = g * c * (p.x * p.x - p.y * p.y);
    const double d1b = g * c * (2.0 * p.x * p.y);
    if (jblk) {
      jblk[row] = d + al + rx2;
      jblk[n + row] = d + al - rx2;
      jblk[2 * n + row] = d1b;
    }
    double a0 = d, a1 = d, b = 0.0;
    if (g > 0.0) {
      const double pa = g * c * 2.0 * r2;
      a0 = d + pa + rx2;
      a1 = d + pa - rx2;
      b = d1b;
    }
    if (pblk) {
      pblk[row] = a0;
      pblk[n + row] = a1;
      pblk[2 * n + row] = b;
    }
    const double lmin = 0.5 * (a0 + a1) - hypot(0.5 * (a0 - a1), b);
    ratio = lmin > 0.0 ? rowabs[row] / lmin : 1.0;
  }
  __shared__ double wmax[kBlock / 64];
  ratio = wave_max(ratio);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = ratio;
  __syncthreads();
  if (threadIdx.x == 0 && bound) {
    double m = wmax[0];
    for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, wmax[w]);
    atomicMax(bound, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_block_apply(
    int64_t n, SellView S,
    const V2<T>* __restrict__ val, const T* __restrict__ blk, const V2<T>* __restrict__ x,
    V2<T>* __restrict__ y) {
  const int64_t row = xcd_block(true) * kBlock + threadIdx.x;
  if (row >= n) return;
  const V2<T> acc = sell_row<T>(row, S, val, x);
  const V2<T> xi = x[row];
  const T b0 = blk[row], b1 = blk[n + row], b2 = blk[2 * n + row];
  V2<T> r;
  r.x = acc.x + b0 * xi.x + b2 * xi.y;
  r.y = acc.y + b2 * xi.x + b1 * xi.y;
  y[row] = r;
}

}  // namespace

void launch_atilde(const MeshCache& m, const double* A_dev, double* atilde, hipStream_t s) {
  k_atilde<<<grid_for(m.E, kBlock), kBlock, 0, s>>>(m.coords.p, m.edges.p, A_dev, m.E, atilde);
}

void launch_alpha(const MeshCache& m, const double* thickness_ref, double* alpha, hipStream_t s) {
  k_alpha<<<grid_for(m.E, kBlock), kBlock, 0, s>>>(m.covol.p, m.len.p, m.edges.p, thickness_ref,
                                                   m.E, alpha);
}

void launch_keo_fill(const MeshCache& m, const double* alpha, const double* atilde, double mu,
                     bool dkeo, bool deriv_is_mu, KeoValues& kv, hipStream_t s) {
  const size_t nv = size_t(m.nnz_u) + 1;
  if (kv.val.n != nv) kv.val.alloc(nv);
  if (dkeo && kv.dval.n != nv) kv.dval.alloc(nv);
  if (kv.diag.n != size_t(m.n)) kv.diag.alloc(m.n);
  if (kv.rowabs.n != size_t(m.n)) kv.rowabs.alloc(m.n);
  k_keo_fill<<<grid_for(int64_t(nv), kBlock), kBlock, 0, s>>>(
      m.uedge.p, alpha, atilde, mu, m.nnz_u, kv.val.p, dkeo ? kv.dval.p : nullptr,
      deriv_is_mu ? 1 : 0);
  k_keo_diag<<<grid_for(m.n, kBlock), kBlock, 0, s>>>(m.slice_ptr.p, m.sedge.p, alpha, m.n,
                                                      kv.diag.p, kv.rowabs.p);
  kv.mu = mu;
  kv.valid = true;
  kv.dvalid = dkeo;
  kv.valid32 = false;
}

void launch_keo_to_fp32(const MeshCache& m, KeoValues& kv, hipStream_t s) {
  const int64_t nv = m.nnz_u + 1;
  if (kv.val32.n != size_t(nv)) kv.val32.alloc(nv);
  k_to_fp32<<<grid_for(nv, kBlock), kBlock, 0, s>>>(kv.val.p, nv, kv.val32.p);
  kv.valid32 = true;
}

void launch_permute_in(int64_t n, const int32_t* perm, const double* x_ref, double* x_int,
                       hipStream_t s) {
  k_permute_in<<<grid_for(n, kBlock), kBlock, 0, s>>>(
      n, perm, reinterpret_cast<const double2*>(x_ref), reinterpret_cast<double2*>(x_int));
}

void launch_permute_out(int64_t n, const int32_t* perm, const double* x_int, double* x_ref,
                        hipStream_t s) {
  k_permute_out<<<grid_for(n, kBlock), kBlock, 0, s>>>(
      n, perm, reinterpret_cast<const double2*>(x_int), reinterpret_cast<double2*>(x_ref));
}

void launch_gather_ref(int64_t n, const int32_t* perm, const double* v_ref, double* v_int,
                       hipStream_t s) {
  k_gather<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, perm, v_ref, v_int);
}

void launch_keo_apply(const MeshCache& m, const KeoValues& kv, const double* x, double* y,
                      hipStream_t s) {
  k_keo_apply<<<grid_for(m.n, kBlock), kBlock, 0, s>>>(
      m.n, sell_view(m), kv.val.p, kv.diag.p, reinterpret_cast<const double2*>(x),
      reinterpret_cast<double2*>(y));
}

void launch_dfdp(const MeshCache& m, const KeoValues& kv, const double* ct, const double* x,
                 bool is_g, double dvdp, const double* dvdp_vec, double* y, hipStream_t s) {
  k_dfdp<<<grid_for(m.n, kBlock), kBlock, 0, s>>>(
      m.n, sell_view(m), kv.dval.p, ct, is_g ? 1 : 0, dvdp, dvdp_vec,
      reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y));
}

// Reduction scratch for k_compute_f (one per process; calls on the library
// stream are serialised).
namespace {
DBuf<double>& fscratch(int64_t nblocks) {
  // Never freed: device memory must not be released from static destructors
  // that may run after the HIP runtime has shut down.
  static DBuf<double>& partials = *new DBuf<double>;
  if (partials.n < size_t(nblocks)) partials.alloc(nblocks);
  return partials;
}
}  // namespace

void launch_compute_f(const MeshCache& m, const KeoValues& kv, const double* ct,
                      const double* V, double g, const double* x, double* f, double* nrm2,
                      hipStream_t s) {
  const unsigned nb = grid_for(m.n, kBlock);
  double* partials = nrm2 ? fscratch(nb).p : nullptr;
  k_compute_f<<<nb, kBlock, 0, s>>>(m.n, sell_view(m), kv.val.p, kv.diag.p, ct, V, g,
                                    reinterpret_cast<const double2*>(x),
                                    reinterpret_cast<double2*>(f), partials);
  if (nrm2) k_finalize_store<<<1, kFinalizeThreads, 0, s>>>(partials, int(nb), nrm2);
}

void launch_rebuild_blocks(const MeshCache& m, const KeoValues& kv, const double* ct,
                           const double* V, double g, const double* psi, double* jblk,
                           double* pblk, double* bound, hipStream_t s) {
  if (bound) NOSH_HIP(hipMemsetAsync(bound, 0, sizeof(double), s));
  k_rebuild_blocks<<<grid_for(m.n, kBlock), kBlock, 0, s>>>(
      m.n, kv.diag.p, kv.rowabs.p, ct, V, g, reinterpret_cast<const double2*>(psi), jblk, pblk,
      reinterpret_cast<unsigned long long*>(bound));
}

void launch_block_apply(const MeshCache& m, const KeoValues& kv, const double* blk,
                        const double* x, double* y, hipStream_t s, hipEvent_t t0,
                        hipEvent_t t1) {
  const auto xv = reinterpret_cast<const double2*>(x);
  const auto yv = reinterpret_cast<double2*>(y);
  if (t0 || t1)
    hipExtLaunchKernelGGL(k_block_apply<double>, dim3(grid_for(m.n, kBlock)), dim3(kBlock), 0, s,
                          t0, t1, 0, m.n, sell_view(m), static_cast<const double2*>(kv.val.p), blk, xv, yv);
  else
    k_block_apply<double><<<grid_for(m.n, kBlock), kBlock, 0, s>>>(m.n, sell_view(m), kv.val.p,
                                                                   blk, xv, yv);
}

namespace {
__global__ void k_fill(double* __restrict__ a, int64_t n, double v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}
__global__ void k_scale(const double* __restrict__ a, double v, double* __restrict__ c,
                        int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) c[i] = v * a[i];
}
}  // namespace

void launch_scale(const double* a, double v, double* c, int64_t n, hipStream_t s) {
  k_scale<<<grid_for(n, kBlock), kBlock, 0, s>>>(a, v, c, n);
}

void launch_fill(double* a, int64_t n, double v, hipStream_t s) {
  k_fill<<<grid_for(n, kBlock), kBlock, 0, s>>>(a, n, v);
}

void launch_mul(const double* a, const double* b, double* c, int64_t n, hipStream_t s) {
  k_mul<<<grid_for(n, kBlock), kBlock, 0, s>>>(a, b, c, n);
}

void launch_block_apply32(const MeshCache& m, const KeoValues& kv, const float* blk,
                          const float* x, float* y, hipStream_t s, hipEvent_t t0,
                          hipEvent_t t1) {
  const auto xv = reinterpret_cast<const float2*>(x);
  const auto yv = reinterpret_cast<float2*>(y);
  if (t0 || t1)
    hipExtLaunchKernelGGL(k_block_apply<float>, dim3(grid_for(m.n, kBlock)), dim3(kBlock), 0, s,
                          t0, t1, 0, m.n, sell_view(m), kv.val32.p, blk, xv, yv);
  else
    k_block_apply<float><<<grid_for(m.n, kBlock), kBlock, 0, s>>>(m.n, sell_view(m),
                                                                  kv.val32.p, blk, xv, yv);
}

}  // namespace nosh

namespace nosh {
namespace {
// Partial sums of (sum c |psi|^4, sum c) or (sum c Re(conj(phi) psi), sum c).
__global__ void k_energy_partials(int64_t n, const double* __restrict__ cv,
                                  const double2* __restrict__ a, const double2* __restrict__ b,
                                  double* __restrict__ partials) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double c = cv[i];
    const double2 x = a[i];
    if (b) {
      const double2 y = b[i];
      v[0] = c * (x.x * y.x + x.y * y.y);
    } else {
      const double r2 = x.x * x.x + x.y * x.y;
      v[0] = c * r2 * r2;
    }
    v[1] = c;
  }
  block_partials<2>(v, partials);
}

__global__ void k_energy_finalize(const double* __restrict__ partials, int nb, double sign,
                                  double* __restrict__ out) {
  double v[2];
  finalize_sum<2>(v, partials, nb);
  if (threadIdx.x == 0) *out = sign * v[0] / v[1];
}
}  // namespace

void launch_cv_weighted(const MeshCache& m, const double* a, const double* b, double* partials,
                        double* out, hipStream_t s) {
  const unsigned nb = grid_for(m.n, kBlock);
  k_energy_partials<<<nb, kBlock, 0, s>>>(m.n, m.cv_int.p, reinterpret_cast<const double2*>(a),
                                          reinterpret_cast<const double2*>(b), partials);
  k_energy_finalize<<<1, kFinalizeThreads, 0, s>>>(partials, int(nb), b ? 1.0 : -1.0, out);
}

}  // namespace nosh

namespace nosh {
namespace {
__global__ void k_slot_fields(const int32_t* __restrict__ sedge, const int32_t* __restrict__ scol,
                              const double* __restrict__ alpha,
                              const double* __restrict__ atilde, int64_t nnz,
                              SlotMeta* __restrict__ s32, double* __restrict__ a64,
                              double* __restrict__ t64) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  const int32_t code = sedge[j];
  double al = 0.0, at = 0.0;
  if (code >= 0) {
    al = alpha[code >> 1];
    at = (code & 1) ? -atilde[code >> 1] : atilde[code >> 1];
  }
  SlotMeta m;
  m.col = scol[j];
  m.alpha = float(al);
  m.at = float(at);
  s32[j] = m;
  a64[j] = al;
  t64[j] = at;
}

__global__ void k_scatter_lane(int64_t n, int bp, int b, const int32_t* __restrict__ perm,
                               const double2* __restrict__ src, double2* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i * bp + b] = src[perm[i]];
}

__global__ void k_gather_lane(int64_t n, int bp, int b, const int32_t* __restrict__ perm,
                              const double2* __restrict__ src, double2* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i * bp + b];
}
}  // namespace

void launch_slot_fields(const MeshCache& m, const double* alpha, const double* atilde,
                        SlotMeta* s32, double* a64, double* t64, double* diag, double* rowabs,
                        hipStream_t s) {
  k_slot_fields<<<grid_for(m.nnz_pad, kBlock), kBlock, 0, s>>>(
      m.sedge.p, m.scol.p, alpha, atilde, m.nnz_pad, s32, a64, t64);
  k_keo_diag<<<grid_for(m.n, kBlock), kBlock, 0, s>>>(m.slice_ptr.p, m.sedge.p, alpha, m.n, diag,
                                                      rowabs);
}

void launch_scatter_lane(int64_t n, int bp, int b, const int32_t* perm, const double* src,
                         double* dst, hipStream_t s) {
  k_scatter_lane<<<grid_for(n, kBlock), kBlock, 0, s>>>(
      n, bp, b, perm, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst));
}

void launch_gather_lane(int64_t n, int bp, int b, const int32_t* perm, const double* src,
                        double* dst, hipStream_t s) {
  k_gather_lane<<<grid_for(n, kBlock), kBlock, 0, s>>>(
      n, bp, b, perm, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst));
}

}  // namespace nosh
