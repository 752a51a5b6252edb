// batched.hip -- many Newton-MINRES solves, on meshes of uneven size, in one
// launch per phase.
//
// An instance is (mesh, mu, g, psi). Instances on the same mesh form a group;
// one kernel launch covers every instance of every group. Within a group the
// vectors are instance-interleaved ([row][lane], lane = instance, padded to a
// power of two bp), so the gather of x[col] for one row is one contiguous
// read of bp complex values. The KEO is never stored per instance: each padded
// slot keeps (alpha, a~_dir) once per mesh (a~_dir = +a~ for the lower
// endpoint's row, -a~ for the higher one's) and the value
//   K_ij = -alpha e^{-i mu_b a~_dir}
// is rebuilt in registers (FP32 __sincosf in the inner solver, FP64 sincos
// for F), trading ~20x fewer matrix bytes for cheap ALU work.
//
// Precision: F(psi), the Newton update and all MINRES scalars are FP64; the
// Krylov vectors, blocks and inner SpMVs are FP32 (the FP32-inner / FP64-outer
// scheme). MINRES runs in lockstep over the batch with per-instance scalars
// and done flags; one iteration is either a captured hipGraph (replayed) or
// a persistent cooperative kernel with grid-wide barriers between phases.
//
// Reference operators: keo::refill_ (src/parameter_matrix_keo.cpp:74-184),
// nls::compute_f_ (src/model_evaluator_nls.cpp:527-628), jacobian_operator
// (src/jacobian_operator.cpp:143-200), keo_regularized::rebuild
// (src/keo_regularized.cpp:181-264).

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>

#include "batched.hpp"
#include "kernels.hpp"
#include "minres_device.hpp"

namespace cg = cooperative_groups;

namespace nosh {
namespace {

struct Lane {
  const BGroup* G;
  int64_t row;
  int b;
  int inst;
  int64_t idx;  // row * bp + b
  bool valid;
};

__device__ __forceinline__ int group_of(const BatchArgs& a, int64_t vb) {
  int gi = 0;
  while (gi + 1 < a.ngroups && vb >= a.groups[gi + 1].block0) ++gi;
  return gi;
}

__device__ __forceinline__ Lane lane_ctx(const BatchArgs& a, int64_t vb) {
  Lane L;
  const int gi = group_of(a, vb);
  L.G = &a.groups[gi];
  const int bp = L.G->bp;
  L.b = threadIdx.x & (bp - 1);
  L.row = (vb - L.G->block0) * (blockDim.x / bp) + threadIdx.x / bp;
  L.inst = L.G->inst0 + L.b;
  L.idx = L.row * bp + L.b;
  L.valid = L.row < L.G->n && L.b < L.G->B;
  return L;
}

// Per-instance partial sums of one launched block, which covers the virtual
// blocks v0 .. v0 + fold - 1 of one group; v is the thread's sum over them
// (0 for padding lanes). Rows are reduced in order for every lane b, giving
// partial (v0 - block0) / fold of the instance.
__device__ __forceinline__ void lane_partials(const BatchArgs& a, int64_t v0, double v) {
  __shared__ double lds[1024];
  const BGroup& G = a.groups[group_of(a, v0)];
  lds[threadIdx.x] = v;
  __syncthreads();
  const int bp = G.bp;
  if (threadIdx.x < bp && threadIdx.x < G.B) {
    double s = 0.0;
    for (int r = 0; r < int(blockDim.x) / bp; ++r) s += lds[r * bp + threadIdx.x];
    a.partials[int64_t(G.inst0 + threadIdx.x) * a.maxblocks + (v0 - G.block0) / a.fold] = s;
  }
  __syncthreads();
}

template <class T>
struct Vt;
template <>
struct Vt<float> {
  using v2 = float2;
};
template <>
struct Vt<double> {
  using v2 = double2;
};

// Streaming accesses of a row's own entries (blocks, residuals, Chebyshev
// and MINRES work vectors) are non-temporal, so that they do not evict from
// L2 the x entries the neighbouring rows gather: with 8 lanes per row the
// streamed bytes per row are 8x those of a single instance, enough to push
// the gather window out of a 4 MB L2.
#ifndef NOSH_BATCH_NT
#define NOSH_BATCH_NT 1
#endif
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldn(const float2* p) {
#if NOSH_BATCH_NT
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(v.x, v.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ double2 ldn(const double2* p) {
#if NOSH_BATCH_NT
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
  return make_double2(v.x, v.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ float ldn(const float* p) {
#if NOSH_BATCH_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void stn(float2* p, float2 v) {
#if NOSH_BATCH_NT
  f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
#else
  *p = v;
#endif
}
__device__ __forceinline__ void stn(double2* p, double2 v) {
#if NOSH_BATCH_NT
  d2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<d2v*>(p));
#else
  *p = v;
#endif
}

__device__ __forceinline__ void sc(float x, float* s, float* c) { __sincosf(x, s, c); }
__device__ __forceinline__ void sc(double x, double* s, double* c) { sincos(x, s, c); }

// Off-diagonal product of one (row, instance) with values rebuilt in registers.
// The slot count of a row is its slice's width, the same for the whole
// wave; slots are taken 4 at a time so that their metadata loads, and then
// the x gathers, are all in flight together (4 measured faster than 1 or 8
// at 96^3 x 8 lanes). FP32: one 12-byte {column, alpha, a~} load per slot;
// FP64: column, alpha and a~ from separate arrays.
template <class T>
__device__ __forceinline__ void slot_meta(const BGroup& G, int32_t j, int32_t& c, T& al, T& at) {
  if constexpr (std::is_same<T, float>::value) {
    const SlotMeta m = G.s32[j];
    c = m.col;
    al = m.alpha;
    at = m.at;
  } else {
    c = G.scol[j];
    al = G.a64[j];
    at = G.t64[j];
  }
}

template <class T, int K>
__device__ __forceinline__ void brow_slots(const BGroup& G, int32_t j, int b, T mu,
                                           const typename Vt<T>::v2* __restrict__ x, T& ar,
                                           T& ai) {
  int32_t c[K];
  T av[K], tv[K];
#pragma unroll
  for (int q = 0; q < K; ++q) slot_meta<T>(G, j + q * kSliceH, c[q], av[q], tv[q]);
  typename Vt<T>::v2 xv[K];
  const int bp = G.bp;
#pragma unroll
  for (int q = 0; q < K; ++q) xv[q] = x[int64_t(c[q]) * bp + b];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    T sn, cs;
    sc(mu * tv[q], &sn, &cs);
    const T vr = -av[q] * cs, vi = av[q] * sn;
    ar += vr * xv[q].x - vi * xv[q].y;
    ai += vr * xv[q].y + vi * xv[q].x;
  }
}

constexpr int kSlotBatch = 4;

#ifndef NOSH_META_SHFL
#define NOSH_META_SHFL 1
#endif

// Groups whose rows take their slot metadata by lane shuffles: every lane is
// a real instance and the callers keep all lanes of a row in brow (lanes of
// converged instances compute and discard).
__device__ __forceinline__ bool meta_shfl(const BGroup& G) {
  return NOSH_META_SHFL && G.B == G.bp && G.bp >= 4;
}

// FP32 rows of a full group (B == bp >= 4 lanes, all lanes active): lane q of
// the row's lanes loads the metadata of slot k + q once and the others take
// it by a lane shuffle, instead of every lane loading the same 12 bytes
// (which costs the texture data path bp times the useful bytes).
__device__ __forceinline__ void brow_shfl4(const BGroup& G, int32_t j, int32_t k, int32_t w,
                                           int b, float mu, const float2* __restrict__ x,
                                           float& ar, float& ai) {
  const int lane = int(threadIdx.x & 63);
  const int g0 = lane & ~(G.bp - 1);
  SlotMeta mine = {0, 0.0f, 0.0f};
  if (b < 4 && k + b < w) mine = G.s32[j + b * kSliceH];
  int32_t c[4];
  float av[4], tv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    c[q] = __shfl(mine.col, g0 + q, 64);
    av[q] = __shfl(mine.alpha, g0 + q, 64);
    tv[q] = __shfl(mine.at, g0 + q, 64);
  }
  float2 xv[4];
  const int bp = G.bp;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (k + q < w) xv[q] = x[int64_t(c[q]) * bp + b];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (k + q < w) {
      float sn, cs;
      sc(mu * tv[q], &sn, &cs);
      const float vr = -av[q] * cs, vi = av[q] * sn;
      ar += vr * xv[q].x - vi * xv[q].y;
      ai += vr * xv[q].y + vi * xv[q].x;
    }
  }
}

#ifndef NOSH_META_PIPE
#define NOSH_META_PIPE 1
#endif

// The shuffled row product over a whole row, NS slots per batch (lanes
// b < NS of the row each load one slot's metadata; NS <= bp; 8 measured
// slower than 4 at 96^3 x 8 lanes: 178 vs 164 us, occupancy 6 instead of 8),
// and software-pipelined: the metadata load of the next batch is issued right
// after the x gathers of the current one, so the two dependent memory
// latencies of consecutive batches overlap. Same arithmetic in the same
// order as the unpipelined loop.
template <int NS>
__device__ __forceinline__ void brow_shfl_row(const BGroup& G, int32_t j, int32_t w, int b,
                                              float mu, const float2* __restrict__ x, float& ar,
                                              float& ai) {
  const int lane = int(threadIdx.x & 63);
  const int g0 = lane & ~(G.bp - 1);
  const int bp = G.bp;
  SlotMeta mine = {0, 0.0f, 0.0f};
  if (b < NS && b < w) mine = G.s32[j + b * kSliceH];
  for (int32_t k = 0; k < w; k += NS, j += NS * kSliceH) {
    int32_t c[NS];
    float av[NS], tv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      c[q] = __shfl(mine.col, g0 + q, 64);
      av[q] = __shfl(mine.alpha, g0 + q, 64);
      tv[q] = __shfl(mine.at, g0 + q, 64);
    }
    float2 xv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (k + q < w) xv[q] = x[int64_t(c[q]) * bp + b];
    if (b < NS && k + NS + b < w) mine = G.s32[j + (NS + b) * kSliceH];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (k + q < w) {
        float sn, cs;
        sc(mu * tv[q], &sn, &cs);
        const float vr = -av[q] * cs, vi = av[q] * sn;
        ar += vr * xv[q].x - vi * xv[q].y;
        ai += vr * xv[q].y + vi * xv[q].x;
      }
    }
  }
}

// The shuffled, pipelined row product of brow_shfl_row over the group's
// precomputed slot values (every active lane shares mu): no sine and cosine
// per slot and lane. Same values, same order, same bits.
template <int NS>
__device__ __forceinline__ void brow_shfl_val(const BGroup& G, int32_t j, int32_t w, int b,
                                              const float2* __restrict__ x, float& ar,
                                              float& ai) {
  const int lane = int(threadIdx.x & 63);
  const int g0 = lane & ~(G.bp - 1);
  const int bp = G.bp;
  SlotVal mine = {0, 0.0f, 0.0f};
  if (b < NS && b < w) mine = G.kval[j + b * kSliceH];
  for (int32_t k = 0; k < w; k += NS, j += NS * kSliceH) {
    int32_t c[NS];
    float vr[NS], vi[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      c[q] = __shfl(mine.col, g0 + q, 64);
      vr[q] = __shfl(mine.vr, g0 + q, 64);
      vi[q] = __shfl(mine.vi, g0 + q, 64);
    }
    float2 xv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (k + q < w) xv[q] = x[int64_t(c[q]) * bp + b];
    if (b < NS && k + NS + b < w) mine = G.kval[j + (NS + b) * kSliceH];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (k + q < w) {
        ar += vr[q] * xv[q].x - vi[q] * xv[q].y;
        ai += vr[q] * xv[q].y + vi[q] * xv[q].x;
      }
    }
  }
}

template <class T>
__device__ __forceinline__ typename Vt<T>::v2 brow(const BGroup& G, int64_t row, int b, T mu,
                                                    const typename Vt<T>::v2* __restrict__ x) {
  // A wave's rows lie in one 64-row slice (it holds 64 / bp consecutive rows
  // starting at a multiple of 64 / bp), so the slice bounds are wave-uniform
  // and come through the scalar cache.
  const int32_t s = uniform(int32_t(row >> 6));
  const int32_t beg = ld_uniform(G.slice_ptr + s), end = ld_uniform(G.slice_ptr + s + 1);
  const int32_t w = (end - beg) / kSliceH;
  T ar = T(0), ai = T(0);
  int32_t j = beg + int32_t(row & 63);
  int32_t k = 0;
  if constexpr (std::is_same<T, float>::value) {
    if (meta_shfl(G)) {
      if (G.kval)
        brow_shfl_val<4>(G, j, w, b, x, ar, ai);
      else if (NOSH_META_PIPE)
        brow_shfl_row<4>(G, j, w, b, mu, x, ar, ai);
      else
        for (; k < w; k += 4, j += 4 * kSliceH) brow_shfl4(G, j, k, w, b, mu, x, ar, ai);
      typename Vt<T>::v2 r;
      r.x = ar;
      r.y = ai;
      return r;
    }
  }
  for (; k + kSlotBatch <= w; k += kSlotBatch, j += kSlotBatch * kSliceH)
    brow_slots<T, kSlotBatch>(G, j, b, mu, x, ar, ai);
  for (; k < w; ++k, j += kSliceH) brow_slots<T, 1>(G, j, b, mu, x, ar, ai);
  typename Vt<T>::v2 r;
  r.x = ar;
  r.y = ai;
  return r;
}

// ---------------------------------------------------------------- phases
// Every phase is a device function over one virtual block, used both by the
// per-phase kernels (graph mode) and by the persistent kernel.

__device__ double ph_compute_f(const BatchArgs& a, int64_t vb) {
  const Lane L = lane_ctx(a, vb);
  double s2 = 0.0;
  if (L.valid) {
    const BGroup& G = *L.G;
    const double mu = a.mu[L.inst], g = a.g[L.inst];
    const double2 acc = brow<double>(G, L.row, L.b, mu, G.psi);
    const double2 xi = G.psi[L.idx];
    const double d = G.diag[L.row] + G.ct[L.row] * (G.V[L.row] + g * (xi.x * xi.x + xi.y * xi.y));
    const double2 r = make_double2(acc.x + d * xi.x, acc.y + d * xi.y);
    G.f[L.idx] = r;
    s2 = r.x * r.x + r.y * r.y;
  }
  return s2;
}

// J and P blocks (FP32, including diag K) and the per-instance Chebyshev bound.
__device__ void ph_rebuild(const BatchArgs& a, int64_t vb) {
  const Lane L = lane_ctx(a, vb);
  __shared__ double rmax[1024];
  rmax[threadIdx.x] = 0.0;
  if (L.valid) {
  const BGroup& G = *L.G;
  const int64_t nb3 = G.n * G.bp;
  const double g = a.g[L.inst];
  const double2 p = G.psi[L.idx];
  const double c = G.ct[L.row], d = G.diag[L.row];
  const double r2 = p.x * p.x + p.y * p.y;
  const double al = c * (G.V[L.row] + g * 2.0 * r2);
  const double rx2 = g * c * (p.x * p.x - p.y * p.y);
  const double d1b = g * c * (2.0 * p.x * p.y);
  G.jblk[L.idx] = float(d + al + rx2);
  G.jblk[nb3 + L.idx] = float(d + al - rx2);
  G.jblk[2 * nb3 + L.idx] = float(d1b);
  double a0 = d, a1 = d, bb = 0.0;
  if (g > 0.0) {
    const double pa = g * c * 2.0 * r2;
    a0 = d + pa + rx2;
    a1 = d + pa - rx2;
    bb = d1b;
  }
  G.pblk[L.idx] = float(a0);
  G.pblk[nb3 + L.idx] = float(a1);
  G.pblk[2 * nb3 + L.idx] = float(bb);
  const double lmin = 0.5 * (a0 + a1) - hypot(0.5 * (a0 - a1), bb);
  rmax[threadIdx.x] = lmin > 0.0 ? G.rowabs[L.row] / lmin : 1.0;
  }
  __syncthreads();
  // One atomic per (block, instance): the max over the block's rows.
  const int bp = L.G->bp;
  if (threadIdx.x < bp && threadIdx.x < L.G->B) {
    double m = 0.0;
    for (int r = 0; r < int(blockDim.x) / bp; ++r) m = fmax(m, rmax[r * bp + threadIdx.x]);
    atomicMax(reinterpret_cast<unsigned long long*>(a.bound) + L.G->inst0 + threadIdx.x,
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
  // rmax is rewritten by the next row tile of a folded block.
  __syncthreads();
}

__device__ void ph_setup(const BatchArgs& a, int64_t vb) {
  const Lane L = lane_ctx(a, vb);
  if (!L.valid) return;
  const BGroup& G = *L.G;
  const double2 fv = G.f[L.idx];
  const float2 r = make_float2(float(-fv.x), float(-fv.y));
  G.R0[L.idx] = r;
  G.R1[L.idx] = r;
  G.W0[L.idx] = make_float2(0.f, 0.f);
  G.W2[L.idx] = make_float2(0.f, 0.f);
  G.x[L.idx] = make_double2(0.0, 0.0);
}

__device__ __forceinline__ double hi_of(const BatchArgs& a, int inst) {
  return a.upper > 0.0 ? a.upper : 1.0 + a.bound[inst];
}

// Buffer selectors choose between pointer values (not between the record's
// fields), so that a record held in registers stays there when the choice
// differs between lanes.
__device__ __forceinline__ float2* pick(bool c, float2* a, float2* b) { return c ? a : b; }
__device__ __forceinline__ float2* rbuf(const BGroup& G, int k) { return pick(k & 1, G.R1, G.R0); }
__device__ __forceinline__ float2* wbuf(const BGroup& G, int k) {
  const int q = ((k % 3) + 3) % 3;
  return pick(q == 0, G.W0, pick(q == 1, G.W1, G.W2));
}
__device__ __forceinline__ float2* zbuf(const BGroup& G, int k) { return pick(k & 1, G.Z2, G.Y); }
__device__ __forceinline__ float2* vbuf(const BGroup& G, int k) { return pick(k & 1, G.v1, G.v); }

// Per-thread context of a launched block. All fold row tiles of a block
// belong to one group and a thread keeps its lane (instance) in all of them,
// so the group record and the instance's MINRES scalars are read once per
// thread, not once per tile. The record is copied into registers by scalar
// loads at kernel entry (its index made wave-uniform): read through a pointer
// it would be reloaded by vector loads after every store of a row tile, which
// the compiler cannot prove does not alias it, with a wait on each reload.
struct Ctx {
  BGroup G;
  int b, inst, rows_per_vb;
  int64_t row0;  // the thread's row in the block's first tile
  bool lane_ok;  // a real instance (b < B)
};

__device__ __forceinline__ Ctx make_ctx(const BatchArgs& a, int64_t v0) {
  Ctx c;
  // Read through the constant address space: the record is never written by
  // a kernel, and pointers loaded from constant memory are known to address
  // global memory, so the row loops use global (not flat) loads.
  const int gi = __builtin_amdgcn_readfirstlane(group_of(a, v0));
#if __HIP_DEVICE_COMPILE__
  typedef const __attribute__((address_space(4))) BGroup* ConstGroups;
  c.G = ((ConstGroups)a.groups)[gi];
#else
  c.G = a.groups[gi];  // host pass only parses device functions
#endif
  const int bp = c.G.bp;
  c.b = threadIdx.x & (bp - 1);
  c.inst = c.G.inst0 + c.b;
  c.rows_per_vb = int(blockDim.x) / bp;
  c.row0 = (v0 - c.G.block0) * c.rows_per_vb + threadIdx.x / bp;
  c.lane_ok = c.b < c.G.B;
  return c;
}

// Pending update of iteration k: w_k = (v_k - oldeps w_{k-2} - delta w_{k-1}) / gamma,
// x += phi w_k.
struct WUpd {
  float oe, de, dn;
  double phi;
};

__device__ __forceinline__ void w_update(const BGroup& G, const WUpd& q, int k, int64_t idx,
                                         float2 vk) {
  float2* wk = wbuf(G, k);
  const float2 w1 = ldn(wbuf(G, k - 2) + idx), w2 = ldn(wbuf(G, k - 1) + idx);
  const float2 w = make_float2((vk.x - q.oe * w1.x - q.de * w2.x) * q.dn,
                               (vk.y - q.oe * w1.y - q.de * w2.y) * q.dn);
  stn(wk + idx, w);
  double2 xv = ldn(G.x + idx);
  xv.x += q.phi * double(w.x);
  xv.y += q.phi * double(w.y);
  stn(G.x + idx, xv);
}

__device__ __forceinline__ WUpd wupd_of(const MinresScalars& S) {
  WUpd q;
  q.oe = float(S.p_oldeps);
  q.de = float(S.p_delta);
  q.dn = float(S.p_denom);
  q.phi = S.p_phi;
  return q;
}

// Chebyshev step 0 over the block's rows. mode 1: r = R0 (MINRES start);
// mode 2: r_new = u - (alfa/beta) r2. Returns the thread's <r, z> sum.
__device__ double rows_cheb_first(const BatchArgs& a, const Ctx& c, int mode) {
  double s = 0.0;
  if (!c.lane_ok) return s;
  const MinresScalars& S = a.S[c.inst];
  const int done = S.done, it = mode == 2 ? S.it : 0;
  const double alfa = S.alfa, beta = S.beta;
  if (done) return s;
  const BGroup& G = c.G;
  // The previous iteration's pending update of w and x (Givens coefficients
  // from its last reduction, v from its Lanczos step), kept out of the
  // gathering kernel to spare its registers.
  const bool pending = mode == 2 && S.pending;
  const WUpd q = wupd_of(S);
  const float2* vprev = vbuf(G, it - 1);
  float2* r = rbuf(G, it);
  const float2* r2 = rbuf(G, it + 1);
  const float cab = mode == 2 ? float(alfa / beta) : 0.0f;
  const double hi = hi_of(a, c.inst);
  const float th = float(0.5 * (hi + hi / a.ratio));
  const int64_t nb3 = G.n * G.bp;
  float2* z0 = zbuf(G, a.z0);
  for (int t = 0; t < a.fold; ++t) {
    const int64_t row = c.row0 + int64_t(t) * c.rows_per_vb;
    if (row >= G.n) break;
    const int64_t idx = row * G.bp + c.b;
    if (pending) w_update(G, q, it, idx, ldn(vprev + idx));
    float2 rv;
    if (mode == 2) {
      const float2 uu = ldn(G.u + idx), r2v = ldn(r2 + idx);
      rv = make_float2(uu.x - cab * r2v.x, uu.y - cab * r2v.y);
      stn(r + idx, rv);
    } else {
      rv = ldn(r + idx);
    }
    float2 z;
    if (a.iters == 0) {
      z = rv;
    } else {
      float o0, o1;
      dinv<float>(ldn(G.pblk + idx), ldn(G.pblk + nb3 + idx), ldn(G.pblk + 2 * nb3 + idx), rv.x,
                  rv.y, &o0, &o1);
      z = make_float2(o0 / th, o1 / th);
    }
    stn(z0 + idx, z);
    s += double(rv.x) * double(z.x) + double(rv.y) * double(z.y);
  }
  return s;
}

// Chebyshev step j over the block's rows; returns the thread's <r, z_j> sum.
// Three-term form z_j = z_{j-1} + c1 (z_{j-1} - z_{j-2}) + c2 D^-1 (r - A z_{j-1})
// with z_{-1} = 0: z_{j-2} sits in z_j's buffer (same parity), is read and
// overwritten in place element by element (no other thread reads it), so no
// separate direction vector is kept: 8 bytes less per row and lane for
// every step, 16 for the first, and none for the first step's z_{j-2}.
__device__ double rows_cheb_step(const BatchArgs& a, const Ctx& c, int mode, int j) {
  double s = 0.0;
  if (!c.lane_ok) return s;
  const MinresScalars& S = a.S[c.inst];
  const int done = S.done, it = mode == 2 ? S.it : 0;
  const BGroup& G = c.G;
  if (meta_shfl(G) ? __all(done) : done) return s;
  const bool last = (j == a.iters - 1);
  const float2* r = rbuf(G, it);
  const float2* zin = zbuf(G, a.z0 + j - 1);
  float2* zout = zbuf(G, a.z0 + j);
  const Cheb cc = cheb_coef(hi_of(a, c.inst), a.ratio, j);
  const float c1 = float(cc.c1), c2 = float(cc.c2);
  const float mu = float(a.mu[c.inst]);
  const int64_t nb3 = G.n * G.bp;
  for (int t = 0; t < a.fold; ++t) {
    const int64_t row = c.row0 + int64_t(t) * c.rows_per_vb;
    if (row >= G.n) break;
    const int64_t idx = row * G.bp + c.b;
    const float2 acc = brow<float>(G, row, c.b, mu, zin);
    if (done) continue;
    const float2 zi = zin[idx];
    const float b0 = ldn(G.pblk + idx), b1 = ldn(G.pblk + nb3 + idx),
                b2 = ldn(G.pblk + 2 * nb3 + idx);
    const float2 rv = ldn(r + idx);
    const float res0 = rv.x - (acc.x + b0 * zi.x + b2 * zi.y);
    const float res1 = rv.y - (acc.y + b2 * zi.x + b1 * zi.y);
    float u0, u1;
    dinv<float>(b0, b1, b2, res0, res1, &u0, &u1);
    float2 d = zi;
    if (j > 1) {
      const float2 zpp = ldn(zout + idx);
      d.x -= zpp.x;
      d.y -= zpp.y;
    }
    d.x = c1 * d.x + c2 * u0;
    d.y = c1 * d.y + c2 * u1;
    const float2 z = make_float2(zi.x + d.x, zi.y + d.y);
    stn(zout + idx, z);
    if (last) s += double(rv.x) * double(z.x) + double(rv.y) * double(z.y);
  }
  return s;
}


// v = y / beta, u = J v - (beta/beta_old) r1, the previous w/x update, over
// the block's rows; returns the thread's <v, u> sum.
__device__ double rows_minres_a(const BatchArgs& a, const Ctx& c) {
  double s = 0.0;
  if (!c.lane_ok) return s;
  const MinresScalars& S = a.S[c.inst];
  const int done = S.done, it = S.it;
  const double beta = S.beta, oldb = S.oldb;
  const BGroup& G = c.G;
  if (meta_shfl(G) ? __all(done) : done) return s;
  const float2* y = zbuf(G, 0);
  const float2* r1 = rbuf(G, it);
  const float ib = float(1.0 / beta);
  const float cb = it >= 1 ? float(beta / oldb) : 0.0f;
  const float mu = float(a.mu[c.inst]);
  const int64_t nb3 = G.n * G.bp;
  for (int t = 0; t < a.fold; ++t) {
    const int64_t row = c.row0 + int64_t(t) * c.rows_per_vb;
    if (row >= G.n) break;
    const int64_t idx = row * G.bp + c.b;
    const float2 acc = brow<float>(G, row, c.b, mu, y);
    if (done) continue;
    const float2 yi = y[idx];
    const float b0 = ldn(G.jblk + idx), b1 = ldn(G.jblk + nb3 + idx),
                b2 = ldn(G.jblk + 2 * nb3 + idx);
    float2 uu = make_float2((acc.x + b0 * yi.x + b2 * yi.y) * ib,
                            (acc.y + b2 * yi.x + b1 * yi.y) * ib);
    if (it >= 1) {
      const float2 rv = ldn(r1 + idx);
      uu.x -= cb * rv.x;
      uu.y -= cb * rv.y;
    }
    const float2 vn = make_float2(yi.x * ib, yi.y * ib);
    stn(vbuf(G, it) + idx, vn);
    stn(G.u + idx, uu);
    s += double(vn.x) * double(uu.x) + double(vn.y) * double(uu.y);
  }
  return s;
}

__device__ void ph_flush(const BatchArgs& a, int64_t vb) {
  const Lane L = lane_ctx(a, vb);
  if (!L.valid || !a.active[L.inst]) return;
  const BGroup& G = *L.G;
  const MinresScalars& S = a.S[L.inst];
  if (S.pending) w_update(G, wupd_of(S), S.it, L.idx, vbuf(G, S.it - 1)[L.idx]);
  const double2 d = G.x[L.idx];
  double2 p = G.psi[L.idx];
  p.x += d.x;
  p.y += d.y;
  G.psi[L.idx] = p;
}

// Per-instance finalisation (block = instance): 0 norm^2 -> out, 1 init,
// 2 alfa, 3 Givens.
__device__ void ph_finalize(const BatchArgs& a, int inst, int what, double* out) {
  MinresScalars* S = &a.S[inst];
  if ((what == 2 || what == 3) && S->done) return;
  int gi = 0;
  while (gi + 1 < a.ngroups && inst >= a.groups[gi + 1].inst0) ++gi;
  const int nb = int(a.groups[gi].nblocks / a.fold);
  double v[1];
  finalize_sum<1>(v, a.partials + int64_t(inst) * a.maxblocks, nb);
  if (threadIdx.x == 0) {
    if (what == 0)
      out[inst] = v[0];
    else if (what == 1)
      init_update(S, v[0]);
    else if (what == 2)
      S->alfa = v[0];  // the pending w/x update is applied by the next phase
    else
      givens_update(S, v[0]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- kernels
// Launched block rb covers the virtual blocks rb*fold .. rb*fold + fold - 1,
// all of one group (every group's virtual block count is a multiple of fold).
__device__ __forceinline__ void blk_compute_f(const BatchArgs& a, int64_t rb) {
  const int64_t v0 = rb * a.fold;
  double s = 0.0;
  for (int r = 0; r < a.fold; ++r) s += ph_compute_f(a, v0 + r);
  lane_partials(a, v0, s);
}
__device__ __forceinline__ void blk_cheb_first(const BatchArgs& a, int64_t rb, int mode) {
  const int64_t v0 = rb * a.fold;
  const double s = rows_cheb_first(a, make_ctx(a, v0), mode);
  if (a.iters <= 1) lane_partials(a, v0, s);
}
__device__ __forceinline__ void blk_cheb_step(const BatchArgs& a, int64_t rb, int mode, int j) {
  const int64_t v0 = rb * a.fold;
  const double s = rows_cheb_step(a, make_ctx(a, v0), mode, j);
  if (j == a.iters - 1) lane_partials(a, v0, s);
}
__device__ __forceinline__ void blk_minres_a(const BatchArgs& a, int64_t rb) {
  const int64_t v0 = rb * a.fold;
  lane_partials(a, v0, rows_minres_a(a, make_ctx(a, v0)));
}

// Launched blocks are remapped so that each XCD works on one contiguous
// range of row tiles and its L2 holds the x entries those rows gather.
__global__ void kb_compute_f(BatchArgs a) { blk_compute_f(a, xcd_block(true)); }
__global__ void kb_rebuild(BatchArgs a) {
  const int64_t rb = xcd_block(true);
  for (int r = 0; r < a.fold; ++r) ph_rebuild(a, rb * a.fold + r);
}
__global__ void kb_setup(BatchArgs a) {
  const int64_t rb = xcd_block(true);
  for (int r = 0; r < a.fold; ++r) ph_setup(a, rb * a.fold + r);
}
__global__ void kb_cheb_first(BatchArgs a, int mode) {
  blk_cheb_first(a, xcd_block(true), mode);
}
__global__ void kb_cheb_step(BatchArgs a, int mode, int j) {
  blk_cheb_step(a, xcd_block(true), mode, j);
}
__global__ void kb_minres_a(BatchArgs a) { blk_minres_a(a, xcd_block(true)); }
__global__ void kb_flush(BatchArgs a) {
  const int64_t rb = xcd_block(true);
  for (int r = 0; r < a.fold; ++r) ph_flush(a, rb * a.fold + r);
}
__global__ void kb_finalize(BatchArgs a, int what, double* out) {
  ph_finalize(a, blockIdx.x, what, out);
}

// Per-instance MINRES state reset: active instances start, others are done.
__global__ void kb_reset(BatchArgs a, int maxit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ninst) return;
  MinresScalars* S = &a.S[i];
  S->done = a.active[i] ? 0 : 1;
  S->it = 0;
  S->pending = 0;
  S->rtol = a.rtol[i];
  S->maxit = maxit;
  S->relres = 1.0;
  a.bound[i] = 0.0;
}

// Slot values at one mu for a whole group (see SlotVal).
__global__ void kb_fill_kval(const SlotMeta* __restrict__ s32, int64_t nnz, float mu,
                             SlotVal* __restrict__ kval) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  const SlotMeta m = s32[j];
  float sn, cs;
  sc(mu * m.at, &sn, &cs);
  SlotVal v;
  v.col = m.col;
  v.vr = -m.alpha * cs;
  v.vi = m.alpha * sn;
  kval[j] = v;
}

// Persistent MINRES: all iterations of the batch in one cooperative launch;
// grid-wide barriers separate the phases; the loop ends when every instance
// is done or after maxit iterations, a decision every block takes from the
// same post-barrier state.
__global__ void kb_minres_persistent(BatchArgs a, int maxit) {
  cg::grid_group grid = cg::this_grid();
  const int64_t nrb = a.total_blocks / a.fold;
  for (int it = 0; it < maxit; ++it) {
    for (int64_t rb = blockIdx.x; rb < nrb; rb += gridDim.x) blk_minres_a(a, rb);
    grid.sync();
    for (int inst = blockIdx.x; inst < a.ninst; inst += gridDim.x)
      ph_finalize(a, inst, 2, nullptr);
    grid.sync();
    for (int64_t rb = blockIdx.x; rb < nrb; rb += gridDim.x) blk_cheb_first(a, rb, 2);
    grid.sync();
    for (int j = 1; j < a.iters; ++j) {
      for (int64_t rb = blockIdx.x; rb < nrb; rb += gridDim.x) blk_cheb_step(a, rb, 2, j);
      grid.sync();
    }
    for (int inst = blockIdx.x; inst < a.ninst; inst += gridDim.x)
      ph_finalize(a, inst, 3, nullptr);
    grid.sync();
    int all_done = 1;
    for (int i = 0; i < a.ninst; ++i) all_done &= a.S[i].done;
    if (all_done) break;
  }
}

}  // namespace

// ------------------------------------------------------------------ host
void Batch::launch_phases_iteration(hipStream_t s) {
  const unsigned nb = unsigned(args.total_blocks / args.fold), bt = unsigned(block_threads);
  kb_minres_a<<<nb, bt, 0, s>>>(args);
  kb_finalize<<<args.ninst, 256, 0, s>>>(args, 2, nullptr);
  kb_cheb_first<<<nb, bt, 0, s>>>(args, 2);
  for (int j = 1; j < args.iters; ++j) kb_cheb_step<<<nb, bt, 0, s>>>(args, 2, j);
  kb_finalize<<<args.ninst, 256, 0, s>>>(args, 3, nullptr);
}

void Batch::build_graph(hipStream_t s) {
  if (graph_exec) return;
  NOSH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < graph_iters; ++q) launch_phases_iteration(s);
  hipGraph_t g = nullptr;
  NOSH_HIP(hipStreamEndCapture(s, &g));
  NOSH_HIP(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0));
  NOSH_HIP(hipGraphDestroy(g));
}

void Batch::drop_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  graph_exec = nullptr;
}

int Batch::persistent_grid() {
  int per_cu = 0;
  NOSH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(kb_minres_persistent), block_threads, 0));
  int dev = 0, cus = 0;
  NOSH_HIP(hipGetDevice(&dev));
  NOSH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t cap = int64_t(std::max(per_cu, 1)) * cus;
  return int(std::min<int64_t>(cap, args.total_blocks / args.fold));
}

void Batch::set_mu(const double* m, hipStream_t s) {
  mu_host.assign(m, m + args.ninst);
  mu.upload(m, args.ninst, s);
}

void Batch::prepare_kval(hipStream_t s) {
  bool changed = false;
  for (auto& gh : ghost) {
    BGroup& G = gh.dev;
    bool shared = NOSH_META_SHFL && G.B == G.bp && G.bp >= 4;
    bool any = false;
    double m0 = 0.0;
    for (int b = 0; shared && b < G.B; ++b) {
      const int i = G.inst0 + b;
      if (!active_h[i]) continue;
      if (!any) {
        m0 = mu_host[i];
        any = true;
      } else if (mu_host[i] != m0) {
        shared = false;
      }
    }
    const SlotVal* want = nullptr;
    if (shared && any) {
      if (!gh.kval.p) gh.kval.alloc(gh.mesh->nnz_pad);
      if (!gh.kval_set || gh.kval_mu != m0) {
        kb_fill_kval<<<grid_for(gh.mesh->nnz_pad, 256), 256, 0, s>>>(
            G.s32, gh.mesh->nnz_pad, float(m0), gh.kval.p);
        gh.kval_mu = m0;
        gh.kval_set = true;
      }
      want = gh.kval.p;
    }
    if (G.kval != want) {
      G.kval = want;
      changed = true;
    }
  }
  if (changed) {
    records_host.clear();
    for (auto& gh : ghost) records_host.push_back(gh.dev);
    NOSH_HIP(hipMemcpyAsync(groups.p, records_host.data(), sizeof(BGroup) * records_host.size(),
                            hipMemcpyHostToDevice, s));
    NOSH_HIP(hipStreamSynchronize(s));
  }
}

int Batch::minres(hipStream_t s, int maxit) {
  NOSH_CHECK(mode == 0 || args.total_blocks / args.fold <= kPersistentMaxBlocks, kUnsupported,
             "persistent grid too large");
  const unsigned nb = unsigned(args.total_blocks / args.fold), bt = unsigned(block_threads);
  prepare_kval(s);
  kb_reset<<<grid_for(args.ninst, 128), 128, 0, s>>>(args, maxit);
  kb_rebuild<<<nb, bt, 0, s>>>(args);
  kb_setup<<<nb, bt, 0, s>>>(args);
  // y = M^{-1} r0 and beta1 for every instance.
  kb_cheb_first<<<nb, bt, 0, s>>>(args, 1);
  for (int j = 1; j < args.iters; ++j) kb_cheb_step<<<nb, bt, 0, s>>>(args, 1, j);
  kb_finalize<<<args.ninst, 256, 0, s>>>(args, 1, nullptr);
  int issued = 0;
  if (mode == 1) {
    void* kargs[] = {&args, &maxit};
    NOSH_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kb_minres_persistent),
                                        dim3(persistent_grid()), dim3(bt), kargs, 0, s));
    issued = maxit;
  } else {
    build_graph(s);
    for (;;) {
      for (int q = 0; q < poll; q += graph_iters, issued += graph_iters)
        NOSH_HIP(hipGraphLaunch(graph_exec, s));
      NOSH_HIP(hipMemcpyAsync(Sh.data(), args.S, sizeof(MinresScalars) * args.ninst,
                              hipMemcpyDeviceToHost, s));
      NOSH_HIP(hipStreamSynchronize(s));
      bool all = true;
      for (int i = 0; i < args.ninst; ++i) all = all && Sh[i].done;
      if (all || issued >= maxit) break;
    }
  }
  kb_flush<<<nb, bt, 0, s>>>(args);
  NOSH_HIP(hipMemcpyAsync(Sh.data(), args.S, sizeof(MinresScalars) * args.ninst,
                          hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  int total = 0;
  for (int i = 0; i < args.ninst; ++i)
    if (active_h[i]) total = std::max(total, Sh[i].it);
  (void)issued;
  return total;
}

double Batch::time_cheb_step(hipStream_t s, int reps) {
  NOSH_CHECK(args.iters >= 2 && reps >= 1, kInvalid,
             "timing a Chebyshev step needs degree >= 2 and reps >= 1");
  const unsigned nb = unsigned(args.total_blocks / args.fold), bt = unsigned(block_threads);
  std::vector<int> all(args.ninst, 1);
  NOSH_HIP(hipMemcpyAsync(active.p, all.data(), sizeof(int) * args.ninst, hipMemcpyHostToDevice,
                          s));
  const std::vector<int> active_before = active_h;
  active_h.assign(args.ninst, 1);
  prepare_kval(s);
  kb_reset<<<grid_for(args.ninst, 128), 128, 0, s>>>(args, 1);
  kb_rebuild<<<nb, bt, 0, s>>>(args);
  kb_setup<<<nb, bt, 0, s>>>(args);
  kb_cheb_first<<<nb, bt, 0, s>>>(args, 1);
  kb_cheb_step<<<nb, bt, 0, s>>>(args, 1, 1);
  // A step past the first (it reads z_{j-2}); repeated, it rewrites z_{j-2}
  // in place with bounded values (|c1| < 1).
  const int j = std::min(2, args.iters - 1);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  NOSH_HIP(hipEventCreate(&e0));
  NOSH_HIP(hipEventCreate(&e1));
  NOSH_HIP(hipEventRecord(e0, s));
  for (int q = 0; q < reps; ++q) kb_cheb_step<<<nb, bt, 0, s>>>(args, 1, j);
  NOSH_HIP(hipEventRecord(e1, s));
  NOSH_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  NOSH_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // The instances are idle again.
  active_h = active_before;
  std::vector<int> none(args.ninst, 0);
  NOSH_HIP(hipMemcpyAsync(active.p, none.data(), sizeof(int) * args.ninst, hipMemcpyHostToDevice,
                          s));
  kb_reset<<<grid_for(args.ninst, 128), 128, 0, s>>>(args, 1);
  NOSH_HIP(hipStreamSynchronize(s));
  return 1e3 * double(ms) / reps;
}

void Batch::compute_norms(hipStream_t s, std::vector<double>& nrm) {
  const unsigned nb = unsigned(args.total_blocks / args.fold), bt = unsigned(block_threads);
  kb_compute_f<<<nb, bt, 0, s>>>(args);
  kb_finalize<<<args.ninst, 256, 0, s>>>(args, 0, norms.p);
  nrm.resize(args.ninst);
  NOSH_HIP(hipMemcpyAsync(nrm.data(), norms.p, sizeof(double) * args.ninst,
                          hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  for (double& v : nrm) v = std::sqrt(v);
}

void Batch::relayout(int threads, int fold_) {
  block_threads = threads;
  fold = fold_;
  args.fold = fold_;
  int64_t b0 = 0, maxb = 0;
  for (auto& gh : ghost) {
    BGroup& G = gh.dev;
    G.block0 = b0;
    const int rows = threads / G.bp;
    G.nblocks = (G.n + rows - 1) / rows;
    G.nblocks = (G.nblocks + fold - 1) / fold * fold;
    b0 += G.nblocks;
    maxb = std::max(maxb, G.nblocks);
  }
  args.total_blocks = b0;
  args.maxblocks = maxb;
  partials.alloc(size_t(args.ninst) * maxb + 64);
  args.partials = partials.p;
  std::vector<BGroup> gv;
  for (auto& gh : ghost) gv.push_back(gh.dev);
  groups.alloc(gv.size());
  NOSH_HIP(hipMemcpy(groups.p, gv.data(), sizeof(BGroup) * gv.size(), hipMemcpyHostToDevice));
  args.groups = groups.p;
  drop_graph();
}

Batch::~Batch() { drop_graph(); }

}  // namespace nosh
