// eigen.hip -- LOBPCG for the Jacobian's extreme eigenpairs (the stability
// eigenvalues of a continuation branch), with block vector kernels.
//
// Replaces the Anasazi eigensolve of executables/nosh-eig/nosh-eig.cpp:
// 212-333 (LOBPCG on the symmetric W = J(psi) with W_prec as
// preconditioner). The method is Knyazev's locally optimal block
// preconditioned conjugate gradient (SIAM J. Sci. Comput. 23 (2001)) with
// the basis kept orthonormal by SVQB (Stathopoulos & Wu, SIAM J. Sci.
// Comput. 23 (2002)): per iteration the search space is S = [X, W, P]
// (Ritz vectors, preconditioned residuals, previous directions), the
// Rayleigh-Ritz problem S^T J S is solved on the host, and J S is carried
// along by the same linear combinations so that each iteration costs one
// J apply and one preconditioner apply per block column.
//
// Vectors live on the device as columns of 2n doubles (internal vertex
// order); a block is a list of column pointers. The dot-product kernel
// reads each column once per 4-column tile pair and writes per-block
// partial sums that a second kernel adds in fixed order (deterministic).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "eigen.hpp"
#include "minres_device.hpp"

namespace nosh {

namespace {

struct ColsC {
  const double* p[kMaxCols];
};
struct ColsW {
  double* p[kMaxCols];
};

constexpr int kDotThreads = 256;
constexpr int kCombThreads = 256;

// Partial sums of A_i . B_j over one row range for one 4 x 4 tile of (i, j).
// One 1-D grid of nblk * ntiles workgroups; the logical index (XCD-remapped,
// so that consecutive logical indices run on one XCD) is rb * ntiles + tile:
// the tiles of one row range run side by side and re-read its columns from
// that XCD's L2 rather than from HBM.
__global__ __launch_bounds__(kDotThreads) void k_block_dots(ColsC A, int ma, ColsC B, int mb,
                                                            int64_t L2, int64_t per_block,
                                                            int nblk,
                                                            double* __restrict__ partials) {
  const int ntj = (mb + 3) >> 2;
  const int ntiles = ((ma + 3) >> 2) * ntj;
  const int64_t logical = xcd_block(true);
  const int tile = int(logical % ntiles);
  const int64_t rb = logical / ntiles;
  const int ti = tile / ntj, tj = tile % ntj;
  const int na = min(4, ma - 4 * ti), nb = min(4, mb - 4 * tj);
  const double2* a[4];
  const double2* b[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a[q] = reinterpret_cast<const double2*>(A.p[4 * ti + (q < na ? q : 0)]);
    b[q] = reinterpret_cast<const double2*>(B.p[4 * tj + (q < nb ? q : 0)]);
  }
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  const int64_t r0 = rb * per_block;
  const int64_t r1 = min(L2, r0 + per_block);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kDotThreads) {
    double2 av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = a[q][r];
      bv[q] = b[q][r];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[4 * i + j] += av[i].x * bv[j].x + av[i].y * bv[j].y;
  }
  __shared__ double lds[16 * 16];
  block_sum<16>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) partials[(int64_t(tile) * 16 + q) * nblk + rb] = acc[q];
  }
}

// out[i mb + j] = sum of the partials of (i, j) in block order; one block
// per (tile, q).
__global__ __launch_bounds__(kFinalizeThreads) void k_block_dots_final(
    const double* __restrict__ partials, int nblk, int ma, int mb, double* __restrict__ out) {
  const int tile = blockIdx.x >> 4, q = blockIdx.x & 15;
  const int ntj = (mb + 3) >> 2;
  const int i = 4 * (tile / ntj) + (q >> 2), j = 4 * (tile % ntj) + (q & 3);
  double v[1];
  finalize_sum<1>(v, partials + (int64_t(tile) * 16 + q) * nblk, nblk);
  if (threadIdx.x == 0 && i < ma && j < mb) out[i * mb + j] = v[0];
}

// Coefficients passed by value (kernel arguments) when m k fits.
constexpr int kArgCoefs = 256;
struct CoefArg {
  double c[kArgCoefs];
};

// Y_j = beta Y_j + sum_c C[c k + j] S_c, two rows (one double2) per thread.
// C is the by-value CA.c when BYVAL, else the device array Cd.
template <bool BYVAL>
__global__ __launch_bounds__(kCombThreads) void k_block_combine(ColsC S, int m, CoefArg CA,
                                                                const double* __restrict__ Cd,
                                                                ColsW Y, int k, double beta,
                                                                int64_t L2) {
  const double* C = BYVAL ? CA.c : Cd;
  const int64_t r = int64_t(blockIdx.x) * kCombThreads + threadIdx.x;
  if (r >= L2) return;
  double2 acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = make_double2(0.0, 0.0);
  for (int c = 0; c < m; ++c) {
    const double2 v = reinterpret_cast<const double2*>(S.p[c])[r];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j < k) {
        const double w = C[c * k + j];
        acc[j].x += w * v.x;
        acc[j].y += w * v.y;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j < k) {
      double2* y = reinterpret_cast<double2*>(Y.p[j]) + r;
      double2 o = acc[j];
      if (beta != 0.0) {
        const double2 yo = *y;
        o.x += beta * yo.x;
        o.y += beta * yo.y;
      }
      *y = o;
    }
  }
}

// ---- Multi-column operators: J or P's off-diagonal part read once for up to
// kBV columns (each slot's column index and value shared by all of them).
constexpr int kBV = 8;
constexpr int kSB = 256;

struct BlkOp {
  int64_t n;
  SellView sv;          // split off-diagonal layout
  const double2* val;   // off-diagonal values of K (shared by J and P)
  const double* blk;    // 3 n per-vertex blocks (J or P), diagonal of K included
  const double* bound;  // Chebyshev bound from the rebuild
  double upper, ratio;
  int iters;
};

// Off-diagonal products of one row for nv <= NV columns.
template <int NV>
__device__ __forceinline__ void blk_row(const BlkOp& o, int64_t row, int nv, const ColsC& X,
                                        double2 (&acc)[NV]) {
  const int32_t sl = uniform(int32_t(row >> 6));
  const int lane = int(row & 63);
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = make_double2(0.0, 0.0);
  // Upper part: values in place.
  for (int32_t j = uniform(o.sv.uptr[sl]) + lane, e = uniform(o.sv.uptr[sl + 1]); j < e;
       j += kSliceH) {
    const int32_t c = __builtin_nontemporal_load(o.sv.ucol + j);
    const double2 v = o.val[j];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      if (q < nv) {
        const double2 xv = reinterpret_cast<const double2*>(X.p[q])[c];
        acc[q].x += v.x * xv.x - v.y * xv.y;
        acc[q].y += v.x * xv.y + v.y * xv.x;
      }
    }
  }
  // Lower part: the conjugate of the mirrored slot's value.
  for (int32_t j = uniform(o.sv.lptr[sl]) + lane, e = uniform(o.sv.lptr[sl + 1]); j < e;
       j += kSliceH) {
    const int32_t c = __builtin_nontemporal_load(o.sv.lcol + j);
    const double2 v = o.val[__builtin_nontemporal_load(o.sv.lpos + j)];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      if (q < nv) {
        const double2 xv = reinterpret_cast<const double2*>(X.p[q])[c];
        acc[q].x += v.x * xv.x + v.y * xv.y;
        acc[q].y += v.x * xv.y - v.y * xv.x;
      }
    }
  }
}

// Y_q = Koff X_q + B X_q.
template <int NV>
__global__ __launch_bounds__(kSB) void k_blk_apply(BlkOp o, int nv, ColsC X, ColsW Y) {
  const int64_t row = xcd_block(true) * kSB + threadIdx.x;
  if (row >= o.n) return;
  double2 acc[NV];
  blk_row<NV>(o, row, nv, X, acc);
  const double b0 = o.blk[row], b1 = o.blk[o.n + row], b2 = o.blk[2 * o.n + row];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q < nv) {
      const double2 xi = reinterpret_cast<const double2*>(X.p[q])[row];
      reinterpret_cast<double2*>(Y.p[q])[row] =
          make_double2(acc[q].x + b0 * xi.x + b2 * xi.y, acc[q].y + b2 * xi.x + b1 * xi.y);
    }
  }
}

// Chebyshev step 0 on P: Z_q = D^{-1} R_q / theta (and the direction D_q).
template <int NV>
__global__ __launch_bounds__(kSB) void k_blk_cheb_first(BlkOp o, int nv, ColsC R, ColsW Z,
                                                        ColsW Dd) {
  const int64_t row = int64_t(blockIdx.x) * kSB + threadIdx.x;
  if (row >= o.n) return;
  const double hi = cheb_hi(o.bound, o.upper);
  const double th = 0.5 * (hi + hi / o.ratio);
  const double b0 = o.blk[row], b1 = o.blk[o.n + row], b2 = o.blk[2 * o.n + row];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q < nv) {
      const double2 rv = reinterpret_cast<const double2*>(R.p[q])[row];
      double z0, z1;
      dinv<double>(b0, b1, b2, rv.x, rv.y, &z0, &z1);
      const double2 z = make_double2(z0 / th, z1 / th);
      reinterpret_cast<double2*>(Z.p[q])[row] = z;
      if (o.iters > 1) reinterpret_cast<double2*>(Dd.p[q])[row] = z;
    }
  }
}

// Chebyshev step j >= 1: Zout = Zin + d, d = c1 d + c2 D^{-1}(R - P Zin).
template <int NV>
__global__ __launch_bounds__(kSB) void k_blk_cheb_step(BlkOp o, int nv, int j, ColsC R, ColsC Zin,
                                                       ColsW Zout, ColsW Dd) {
  const int64_t row = xcd_block(true) * kSB + threadIdx.x;
  if (row >= o.n) return;
  const Cheb c = cheb_coef(cheb_hi(o.bound, o.upper), o.ratio, j);
  const bool last = j == o.iters - 1;
  double2 acc[NV];
  blk_row<NV>(o, row, nv, Zin, acc);
  const double b0 = o.blk[row], b1 = o.blk[o.n + row], b2 = o.blk[2 * o.n + row];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q < nv) {
      const double2 zi = reinterpret_cast<const double2*>(Zin.p[q])[row];
      const double2 rv = reinterpret_cast<const double2*>(R.p[q])[row];
      const double res0 = rv.x - (acc[q].x + b0 * zi.x + b2 * zi.y);
      const double res1 = rv.y - (acc[q].y + b2 * zi.x + b1 * zi.y);
      double u0, u1;
      dinv<double>(b0, b1, b2, res0, res1, &u0, &u1);
      double2* dp = reinterpret_cast<double2*>(Dd.p[q]) + row;
      double2 d = *dp;
      d.x = c.c1 * d.x + c.c2 * u0;
      d.y = c.c1 * d.y + c.c2 * u1;
      if (!last) *dp = d;
      reinterpret_cast<double2*>(Zout.p[q])[row] = make_double2(zi.x + d.x, zi.y + d.y);
    }
  }
}

template <class C, class V>
C cols_of(const V& v, size_t from, int cnt) {
  C c{};
  for (int q = 0; q < cnt; ++q) c.p[q] = v[from + q];
  return c;
}

// Y = J X (or P X) for any number of columns, kBV at a time.
void blk_apply(const BlkOp& o, const std::vector<double*>& X, const std::vector<double*>& Y,
               hipStream_t s) {
  for (size_t f = 0; f < X.size(); f += kBV) {
    const int nv = int(std::min<size_t>(kBV, X.size() - f));
    k_blk_apply<kBV><<<grid_for(o.n, kSB), kSB, 0, s>>>(o, nv, cols_of<ColsC>(X, f, nv),
                                                        cols_of<ColsW>(Y, f, nv));
  }
}

// Z = Chebyshev(P) R for every column; T and Dd are scratch columns.
void blk_cheb(const BlkOp& o, const std::vector<double*>& R, const std::vector<double*>& Z,
              const std::vector<double*>& T, const std::vector<double*>& Dd, hipStream_t s) {
  const unsigned nb = grid_for(o.n, kSB);
  const int steps = o.iters < 1 ? 1 : o.iters;
  for (size_t f = 0; f < R.size(); f += kBV) {
    const int nv = int(std::min<size_t>(kBV, R.size() - f));
    const ColsC r = cols_of<ColsC>(R, f, nv);
    const ColsW d = cols_of<ColsW>(Dd, f, nv);
    const ColsW zb[2] = {cols_of<ColsW>(Z, f, nv), cols_of<ColsW>(T, f, nv)};
    int cur = (steps - 1) % 2;
    k_blk_cheb_first<kBV><<<nb, kSB, 0, s>>>(o, nv, r, zb[cur], d);
    for (int j = 1; j < steps; ++j) {
      ColsC zin{};
      for (int q = 0; q < nv; ++q) zin.p[q] = zb[cur].p[q];
      k_blk_cheb_step<kBV><<<nb, kSB, 0, s>>>(o, nv, j, r, zin, zb[1 - cur], d);
      cur = 1 - cur;
    }
  }
}

// SplitMix64 uniform(-1, 1) values, index i of the stream seed.
__global__ void k_random_block(double* __restrict__ x, int64_t count, uint64_t seed) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t z = seed + uint64_t(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  x[i] = 2.0 * (double(z >> 11) * 0x1.0p-53) - 1.0;
}

// q = i psi (interleaved: (Re, Im) -> (-Im, Re)).
__global__ void k_times_i(int64_t n, const double2* __restrict__ psi, double2* __restrict__ q) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 v = psi[i];
    q[i] = make_double2(-v.y, v.x);
  }
}

// Column pool over one allocation of ncols columns of L doubles.
struct Pool {
  DBuf<double> buf;
  std::vector<double*> free;
  void init(int ncols, int64_t L) {
    buf.alloc(size_t(ncols) * L);
    free.clear();
    for (int c = ncols - 1; c >= 0; --c) free.push_back(buf.p + int64_t(c) * L);
  }
  std::vector<double*> take(int k) {
    NOSH_CHECK(int(free.size()) >= k, kInvalid, "eigensolver column pool exhausted");
    std::vector<double*> v(free.end() - k, free.end());
    free.resize(free.size() - k);
    return v;
  }
  void give(std::vector<double*>& v) {
    free.insert(free.end(), v.begin(), v.end());
    v.clear();
  }
};

std::vector<const double*> cview(const std::vector<double*>& a) {
  return std::vector<const double*>(a.begin(), a.end());
}

std::vector<const double*> cat(std::initializer_list<const std::vector<double*>*> parts) {
  std::vector<const double*> r;
  for (auto* p : parts) r.insert(r.end(), p->begin(), p->end());
  return r;
}

}  // namespace

void block_dots(const double* const* A, int ma, const double* const* B, int mb, int64_t L,
                BlockWork& w, double* out, hipStream_t s) {
  NOSH_CHECK(ma >= 1 && mb >= 1 && ma <= kMaxCols && mb <= kMaxCols && L % 2 == 0, kInvalid,
             "block_dots: bad block shape");
  const int64_t L2 = L / 2;
  const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>(1024, L2 / (kDotThreads * 8)));
  const int64_t per = (L2 + nblk - 1) / nblk;
  const int ntiles = ((ma + 3) / 4) * ((mb + 3) / 4);
  const size_t need = size_t(ntiles) * 16 * nblk;
  if (w.partials.n < need) w.partials.alloc(need);
  if (w.out.n < size_t(ma * mb)) w.out.alloc(kMaxCols * kMaxCols);
  ColsC a{}, b{};
  for (int i = 0; i < ma; ++i) a.p[i] = A[i];
  for (int j = 0; j < mb; ++j) b.p[j] = B[j];
  k_block_dots<<<unsigned(nblk * ntiles), kDotThreads, 0, s>>>(a, ma, b, mb, L2, per, int(nblk),
                                                               w.partials.p);
  k_block_dots_final<<<unsigned(ntiles * 16), kFinalizeThreads, 0, s>>>(w.partials.p, int(nblk),
                                                                       ma, mb, w.out.p);
  NOSH_HIP(hipMemcpyAsync(out, w.out.p, sizeof(double) * ma * mb, hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
}

void block_combine(const double* const* S, int m, const double* C, double* const* Y, int k,
                   double beta, int64_t L, BlockWork& w, hipStream_t s) {
  NOSH_CHECK(m >= 1 && m <= kMaxCols && k >= 1 && k <= 16 && L % 2 == 0, kInvalid,
             "block_combine: bad block shape");
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < m; ++i)
      NOSH_CHECK(Y[j] != S[i], kInvalid, "block_combine: output aliases an input");
  ColsC sc{};
  ColsW yc{};
  for (int i = 0; i < m; ++i) sc.p[i] = S[i];
  for (int j = 0; j < k; ++j) yc.p[j] = Y[j];
  const int64_t L2 = L / 2;
  const unsigned grid = grid_for(L2, kCombThreads);
  if (m * k <= kArgCoefs) {
    CoefArg ca;
    std::memcpy(ca.c, C, sizeof(double) * m * k);
    k_block_combine<true><<<grid, kCombThreads, 0, s>>>(sc, m, ca, nullptr, yc, k, beta, L2);
    return;
  }
  if (w.coef.n < size_t(kMaxCols * 16)) w.coef.alloc(kMaxCols * 16);
  // Synchronous upload: the host array may change right after this call.
  NOSH_HIP(hipStreamSynchronize(s));
  NOSH_HIP(hipMemcpy(w.coef.p, C, sizeof(double) * m * k, hipMemcpyHostToDevice));
  k_block_combine<false><<<grid, kCombThreads, 0, s>>>(sc, m, CoefArg{}, w.coef.p, yc, k, beta,
                                                       L2);
}

void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  // Cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 8.5):
  // each rotation zeroes A_pq with t = tan(phi) the smaller root of
  // t^2 + 2 t zeta - 1 = 0, zeta = (A_qq - A_pp) / (2 A_pq).
  V.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) V[size_t(i) * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a2 = A[size_t(i) * n + j] * A[size_t(i) * n + j];
        tot += a2;
        if (i != j) off += a2;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[size_t(p) * n + q];
        if (std::fabs(apq) <= 1e-300) continue;
        const double zeta = (A[size_t(q) * n + q] - A[size_t(p) * n + p]) / (2.0 * apq);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A[size_t(k) * n + p], akq = A[size_t(k) * n + q];
          A[size_t(k) * n + p] = c * akp - sn * akq;
          A[size_t(k) * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A[size_t(p) * n + k], aqk = A[size_t(q) * n + k];
          A[size_t(p) * n + k] = c * apk - sn * aqk;
          A[size_t(q) * n + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[size_t(k) * n + p], vkq = V[size_t(k) * n + q];
          V[size_t(k) * n + p] = c * vkp - sn * vkq;
          V[size_t(k) * n + q] = sn * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(),
            [&](int a, int b) { return A[size_t(a) * n + a] < A[size_t(b) * n + b]; });
  std::vector<double> V2(size_t(n) * n);
  w.resize(n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[size_t(idx[j]) * n + idx[j]];
    for (int i = 0; i < n; ++i) V2[size_t(i) * n + j] = V[size_t(i) * n + idx[j]];
  }
  V.swap(V2);
}

namespace {

struct Lobpcg {
  MeshCache& m;
  hipStream_t s;
  int64_t L;
  Pool pool;
  BlockWork bw;
  const double* q = nullptr;  // unit gauge direction or null

  Lobpcg(MeshCache& mm, hipStream_t ss) : m(mm), s(ss), L(2 * mm.n) {}

  std::vector<double> dots(const std::vector<const double*>& A,
                           const std::vector<const double*>& B) {
    std::vector<double> r(A.size() * B.size());
    block_dots(A.data(), int(A.size()), B.data(), int(B.size()), L, bw, r.data(), s);
    return r;
  }
  // Y -= Q (Q^T Y) for orthonormal Q; tracked: AY -= AQ (same coefficients).
  void project(const std::vector<const double*>& Q, std::vector<double*>& Y,
               const std::vector<const double*>* AQ = nullptr, std::vector<double*>* AY = nullptr) {
    if (Q.empty() || Y.empty()) return;
    std::vector<double> c = dots(Q, cview(Y));
    for (double& v : c) v = -v;
    block_combine(Q.data(), int(Q.size()), c.data(), Y.data(), int(Y.size()), 1.0, L, bw, s);
    if (AQ && AY)
      block_combine(AQ->data(), int(AQ->size()), c.data(), AY->data(), int(AY->size()), 1.0, L,
                    bw, s);
  }
  // Replaces V (and AV) by an orthonormal basis of span(V) (SVQB); returns
  // the number of columns kept.
  int svqb(std::vector<double*>& V, std::vector<double*>* AV = nullptr) {
    const int mv = int(V.size());
    if (mv == 0) return 0;
    std::vector<double> G = dots(cview(V), cview(V));
    double dmax = 0.0;
    for (int i = 0; i < mv; ++i) dmax = std::max(dmax, G[size_t(i) * mv + i]);
    std::vector<int> keep;
    for (int i = 0; i < mv; ++i)
      if (G[size_t(i) * mv + i] > 1e-28 * dmax && G[size_t(i) * mv + i] > 1e-300) keep.push_back(i);
    const int mk = int(keep.size());
    if (mk == 0) {
      pool.give(V);
      if (AV) pool.give(*AV);
      return 0;
    }
    std::vector<double> D(mk), H(size_t(mk) * mk);
    for (int a = 0; a < mk; ++a) D[a] = 1.0 / std::sqrt(G[size_t(keep[a]) * mv + keep[a]]);
    for (int a = 0; a < mk; ++a)
      for (int b = 0; b < mk; ++b)
        H[size_t(a) * mk + b] = D[a] * G[size_t(keep[a]) * mv + keep[b]] * D[b];
    std::vector<double> th, U;
    sym_eig(mk, H, th, U);
    const double tmax = th.back();
    std::vector<int> cols;
    for (int j = 0; j < mk; ++j)
      if (th[j] > 1e-13 * tmax) cols.push_back(j);
    const int nk = int(cols.size());
    std::vector<double> T(size_t(mv) * nk, 0.0);
    for (int a = 0; a < mk; ++a)
      for (int c = 0; c < nk; ++c)
        T[size_t(keep[a]) * nk + c] = D[a] * U[size_t(a) * mk + cols[c]] / std::sqrt(th[cols[c]]);
    std::vector<double*> nv = pool.take(nk);
    auto cv = cview(V);
    block_combine(cv.data(), mv, T.data(), nv.data(), nk, 0.0, L, bw, s);
    pool.give(V);
    V = nv;
    if (AV) {
      std::vector<double*> na = pool.take(nk);
      auto ca = cview(*AV);
      block_combine(ca.data(), mv, T.data(), na.data(), nk, 0.0, L, bw, s);
      pool.give(*AV);
      *AV = na;
    }
    return nk;
  }
};

}  // namespace

int lobpcg_jacobian(nosh_nls_s* nls, const double* psi_int, const std::map<std::string, double>& p,
                    const nosh_eigen_options& o, double* evals, DBuf<double>* evecs,
                    nosh_eigen_stats* st, hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  MeshCache& m = *nls->mesh->m;
  const int64_t n = m.n, L = 2 * n;
  const int nev = o.nev;
  NOSH_CHECK(nev >= 1 && nev <= 16, kInvalid, "nev must be in 1..16");
  const int k = std::min(16, std::max(nev, o.block_size));
  NOSH_CHECK(o.max_iters >= 1 && o.tol > 0.0, kInvalid, "bad eigensolver options");
  NOSH_CHECK(o.which == 0 || o.which == 1, kInvalid, "which must be 0 or 1");

  // J(psi) and P(psi) at p.
  const double g = need(p, "g");
  need(p, "mu");
  nls->keo->set_parameters(p, s);
  compute_ct(m, nls->thickness, nls->ct, p, s);
  const double* V = nls->V->sf.get_v(p, s);
  if (nls->jblk.n != size_t(3 * n)) {
    nls->jblk.alloc(3 * n);
    nls->pblk.alloc(3 * n);
  }
  if (nls->bound.n < 1) nls->bound.alloc(1);
  launch_rebuild_blocks(m, nls->keo->kv, nls->ct.p, V, g, psi_int, nls->jblk.p, nls->pblk.p,
                        nls->bound.p, s);
  int applies = 0;
  BlkOp jop{n, sell_view(m), nls->keo->kv.val.p, nls->jblk.p, nls->bound.p, 0.0,
            o.cheb_ratio, o.cheb_iters};
  BlkOp pop = jop;
  pop.blk = nls->pblk.p;
  auto applyJ = [&](const std::vector<double*>& X, std::vector<double*>& Y) {
    blk_apply(jop, X, Y, s);
    applies += int(X.size());
  };
  const bool prec = o.use_prec && o.cheb_iters > 0;

  Lobpcg lb(m, s);
  lb.pool.init(10 * k + 1, L);
  // Gauge direction.
  std::vector<double*> qv;
  if (o.deflate_gauge) {
    qv = lb.pool.take(1);
    k_times_i<<<grid_for(n, 256), 256, 0, s>>>(n, reinterpret_cast<const double2*>(psi_int),
                                                reinterpret_cast<double2*>(qv[0]));
    if (lb.svqb(qv) == 1)
      lb.q = qv[0];
  }
  const std::vector<const double*> Q = lb.q ? std::vector<const double*>{lb.q}
                                            : std::vector<const double*>{};

  // Start block.
  std::vector<double*> X = lb.pool.take(k), AX, W, AW, Pd, AP;
  for (int j = 0; j < k; ++j)
    k_random_block<<<grid_for(L, 256), 256, 0, s>>>(X[j], L, o.seed + 0x51ED270B27ull * (j + 1));
  lb.project(Q, X);
  NOSH_CHECK(lb.svqb(X) == k, kInvalid, "eigensolver start block is rank deficient");
  AX = lb.pool.take(k);
  applyJ(X, AX);
  std::vector<double> theta(k);
  double scale = 0.0;
  // Rayleigh-Ritz on span(X).
  {
    std::vector<double> G = lb.dots(cview(X), cview(AX));
    for (int a = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b)
        G[size_t(a) * k + b] = G[size_t(b) * k + a] = 0.5 * (G[size_t(a) * k + b] + G[size_t(b) * k + a]);
    std::vector<double> th, U;
    sym_eig(k, G, th, U);
    if (o.which == 1) {
      std::reverse(th.begin(), th.end());
      for (int a = 0; a < k; ++a) std::reverse(U.begin() + size_t(a) * k, U.begin() + size_t(a + 1) * k);
    }
    std::vector<double*> nx = lb.pool.take(k), nax = lb.pool.take(k);
    auto cx = cview(X), cax = cview(AX);
    block_combine(cx.data(), k, U.data(), nx.data(), k, 0.0, L, lb.bw, s);
    block_combine(cax.data(), k, U.data(), nax.data(), k, 0.0, L, lb.bw, s);
    lb.pool.give(X);
    lb.pool.give(AX);
    X = nx;
    AX = nax;
    theta = th;
    for (double t : th) scale = std::max(scale, std::fabs(t));
  }

  int it = 0, conv = 0;
  double maxres = 0.0;
  bool fresh = true;  // AX recomputed by J applies since the last update
  std::vector<double> rn(k);
  for (;; ++it) {
    // Residuals R = AX - X diag(theta) into W.
    W = lb.pool.take(k);
    {
      std::vector<const double*> S = cat({&AX, &X});
      std::vector<double> C(size_t(2 * k) * k, 0.0);
      for (int j = 0; j < k; ++j) {
        C[size_t(j) * k + j] = 1.0;
        C[size_t(k + j) * k + j] = -theta[j];
      }
      block_combine(S.data(), 2 * k, C.data(), W.data(), k, 0.0, L, lb.bw, s);
      std::vector<double> G = lb.dots(cview(W), cview(W));
      for (int j = 0; j < k; ++j) rn[j] = std::sqrt(std::max(0.0, G[size_t(j) * k + j]));
    }
    conv = 0;
    maxres = 0.0;
    for (int j = 0; j < nev; ++j) {
      const double r = rn[j] / scale;
      maxres = std::max(maxres, r);
      if (r <= o.tol) ++conv;
    }
    if (conv == nev && !fresh) {
      // Confirm with J applied afresh (the carried AX accumulates rounding).
      lb.pool.give(W);
      applyJ(X, AX);
      fresh = true;
      --it;
      continue;
    }
    if (conv == nev || it >= o.max_iters) {
      lb.pool.give(W);
      break;
    }
    fresh = false;
    // Preconditioned residuals.
    if (prec) {
      std::vector<double*> Z = lb.pool.take(k), T = lb.pool.take(k), Dd = lb.pool.take(k);
      blk_cheb(pop, W, Z, T, Dd, s);
      lb.pool.give(T);
      lb.pool.give(Dd);
      lb.pool.give(W);
      W = Z;
    }
    const auto cX = cview(X);
    const auto cAX = cview(AX);
    for (int rep = 0; rep < 2; ++rep) {
      lb.project(Q, W);
      lb.project(cX, W);
    }
    lb.svqb(W);
    if (W.empty()) break;  // the residuals vanished to rounding
    AW = lb.pool.take(int(W.size()));
    applyJ(W, AW);
    if (!Pd.empty()) {
      const auto cW = cview(W), cAW = cview(AW);
      for (int rep = 0; rep < 2; ++rep) {
        lb.project(cX, Pd, &cAX, &AP);
        lb.project(cW, Pd, &cAW, &AP);
      }
      lb.svqb(Pd, &AP);
    }
    // Rayleigh-Ritz on S = [X, W, P].
    const std::vector<const double*> S = cat({&X, &W, &Pd}), AS = cat({&AX, &AW, &AP});
    const int ms = int(S.size());
    std::vector<double> G = lb.dots(S, AS);
    for (int a = 0; a < ms; ++a)
      for (int b = a + 1; b < ms; ++b)
        G[size_t(a) * ms + b] = G[size_t(b) * ms + a] =
            0.5 * (G[size_t(a) * ms + b] + G[size_t(b) * ms + a]);
    std::vector<double> th, U;
    sym_eig(ms, G, th, U);
    for (double t : th) scale = std::max(scale, std::fabs(t));
    std::vector<double> Cx(size_t(ms) * k);
    for (int j = 0; j < k; ++j) {
      const int src = o.which == 0 ? j : ms - 1 - j;
      theta[j] = th[src];
      for (int a = 0; a < ms; ++a) Cx[size_t(a) * k + j] = U[size_t(a) * ms + src];
    }
    std::vector<double*> nX = lb.pool.take(k), nAX = lb.pool.take(k), nP = lb.pool.take(k),
                        nAP = lb.pool.take(k);
    // New X and new directions P (the components of the new X along W and
    // P) in one pass over S, likewise AX and AP over AS.
    if (2 * k <= 16) {
      std::vector<double> C2(size_t(ms) * 2 * k, 0.0);
      for (int a = 0; a < ms; ++a)
        for (int j = 0; j < k; ++j) {
          C2[size_t(a) * 2 * k + j] = Cx[size_t(a) * k + j];
          if (a >= k) C2[size_t(a) * 2 * k + k + j] = Cx[size_t(a) * k + j];
        }
      std::vector<double*> Y(nX);
      Y.insert(Y.end(), nP.begin(), nP.end());
      block_combine(S.data(), ms, C2.data(), Y.data(), 2 * k, 0.0, L, lb.bw, s);
      Y.assign(nAX.begin(), nAX.end());
      Y.insert(Y.end(), nAP.begin(), nAP.end());
      block_combine(AS.data(), ms, C2.data(), Y.data(), 2 * k, 0.0, L, lb.bw, s);
    } else {
      block_combine(S.data(), ms, Cx.data(), nX.data(), k, 0.0, L, lb.bw, s);
      block_combine(AS.data(), ms, Cx.data(), nAX.data(), k, 0.0, L, lb.bw, s);
      block_combine(S.data() + k, ms - k, Cx.data() + size_t(k) * k, nP.data(), k, 0.0, L,
                    lb.bw, s);
      block_combine(AS.data() + k, ms - k, Cx.data() + size_t(k) * k, nAP.data(), k, 0.0, L,
                    lb.bw, s);
    }
    lb.pool.give(X);
    lb.pool.give(AX);
    lb.pool.give(W);
    lb.pool.give(AW);
    lb.pool.give(Pd);
    lb.pool.give(AP);
    X = nX;
    AX = nAX;
    Pd = nP;
    AP = nAP;
    lb.project(Q, X);
  }
  for (int j = 0; j < nev; ++j) evals[j] = theta[j];
  if (evecs) {
    if (evecs->n != size_t(nev) * L) evecs->alloc(size_t(nev) * L);
    for (int j = 0; j < nev; ++j)
      NOSH_HIP(hipMemcpyAsync(evecs->p + int64_t(j) * L, X[j], sizeof(double) * L,
                              hipMemcpyDeviceToDevice, s));
  }
  NOSH_HIP(hipStreamSynchronize(s));
  if (st) {
    st->iters = it;
    st->converged = conv;
    st->applies = applies;
    st->max_residual = maxres;
    st->scale = scale;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return conv;
}

}  // namespace nosh

// ------------------------------------------------------------------ C ABI

using namespace nosh;

namespace {
template <class F>
int eguard(F&& f) {
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
}  // namespace

extern "C" {

int nosh_eigen_default_options(nosh_eigen_options* o) {
  return eguard([&] {
    NOSH_CHECK(o, kInvalid, "null options");
    o->nev = 4;
    o->block_size = 6;
    o->max_iters = 500;
    o->tol = 1e-8;
    o->which = 0;
    o->deflate_gauge = 1;
    o->use_prec = 1;
    o->cheb_iters = 6;
    o->cheb_ratio = 300.0;
    o->seed = 0;
  });
}

int nosh_nls_jacobian_eigenpairs(nosh_nls_t nls, const double* psi, int np,
                                 const char* const* names, const double* values,
                                 const nosh_eigen_options* opts, double* evals, double* evecs,
                                 nosh_eigen_stats* stats) {
  int conv = 0, nev = 0;
  int rc = eguard([&] {
    NOSH_CHECK(nls && psi && evals, kInvalid, "null argument");
    nosh_eigen_options o;
    if (opts)
      o = *opts;
    else
      nosh_eigen_default_options(&o);
    nev = o.nev;
    std::map<std::string, double> p;
    for (int i = 0; i < np; ++i) p[names[i]] = values[i];
    MeshCache& m = *nls->mesh->m;
    const int64_t n = m.n;
    hipStream_t s = default_stream();
    nls->st.ensure(2 * n);
    nls->st.a.upload(psi, 2 * n, s);
    launch_permute_in(n, m.perm.p, nls->st.a.p, nls->st.b.p, s);
    DBuf<double> psi_int(2 * n);
    NOSH_HIP(hipMemcpyAsync(psi_int.p, nls->st.b.p, sizeof(double) * 2 * n,
                            hipMemcpyDeviceToDevice, s));
    DBuf<double> vec;
    conv = lobpcg_jacobian(nls, psi_int.p, p, o, evals, evecs ? &vec : nullptr, stats, s);
    if (evecs) {
      for (int j = 0; j < o.nev; ++j) {
        launch_permute_out(n, m.perm.p, vec.p + int64_t(j) * 2 * n, nls->st.d.p, s);
        nls->st.d.download(evecs + int64_t(j) * 2 * n, 2 * n, s);
      }
      NOSH_HIP(hipStreamSynchronize(s));
    }
    if (conv < o.nev) throw nosh::Error(kNotConverged, "eigensolver: not all pairs converged");
  });
  return rc;
}

}  // extern "C"
