// mesh_cache.hip -- builds the edge/covolume cache of a triangle or
// tetrahedron mesh on the device.
//
// Reference behaviour reproduced (paths relative to the nosh tree):
//  * edge creation and numbering: MOAB get_adjacencies(cells, 1, create=true)
//    at src/mesh_reader.cpp:117-125 and src/mesh.cpp:629-691 numbers edges in
//    order of first appearance over the cells in file order, each cell's
//    local edges in MOAB canonical order (tri 01 12 20; tet 01 12 20 03 13 23);
//    edge endpoints are (lower id, higher id) (src/mesh.cpp:670-677) and each
//    cell's edges are listed in ascending edge id (src/mesh.cpp:650-663).
//  * edge coefficients: per cell, the (3x3 | 6x6) system
//    A_ij = (e_i.e_j)^2, rhs_i = |T| |e_i|^2 solved with complete pivoting
//    (src/mesh_tri.cpp:105-151, src/mesh_tetra.cpp:105-160), covolume_e +=
//    c_{T,e} * length_e over cells in order (src/mesh_tri.cpp:77-100).
//  * control volumes: circumcentric splitting per cell
//    (src/mesh.cpp:895-1027, src/mesh_tetra.cpp:194-415).
//
// Device algorithm: a stable radix sort of the (lower, higher) keys of all
// T*ne cell-edge slots groups each edge's occurrences in cell order; run
// heads give first appearances, and a scan over first-appearance flags in
// slot order numbers the edges. Per-edge and per-vertex sums are segmented
// sums over those runs in cell order (deterministic, no float atomics).
// Reverse Cuthill-McKee runs on the host over the device-built adjacency.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>
#include <queue>

#include "internal.hpp"

namespace nosh {
namespace {

constexpr int kBlock = 256;

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(const double* c, int64_t v) {
  return {c[3 * v], c[3 * v + 1], c[3 * v + 2]};
}
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double nrm(D3 a) { return sqrt(dot(a, a)); }

// Local edge l of a cell: (0,1) (1,2) (2,0) then (0,3) (1,3) (2,3).
__device__ __forceinline__ void local_edge(int l, int* a, int* b) {
  if (l < 3) {
    *a = l;
    *b = l == 2 ? 0 : l + 1;
  } else {
    *a = l - 3;
    *b = 3;
  }
}

__global__ void k_edge_keys(const int32_t* __restrict__ cells, int64_t T, int nv, int ne,
                            uint64_t n, uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= T * ne) return;
  const int64_t t = i / ne;
  const int l = int(i - t * ne);
  int la, lb;
  local_edge(l, &la, &lb);
  uint32_t a = uint32_t(cells[t * nv + la]), b = uint32_t(cells[t * nv + lb]);
  if (a > b) {
    const uint32_t tmp = a;
    a = b;
    b = tmp;
  }
  keys[i] = uint64_t(a) * n + uint64_t(b);
  pos[i] = uint32_t(i);
}

template <class K>
__global__ void k_heads(const K* __restrict__ keys, int64_t N, int32_t* __restrict__ head) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_mark_first(const int32_t* __restrict__ head, const uint32_t* __restrict__ spos,
                             int64_t N, int32_t* __restrict__ isfirst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (head[i]) isfirst[spos[i]] = 1;
}

// For every run head: the edge id of its first slot, the endpoints, the
// start of the run in sorted order.
__global__ void k_run_edges(const int32_t* __restrict__ head, const int32_t* __restrict__ runid,
                            const uint32_t* __restrict__ spos, const uint64_t* __restrict__ skeys,
                            const int32_t* __restrict__ eid_at_pos, int64_t N, uint64_t n,
                            int32_t* __restrict__ run_edge, int32_t* __restrict__ edges,
                            int32_t* __restrict__ edge_start) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N || !head[i]) return;
  const int32_t e = eid_at_pos[spos[i]];
  run_edge[runid[i]] = e;
  edges[2 * e] = int32_t(skeys[i] / n);
  edges[2 * e + 1] = int32_t(skeys[i] % n);
  edge_start[e] = int32_t(i);
}

__global__ void k_cell_edges_canon(const int32_t* __restrict__ runid,
                                   const uint32_t* __restrict__ spos,
                                   const int32_t* __restrict__ run_edge, int64_t N,
                                   int32_t* __restrict__ ce) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  ce[spos[i]] = run_edge[runid[i]];
}

__global__ void k_sort_cell_edges(int32_t* __restrict__ ce, int64_t T, int ne) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= T) return;
  int32_t v[6];
#pragma unroll
  for (int l = 0; l < 6; ++l) v[l] = l < ne ? ce[t * ne + l] : 0x7fffffff;
  // Insertion sort (ne <= 6).
#pragma unroll
  for (int i = 1; i < 6; ++i)
#pragma unroll
    for (int j = i; j > 0; --j)
      if (v[j] < v[j - 1]) {
        const int32_t tmp = v[j];
        v[j] = v[j - 1];
        v[j - 1] = tmp;
      }
  for (int l = 0; l < ne; ++l) ce[t * ne + l] = v[l];
}

__global__ void k_edge_length(const double* __restrict__ X, const int32_t* __restrict__ edges,
                              int64_t E, double* __restrict__ len) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  len[e] = nrm(d3(X, edges[2 * e]) - d3(X, edges[2 * e + 1]));
}

// Gaussian elimination with complete pivoting on an m x m system (m <= 6).
template <int M>
__device__ void solve_full_pivot(double (&A)[M][M], double (&b)[M], double (&x)[M]) {
  int colperm[M];
#pragma unroll
  for (int i = 0; i < M; ++i) colperm[i] = i;
  for (int k = 0; k < M; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int i = k; i < M; ++i)
      for (int j = k; j < M; ++j)
        if (fabs(A[i][j]) > best) {
          best = fabs(A[i][j]);
          pr = i;
          pc = j;
        }
    if (pr != k) {
      for (int j = 0; j < M; ++j) {
        const double t = A[k][j];
        A[k][j] = A[pr][j];
        A[pr][j] = t;
      }
      const double t = b[k];
      b[k] = b[pr];
      b[pr] = t;
    }
    if (pc != k) {
      for (int i = 0; i < M; ++i) {
        const double t = A[i][k];
        A[i][k] = A[i][pc];
        A[i][pc] = t;
      }
      const int t = colperm[k];
      colperm[k] = colperm[pc];
      colperm[pc] = t;
    }
    const double piv = A[k][k];
    for (int i = k + 1; i < M; ++i) {
      const double f = A[i][k] / piv;
      A[i][k] = 0.0;
      for (int j = k + 1; j < M; ++j) A[i][j] -= f * A[k][j];
      b[i] -= f * b[k];
    }
  }
  double y[M];
  for (int k = M - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < M; ++j) s -= A[k][j] * y[j];
    y[k] = s / A[k][k];
  }
  for (int k = 0; k < M; ++k) x[colperm[k]] = y[k];
}

// Edge coefficients of every cell, in ascending-edge-id order.
template <int NE>
__global__ void k_edge_coefficients(const double* __restrict__ X,
                                    const int32_t* __restrict__ edges,
                                    const int32_t* __restrict__ ce, int64_t T,
                                    double* __restrict__ coef) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= T) return;
  D3 e[NE];
#pragma unroll
  for (int l = 0; l < NE; ++l) {
    const int32_t k = ce[t * NE + l];
    e[l] = d3(X, edges[2 * k]) - d3(X, edges[2 * k + 1]);
  }
  double vol;
  if (NE == 3) {
    vol = 0.5 * nrm(cross(e[0], e[1]));
  } else {
    double al = dot(e[0], cross(e[1], e[2]));
    if (fabs(al) / nrm(e[0]) / nrm(e[1]) / nrm(e[2]) < 1.0e-5)
      al = dot(e[0], cross(e[1], e[NE > 3 ? 3 : 2]));
    vol = fabs(al) / 6.0;
  }
  double A[NE][NE], rhs[NE], c[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const double a = dot(e[i], e[i]);
    rhs[i] = vol * a;
    A[i][i] = a * a;
#pragma unroll
    for (int j = i + 1; j < NE; ++j) {
      const double v = dot(e[i], e[j]) * dot(e[j], e[i]);
      A[i][j] = v;
      A[j][i] = v;
    }
  }
  solve_full_pivot<NE>(A, rhs, c);
#pragma unroll
  for (int l = 0; l < NE; ++l) coef[t * NE + l] = c[l];
}

// covolume_e = sum over the cells containing e (in cell order) of c_{T,e} len_e:
// Per-run segmented sum, using run starts (runs numbered in sorted order).
__global__ void k_covolume_runs(const int32_t* __restrict__ run_start,
                                const int32_t* __restrict__ run_edge,
                                const uint32_t* __restrict__ spos, const int32_t* __restrict__ ce,
                                const double* __restrict__ coef, const double* __restrict__ len,
                                int64_t U, int ne, double* __restrict__ covol) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= U) return;
  const int32_t e = run_edge[r];
  const double le = len[e];
  double acc = 0.0;
  for (int64_t i = run_start[r]; i < run_start[r + 1]; ++i) {
    const int64_t t = spos[i] / ne;
    int l = 0;
    for (int q = 0; q < ne; ++q)
      if (ce[t * ne + q] == e) l = q;
    acc += coef[t * ne + l] * le;
  }
  covol[e] = acc;
}

__global__ void k_run_start(const int32_t* __restrict__ head, const int32_t* __restrict__ runid,
                            int64_t N, int32_t* __restrict__ run_start) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (head[i]) run_start[runid[i]] = int32_t(i);
}

__device__ D3 tri_circumcenter(D3 n0, D3 n1, D3 n2) {
  const D3 a = n0 - n1, b = n1 - n2, c = n2 - n0;
  const D3 axb = cross(a, b);
  const double omega = 2.0 * dot(axb, axb);
  const double al = -dot(b, b) * dot(a, c) / omega;
  const double be = -dot(c, c) * dot(b, a) / omega;
  const double ga = -dot(a, a) * dot(c, b) / omega;
  return al * n0 + be * n1 + ga * n2;
}

// Per-cell contributions to the control volumes of its vertices.
__global__ void k_cell_cv(const double* __restrict__ X, const int32_t* __restrict__ cells,
                          int64_t T, int nv, double* __restrict__ cvc) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= T) return;
  D3 x[4];
  for (int i = 0; i < nv; ++i) x[i] = d3(X, cells[t * nv + i]);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (nv == 3) {
    const D3 cc = tri_circumcenter(x[0], x[1], x[2]);
    for (int e0 = 0; e0 < 3; ++e0)
      for (int e1 = e0 + 1; e1 < 3; ++e1) {
        const int other = 3 - e0 - e1;
        const double el = nrm(x[e1] - x[e0]);
        const D3 mp = 0.5 * (x[e0] + x[e1]);
        const double coedge = nrm(mp - cc);
        const D3 cell_normal = cross(x[other] - x[e0], mp - x[e0]);
        const D3 cc_normal = cross(cc - x[e0], mp - x[e0]);
        const double cov = copysign(coedge, dot(cc_normal, cell_normal));
        const double pyr = 0.5 * el * cov / 2;
        acc[e0] += pyr;
        acc[e1] += pyr;
      }
  } else {
    const D3 r0 = x[1] - x[0], r1 = x[2] - x[0], r2 = x[3] - x[0];
    const double omega = 2.0 * dot(r0, cross(r1, r2));
    const double al = dot(r0, r0) / omega, be = dot(r1, r1) / omega, ga = dot(r2, r2) / omega;
    const D3 cc = x[0] + al * cross(r1, r2) + be * cross(r2, r0) + ga * cross(r0, r1);
    for (int e0 = 0; e0 < 4; ++e0)
      for (int e1 = e0 + 1; e1 < 4; ++e1) {
        int o[2], no = 0;
        for (int q = 0; q < 4; ++q)
          if (q != e0 && q != e1) o[no++] = q;
        const D3 x0 = x[e0], x1 = x[e1], o0 = x[o[0]], o1 = x[o[1]];
        const double el = nrm(x1 - x0);
        const D3 mp = 0.5 * (x0 + x1);
        const D3 ccf0 = tri_circumcenter(x0, x1, o0);
        const D3 ccf1 = tri_circumcenter(x0, x1, o1);
        const D3 gauge = cross(o0 - mp, o1 - mp);
        double cov = 0.0;
        const double a0 = 0.5 * nrm(mp - ccf0) * nrm(ccf0 - cc);
        cov += copysign(a0, dot(cross(ccf0 - mp, cc - mp), gauge));
        const double a1 = 0.5 * nrm(mp - ccf1) * nrm(ccf1 - cc);
        cov += copysign(a1, dot(cross(cc - mp, ccf1 - mp), gauge));
        const double pyr = 0.5 * el * cov / 3;
        acc[e0] += pyr;
        acc[e1] += pyr;
      }
  }
  for (int i = 0; i < nv; ++i) cvc[t * nv + i] = acc[i];
}

__global__ void k_add(int32_t* __restrict__ v, int64_t N, int32_t d) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) v[i] += d;
}

__global__ void k_iota(uint32_t* __restrict__ v, int64_t N) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) v[i] = uint32_t(i);
}

__global__ void k_u32_from_i32(const int32_t* __restrict__ a, int64_t N, uint32_t* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) b[i] = uint32_t(a[i]);
}

// start[v] = first index i with key[i] >= v, for v in [0, n]; keys sorted.
__global__ void k_segment_starts(const uint32_t* __restrict__ keys, int64_t N, int64_t n,
                                 int32_t* __restrict__ start) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > N) return;
  const int64_t prev = i == 0 ? -1 : int64_t(keys[i - 1]);
  const int64_t cur = i == N ? n : int64_t(keys[i]);
  for (int64_t v = prev + 1; v <= cur; ++v) start[v] = int32_t(i);
}

__global__ void k_vertex_cv(const int32_t* __restrict__ start, const uint32_t* __restrict__ pos,
                            const double* __restrict__ cvc, int64_t n, double* __restrict__ cv) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double acc = 0.0;
  for (int32_t i = start[v]; i < start[v + 1]; ++i) acc += cvc[pos[i]];
  cv[v] = acc;
}

// Directed adjacency entries: key (row, col) in the chosen numbering, value 2e|dir.
__global__ void k_directed(const int32_t* __restrict__ edges, int64_t E,
                           const int32_t* __restrict__ iperm, uint64_t n,
                           uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint64_t a = uint64_t(edges[2 * e]), b = uint64_t(edges[2 * e + 1]);
  if (iperm) {
    a = uint64_t(iperm[a]);
    b = uint64_t(iperm[b]);
  }
  keys[2 * e] = a * n + b;      // row = lower endpoint
  vals[2 * e] = int32_t(2 * e);
  keys[2 * e + 1] = b * n + a;  // row = higher endpoint
  vals[2 * e + 1] = int32_t(2 * e + 1);
}

__global__ void k_rows_of(const uint64_t* __restrict__ keys, int64_t N, uint64_t n,
                          uint32_t* __restrict__ rows, int32_t* __restrict__ cols) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  rows[i] = uint32_t(keys[i] / n);
  cols[i] = int32_t(keys[i] % n);
}

__global__ void k_slice_width(const int32_t* __restrict__ rowstart, int64_t n, int64_t nslices,
                              int32_t* __restrict__ width64) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int w = 0;
  for (int64_t r = s * kSliceH; r < std::min<int64_t>(n, (s + 1) * kSliceH); ++r)
    w = max(w, rowstart[r + 1] - rowstart[r]);
  width64[s] = w * kSliceH;
}

__global__ void k_sell_pad(const int32_t* __restrict__ slice_ptr, int64_t nslices, int64_t n,
                           int32_t* __restrict__ scol, int32_t* __restrict__ sedge) {
  const int64_t s = blockIdx.x;
  if (s >= nslices) return;
  for (int64_t j = slice_ptr[s] + threadIdx.x; j < slice_ptr[s + 1]; j += blockDim.x) {
    const int64_t r = s * kSliceH + ((j - slice_ptr[s]) & (kSliceH - 1));
    scol[j] = int32_t(r < n ? r : 0);
    sedge[j] = -1;
  }
}

__global__ void k_sell_fill(const uint32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                            const int32_t* __restrict__ vals, const int32_t* __restrict__ rowstart,
                            const int32_t* __restrict__ slice_ptr, int64_t N,
                            int32_t* __restrict__ scol, int32_t* __restrict__ sedge) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t r = rows[i];
  const int64_t k = i - rowstart[r];
  const int64_t dst = slice_ptr[r / kSliceH] + k * kSliceH + (r % kSliceH);
  scol[dst] = cols[i];
  sedge[dst] = vals[i];
}

// Split layout (see MeshCache): per row, the number of columns below the row
// (CSR columns are sorted, so these come first) and above it.
__global__ void k_count_split(const int32_t* __restrict__ rowstart, const int32_t* __restrict__ cols,
                              int64_t n, int32_t* __restrict__ nlow, int32_t* __restrict__ nup) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int32_t k = rowstart[r];
  const int32_t e = rowstart[r + 1];
  while (k < e && cols[k] < r) ++k;
  nlow[r] = k - rowstart[r];
  nup[r] = e - k;
}

__global__ void k_split_width(const int32_t* __restrict__ nlow, const int32_t* __restrict__ nup,
                              int64_t n, int64_t nslices, int32_t* __restrict__ wl64,
                              int32_t* __restrict__ wu64) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int wl = 0, wu = 0;
  for (int64_t r = s * kSliceH; r < std::min<int64_t>(n, (s + 1) * kSliceH); ++r) {
    wl = max(wl, nlow[r]);
    wu = max(wu, nup[r]);
  }
  wl64[s] = wl * kSliceH;
  wu64[s] = wu * kSliceH;
}

// Padding of both parts: column = own row; upper edge code -1 (value 0),
// lower position = zpos (the zero entry after the upper values).
__global__ void k_split_pad(const int32_t* __restrict__ uptr, const int32_t* __restrict__ lptr,
                            int64_t nslices, int64_t n, int32_t zpos, int32_t* __restrict__ ucol,
                            int32_t* __restrict__ uedge, int32_t* __restrict__ lcol,
                            int32_t* __restrict__ lpos) {
  const int64_t s = blockIdx.x;
  if (s >= nslices) return;
  for (int64_t j = uptr[s] + threadIdx.x; j < uptr[s + 1]; j += blockDim.x) {
    const int64_t r = s * kSliceH + ((j - uptr[s]) & (kSliceH - 1));
    ucol[j] = int32_t(r < n ? r : 0);
    uedge[j] = -1;
  }
  for (int64_t j = lptr[s] + threadIdx.x; j < lptr[s + 1]; j += blockDim.x) {
    const int64_t r = s * kSliceH + ((j - lptr[s]) & (kSliceH - 1));
    lcol[j] = int32_t(r < n ? r : 0);
    lpos[j] = zpos;
  }
}

// Every CSR entry to its part: upper entries (column above the row) keep the
// edge code and record where the edge's value will live; lower entries
// keep the code for now (k_split_lpos turns it into that position).
__global__ void k_split_fill(const uint32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                             const int32_t* __restrict__ codes,
                             const int32_t* __restrict__ rowstart,
                             const int32_t* __restrict__ nlow, const int32_t* __restrict__ uptr,
                             const int32_t* __restrict__ lptr, int64_t N,
                             int32_t* __restrict__ ucol, int32_t* __restrict__ uedge,
                             int32_t* __restrict__ lcol, int32_t* __restrict__ lcode,
                             int32_t* __restrict__ upos_of_edge) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t r = rows[i];
  const int64_t k = i - rowstart[r];
  const int64_t s = r / kSliceH, lane = r % kSliceH;
  const int32_t c = cols[i], code = codes[i];
  if (c > r) {
    const int64_t dst = uptr[s] + (k - nlow[r]) * kSliceH + lane;
    ucol[dst] = c;
    uedge[dst] = code;
    upos_of_edge[code >> 1] = int32_t(dst);
  } else {
    const int64_t dst = lptr[s] + k * kSliceH + lane;
    lcol[dst] = c;
    lcode[dst] = code;
  }
}

__global__ void k_split_lpos(const int32_t* __restrict__ lcode,
                             const int32_t* __restrict__ upos_of_edge, int64_t nnz_l,
                             int32_t* __restrict__ lpos) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= nnz_l) return;
  const int32_t code = lcode[j];
  if (code >= 0) lpos[j] = upos_of_edge[code >> 1];
}

// Column offsets (column - row) of one part as 16-bit integers, one block per
// slice; rows past n keep offset 0 (never read). Sets *wide when an offset
// does not fit.
__global__ void k_col_off16(const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            int64_t nslices, int64_t n, int16_t* __restrict__ off,
                            int32_t* __restrict__ wide) {
  const int64_t s = blockIdx.x;
  if (s >= nslices) return;
  for (int64_t j = ptr[s] + threadIdx.x; j < ptr[s + 1]; j += blockDim.x) {
    const int64_t r = s * kSliceH + ((j - ptr[s]) & (kSliceH - 1));
    const int64_t d = r < n ? int64_t(col[j]) - r : 0;
    if (d < -32768 || d > 32767) {
      *wide = 1;
      off[j] = 0;
    } else {
      off[j] = int16_t(d);
    }
  }
}

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_gather_f64(const int32_t* __restrict__ perm, const double* __restrict__ src,
                             int64_t n, double* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_invert_perm(const int32_t* __restrict__ perm, int64_t n,
                              int32_t* __restrict__ iperm) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) iperm[perm[i]] = int32_t(i);
}

template <class K, class V>
void sort_pairs(const K* kin, K* kout, const V* vin, V* vout, int64_t N, int end_bit,
                hipStream_t s) {
  size_t tmp = 0;
  NOSH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, int(N), 0,
                                              end_bit, s));
  DBuf<char> t(tmp + 16);
  NOSH_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, kin, kout, vin, vout, int(N), 0,
                                              end_bit, s));
  NOSH_HIP(hipStreamSynchronize(s));
}

void inclusive_sum(const int32_t* in, int32_t* out, int64_t N, hipStream_t s) {
  size_t tmp = 0;
  NOSH_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, int(N), s));
  DBuf<char> t(tmp + 16);
  NOSH_HIP(hipcub::DeviceScan::InclusiveSum(t.p, tmp, in, out, int(N), s));
  NOSH_HIP(hipStreamSynchronize(s));
}

void exclusive_sum(const int32_t* in, int32_t* out, int64_t N, hipStream_t s) {
  size_t tmp = 0;
  NOSH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, int(N), s));
  DBuf<char> t(tmp + 16);
  NOSH_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, int(N), s));
  NOSH_HIP(hipStreamSynchronize(s));
}

int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// Reverse Cuthill-McKee over a symmetric CSR graph (host). Start vertices
// are pseudo-peripheral (George-Liu); neighbours are visited by ascending
// degree, ties by ascending vertex id; components in order of their
// lowest-numbered vertex.
std::vector<int32_t> rcm_order(int64_t n, const std::vector<int32_t>& ptr,
                               const std::vector<int32_t>& adj) {
  std::vector<int32_t> deg(n);
  for (int64_t v = 0; v < n; ++v) deg[v] = ptr[v + 1] - ptr[v];
  std::vector<char> done(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> level(n, -1), touched;
  auto bfs_last_level = [&](int32_t root, int* ecc) {
    // Returns the min-degree vertex of the last BFS level from root.
    touched.clear();
    std::vector<int32_t> frontier{root}, next;
    level[root] = 0;
    touched.push_back(root);
    int depth = 0;
    std::vector<int32_t> last = frontier;
    while (!frontier.empty()) {
      next.clear();
      for (int32_t v : frontier)
        for (int32_t q = ptr[v]; q < ptr[v + 1]; ++q) {
          const int32_t w = adj[q];
          if (level[w] < 0) {
            level[w] = depth + 1;
            touched.push_back(w);
            next.push_back(w);
          }
        }
      if (next.empty()) break;
      ++depth;
      last = next;
      frontier.swap(next);
    }
    for (int32_t v : touched) level[v] = -1;
    *ecc = depth;
    int32_t best = last[0];
    for (int32_t v : last)
      if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
    return best;
  };
  std::vector<int32_t> nb;
  for (int64_t seed = 0; seed < n; ++seed) {
    if (done[seed]) continue;
    int32_t root = int32_t(seed);
    int ecc = 0;
    int32_t cand = bfs_last_level(root, &ecc);
    for (int iter = 0; iter < 8; ++iter) {
      int ecc2 = 0;
      const int32_t cand2 = bfs_last_level(cand, &ecc2);
      if (ecc2 <= ecc) {
        root = cand;
        break;
      }
      root = cand;
      ecc = ecc2;
      cand = cand2;
    }
    size_t head = order.size();
    order.push_back(root);
    done[root] = 1;
    while (head < order.size()) {
      const int32_t v = order[head++];
      nb.clear();
      for (int32_t q = ptr[v]; q < ptr[v + 1]; ++q)
        if (!done[adj[q]]) nb.push_back(adj[q]);
      std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
        return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
      });
      for (int32_t w : nb) {
        done[w] = 1;
        order.push_back(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

}  // namespace

std::unique_ptr<MeshCache> build_mesh_cache(const double* coords, int64_t n, const int32_t* cells,
                                            int64_t T, int nv, bool reorder, hipStream_t s) {
  NOSH_CHECK(nv == 3 || nv == 4, kInvalid, "vertices_per_cell must be 3 or 4");
  NOSH_CHECK(n > 0 && T > 0, kInvalid, "empty mesh");
  const int ne = nv == 3 ? 3 : 6;
  NOSH_CHECK(T * ne < (int64_t(1) << 31) && n < (int64_t(1) << 31), kInvalid,
             "mesh too large for 32-bit slot indices");
  for (int64_t i = 0; i < T * nv; ++i)
    NOSH_CHECK(cells[i] >= 0 && cells[i] < n, kInvalid, "cell vertex index out of range");
  auto m = std::make_unique<MeshCache>();
  m->nv = nv;
  m->ne = ne;
  m->n = n;
  m->T = T;
  m->coords.alloc(3 * n);
  m->coords.upload(coords, 3 * n, s);
  m->cells.alloc(T * nv);
  m->cells.upload(cells, T * nv, s);

  const int64_t N = T * ne;
  const unsigned gN = grid_for(N, kBlock);
  // 1. Sort cell-edge slots by (lower, higher) key, stably.
  DBuf<uint64_t> keys(N), skeys(N);
  DBuf<uint32_t> pos(N), spos(N);
  k_edge_keys<<<gN, kBlock, 0, s>>>(m->cells.p, T, nv, ne, uint64_t(n), keys.p, pos.p);
  const int kbits = bits_for(uint64_t(n) * uint64_t(n));
  sort_pairs(keys.p, skeys.p, pos.p, spos.p, N, kbits, s);
  keys.release();
  pos.release();
  // 2. Runs and first appearances.
  DBuf<int32_t> head(N), runid(N), isfirst(N), eid(N);
  k_heads<<<gN, kBlock, 0, s>>>(skeys.p, N, head.p);
  // runid[i] = (number of run heads at or before i) - 1.
  inclusive_sum(head.p, runid.p, N, s);
  int32_t nruns = 0;
  NOSH_HIP(hipMemcpy(&nruns, runid.p + N - 1, 4, hipMemcpyDeviceToHost));
  k_add<<<gN, kBlock, 0, s>>>(runid.p, N, -1);
  const int64_t E = nruns;
  m->E = E;
  isfirst.zero(s);
  k_mark_first<<<gN, kBlock, 0, s>>>(head.p, spos.p, N, isfirst.p);
  exclusive_sum(isfirst.p, eid.p, N, s);
  isfirst.release();
  DBuf<int32_t> run_edge(E), edge_start(E), run_start(E + 1);
  m->edges.alloc(2 * E);
  k_run_edges<<<gN, kBlock, 0, s>>>(head.p, runid.p, spos.p, skeys.p, eid.p, N, uint64_t(n),
                                    run_edge.p, m->edges.p, edge_start.p);
  k_run_start<<<gN, kBlock, 0, s>>>(head.p, runid.p, N, run_start.p);
  const int32_t Ni = int32_t(N);
  NOSH_HIP(hipMemcpyAsync(run_start.p + E, &Ni, 4, hipMemcpyHostToDevice, s));
  eid.release();
  skeys.release();
  // 3. Cell -> edges, ascending.
  m->cell_edges.alloc(N);
  k_cell_edges_canon<<<gN, kBlock, 0, s>>>(runid.p, spos.p, run_edge.p, N, m->cell_edges.p);
  k_sort_cell_edges<<<grid_for(T, kBlock), kBlock, 0, s>>>(m->cell_edges.p, T, ne);
  head.release();
  runid.release();
  // 4. Edge lengths, coefficients, covolumes.
  m->len.alloc(E);
  m->covol.alloc(E);
  k_edge_length<<<grid_for(E, kBlock), kBlock, 0, s>>>(m->coords.p, m->edges.p, E, m->len.p);
  DBuf<double> coef(N);
  if (ne == 3)
    k_edge_coefficients<3><<<grid_for(T, 128), 128, 0, s>>>(m->coords.p, m->edges.p,
                                                            m->cell_edges.p, T, coef.p);
  else
    k_edge_coefficients<6><<<grid_for(T, 128), 128, 0, s>>>(m->coords.p, m->edges.p,
                                                            m->cell_edges.p, T, coef.p);
  k_covolume_runs<<<grid_for(E, kBlock), kBlock, 0, s>>>(run_start.p, run_edge.p, spos.p,
                                                         m->cell_edges.p, coef.p, m->len.p, E, ne,
                                                         m->covol.p);
  NOSH_HIP(hipStreamSynchronize(s));
  coef.release();
  spos.release();
  run_edge.release();
  run_start.release();
  edge_start.release();
  // 5. Control volumes: per-cell contributions summed per vertex in cell order.
  {
    const int64_t Nv = T * nv;
    DBuf<double> cvc(Nv);
    k_cell_cv<<<grid_for(T, 128), 128, 0, s>>>(m->coords.p, m->cells.p, T, nv, cvc.p);
    DBuf<uint32_t> vk(Nv), svk(Nv), vp(Nv), svp(Nv);
    k_u32_from_i32<<<grid_for(Nv, kBlock), kBlock, 0, s>>>(m->cells.p, Nv, vk.p);
    k_iota<<<grid_for(Nv, kBlock), kBlock, 0, s>>>(vp.p, Nv);
    sort_pairs(vk.p, svk.p, vp.p, svp.p, Nv, bits_for(uint64_t(n)), s);
    DBuf<int32_t> vstart(n + 1);
    k_segment_starts<<<grid_for(Nv + 1, kBlock), kBlock, 0, s>>>(svk.p, Nv, n, vstart.p);
    m->cv.alloc(n);
    k_vertex_cv<<<grid_for(n, kBlock), kBlock, 0, s>>>(vstart.p, svp.p, cvc.p, n, m->cv.p);
    NOSH_HIP(hipStreamSynchronize(s));
  }
  // 6. Vertex order: reverse Cuthill-McKee over the reference-order graph.
  const int64_t ND = 2 * E;
  std::vector<int32_t> perm_h(n);
  if (reorder) {
    DBuf<uint64_t> dk(ND), sdk(ND);
    DBuf<int32_t> dv(ND), sdv(ND);
    k_directed<<<grid_for(E, kBlock), kBlock, 0, s>>>(m->edges.p, E, nullptr, uint64_t(n), dk.p,
                                                      dv.p);
    sort_pairs(dk.p, sdk.p, dv.p, sdv.p, ND, kbits, s);
    DBuf<uint32_t> rows(ND);
    DBuf<int32_t> cols(ND), rstart(n + 1);
    k_rows_of<<<grid_for(ND, kBlock), kBlock, 0, s>>>(sdk.p, ND, uint64_t(n), rows.p, cols.p);
    k_segment_starts<<<grid_for(ND + 1, kBlock), kBlock, 0, s>>>(rows.p, ND, n, rstart.p);
    std::vector<int32_t> ptr_h(n + 1), adj_h(ND);
    NOSH_HIP(hipMemcpyAsync(ptr_h.data(), rstart.p, 4 * (n + 1), hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipMemcpyAsync(adj_h.data(), cols.p, 4 * ND, hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipStreamSynchronize(s));
    perm_h = rcm_order(n, ptr_h, adj_h);
  } else {
    std::iota(perm_h.begin(), perm_h.end(), 0);
  }
  m->perm.alloc(n);
  m->perm.upload(perm_h.data(), n, s);
  m->iperm.alloc(n);
  k_invert_perm<<<grid_for(n, kBlock), kBlock, 0, s>>>(m->perm.p, n, m->iperm.p);
  m->cv_int.alloc(n);
  k_gather_f64<<<grid_for(n, kBlock), kBlock, 0, s>>>(m->perm.p, m->cv.p, n, m->cv_int.p);
  // 7. Internal sliced ELLPACK of the off-diagonal graph.
  {
    DBuf<uint64_t> dk(ND), sdk(ND);
    DBuf<int32_t> dv(ND), sdv(ND);
    k_directed<<<grid_for(E, kBlock), kBlock, 0, s>>>(m->edges.p, E, m->iperm.p, uint64_t(n),
                                                      dk.p, dv.p);
    sort_pairs(dk.p, sdk.p, dv.p, sdv.p, ND, kbits, s);
    DBuf<uint32_t> rows(ND);
    DBuf<int32_t> cols(ND), rstart(n + 1);
    k_rows_of<<<grid_for(ND, kBlock), kBlock, 0, s>>>(sdk.p, ND, uint64_t(n), rows.p, cols.p);
    k_segment_starts<<<grid_for(ND + 1, kBlock), kBlock, 0, s>>>(rows.p, ND, n, rstart.p);
    const int64_t ns = (n + kSliceH - 1) / kSliceH;
    m->nslices = ns;
    DBuf<int32_t> w64(ns + 1);
    w64.zero(s);
    k_slice_width<<<grid_for(ns, kBlock), kBlock, 0, s>>>(rstart.p, n, ns, w64.p);
    m->slice_ptr.alloc(ns + 1);
    exclusive_sum(w64.p, m->slice_ptr.p, ns + 1, s);
    int32_t total = 0;
    NOSH_HIP(hipMemcpy(&total, m->slice_ptr.p + ns, 4, hipMemcpyDeviceToHost));
    m->nnz_pad = total;
    std::vector<int32_t> wh(ns);
    NOSH_HIP(hipMemcpy(wh.data(), w64.p, 4 * ns, hipMemcpyDeviceToHost));
    m->max_width = 0;
    for (int32_t w : wh) m->max_width = std::max(m->max_width, w / kSliceH);
    m->scol.alloc(std::max<int64_t>(total, 1));
    m->sedge.alloc(std::max<int64_t>(total, 1));
    k_sell_pad<<<unsigned(ns), kBlock, 0, s>>>(m->slice_ptr.p, ns, n, m->scol.p, m->sedge.p);
    k_sell_fill<<<grid_for(ND, kBlock), kBlock, 0, s>>>(rows.p, cols.p, sdv.p, rstart.p,
                                                        m->slice_ptr.p, ND, m->scol.p,
                                                        m->sedge.p);
    // 8. The split layout of the same graph for the operator kernels.
    DBuf<int32_t> nlow(n), nup(n), wl64(ns + 1), wu64(ns + 1);
    wl64.zero(s);
    wu64.zero(s);
    k_count_split<<<grid_for(n, kBlock), kBlock, 0, s>>>(rstart.p, cols.p, n, nlow.p, nup.p);
    k_split_width<<<grid_for(ns, kBlock), kBlock, 0, s>>>(nlow.p, nup.p, n, ns, wl64.p, wu64.p);
    m->uptr.alloc(ns + 1);
    m->lptr.alloc(ns + 1);
    exclusive_sum(wu64.p, m->uptr.p, ns + 1, s);
    exclusive_sum(wl64.p, m->lptr.p, ns + 1, s);
    int32_t tu = 0, tl = 0;
    NOSH_HIP(hipMemcpy(&tu, m->uptr.p + ns, 4, hipMemcpyDeviceToHost));
    NOSH_HIP(hipMemcpy(&tl, m->lptr.p + ns, 4, hipMemcpyDeviceToHost));
    m->nnz_u = tu;
    m->nnz_l = tl;
    m->ucol.alloc(std::max<int64_t>(tu, 1));
    m->uedge.alloc(std::max<int64_t>(tu, 1));
    m->lcol.alloc(std::max<int64_t>(tl, 1));
    m->lpos.alloc(std::max<int64_t>(tl, 1));
    DBuf<int32_t> lcode(std::max<int64_t>(tl, 1)), upos(std::max<int64_t>(E, 1));
    k_fill_i32<<<grid_for(tl, kBlock), kBlock, 0, s>>>(lcode.p, tl, -1);
    k_split_pad<<<unsigned(ns), kBlock, 0, s>>>(m->uptr.p, m->lptr.p, ns, n, tu, m->ucol.p,
                                                m->uedge.p, m->lcol.p, m->lpos.p);
    k_split_fill<<<grid_for(ND, kBlock), kBlock, 0, s>>>(
        rows.p, cols.p, sdv.p, rstart.p, nlow.p, m->uptr.p, m->lptr.p, ND, m->ucol.p,
        m->uedge.p, m->lcol.p, lcode.p, upos.p);
    k_split_lpos<<<grid_for(tl, kBlock), kBlock, 0, s>>>(lcode.p, upos.p, tl, m->lpos.p);
    // 16-bit column offsets when the bandwidth allows (2 bytes less per slot
    // for every kernel that reads the split layout).
    m->uoff.alloc(std::max<int64_t>(tu, 1));
    m->loff.alloc(std::max<int64_t>(tl, 1));
    DBuf<int32_t> wide(1);
    k_fill_i32<<<1, 1, 0, s>>>(wide.p, 1, 0);
    k_col_off16<<<unsigned(ns), kBlock, 0, s>>>(m->uptr.p, m->ucol.p, ns, n, m->uoff.p, wide.p);
    k_col_off16<<<unsigned(ns), kBlock, 0, s>>>(m->lptr.p, m->lcol.p, ns, n, m->loff.p, wide.p);
    int32_t wide_h = 0;
    NOSH_HIP(hipMemcpyAsync(&wide_h, wide.p, 4, hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipStreamSynchronize(s));
    if (wide_h) {
      m->uoff.release();
      m->loff.release();
    }
  }
  return m;
}

}  // namespace nosh
