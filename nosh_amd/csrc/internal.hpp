// internal.hpp -- device-resident objects of the nosh_amd library.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace nosh {

// Rows per SELL slice: one wavefront, one row per lane.
constexpr int kSliceH = 64;

// Edge/covolume cache and the internal sparse layout of the complex graph.
//
// Reference-order arrays follow the caller's vertex and cell numbering.
// Internal arrays are in the bandwidth-reduced vertex order (perm).
// The off-diagonal structure of every operator is a sliced ELLPACK over the
// internal rows: slice s holds rows [64 s, 64 s + 64); entry k of row r is
// at slice_ptr[s] + 64 k + (r % 64). scol is the internal column, sedge the
// reference edge id e coded as 2e (row = lower endpoint) or 2e+1 (row =
// higher endpoint), -1 for padding (col = own row, value 0).
struct MeshCache {
  int nv = 3, ne = 3;
  int64_t n = 0, E = 0, T = 0;
  DBuf<double> coords;      // 3n
  DBuf<int32_t> cells;      // T nv
  DBuf<int32_t> edges;      // 2E
  DBuf<int32_t> cell_edges; // T ne
  DBuf<double> len, covol;  // E
  DBuf<double> cv;          // n, reference order
  DBuf<int32_t> perm;       // n: internal -> reference
  DBuf<int32_t> iperm;      // n: reference -> internal
  int64_t nslices = 0, nnz_pad = 0;
  int max_width = 0;
  DBuf<int32_t> slice_ptr;  // nslices + 1
  DBuf<int32_t> scol;       // nnz_pad
  DBuf<int32_t> sedge;      // nnz_pad
  DBuf<double> cv_int;      // n, internal order
  // Split layout of the same graph, read by the single-instance operator
  // kernels: each undirected edge appears in two rows, and K_ji = conj(K_ij).
  // The slot in the row with the smaller internal index (the "upper" part:
  // column above the row) holds the value; the slot in the other row (the
  // "lower" part) holds the position of that value. Each part is a sliced
  // ELLPACK with its own slice widths, so a wave reads either contiguous
  // values or contiguous positions. Padding: column = own row; upper edge
  // code -1 (value 0); lower position nnz_u (the zero entry after the
  // upper values).
  int64_t nnz_u = 0, nnz_l = 0;
  DBuf<int32_t> uptr, ucol, uedge;  // nslices + 1, nnz_u, nnz_u
  DBuf<int32_t> lptr, lcol, lpos;   // nslices + 1, nnz_l, nnz_l
  // The columns again as 16-bit offsets from the row (column - row), when
  // every one fits (after RCM the offsets are bounded by the bandwidth:
  // 9,507 for the 96^3 cube, 26,083 for 160^3); empty otherwise.
  DBuf<int16_t> uoff, loff;  // nnz_u, nnz_l
};

// Device view of the split layout (kernel argument); uoff / loff are null
// when the columns are read as 32-bit indices.
struct SellView {
  const int32_t* uptr;
  const int32_t* ucol;
  const int32_t* lptr;
  const int32_t* lcol;
  const int32_t* lpos;
  const int16_t* uoff;
  const int16_t* loff;
};
inline SellView sell_view(const MeshCache& m) {
  return SellView{m.uptr.p, m.ucol.p, m.lptr.p, m.lcol.p, m.lpos.p, m.uoff.p, m.loff.p};
}

std::unique_ptr<MeshCache> build_mesh_cache(const double* coords, int64_t n, const int32_t* cells,
                                            int64_t T, int nv, bool reorder, hipStream_t s);

// Scalar field on vertices (constant with optional parameter, or explicit).
// scalar_field::constant (src/scalar_field_constant.cpp:43-75): V = c, or
// c + params[param] when the field has a parameter and the map holds it.
// scalar_field::explicit_values (src/scalar_field_explicit_values.cpp:33-57):
// V = params.at("beta") * values (std::out_of_range without "beta"), with
// the scalar parameter {"beta": 1}.
struct ScalarField {
  MeshCache* mesh = nullptr;
  bool constant = true;
  double c = 0.0;
  std::string param;  // "" = none; "beta" for explicit values
  double param_init = 0.0;
  std::vector<double> values;  // explicit values, reference order (host copy)
  DBuf<double> base;           // explicit values, internal order (unscaled)
  DBuf<double> dev;            // internal order, current parameters
  double dev_value = 0.0;      // constant (or beta) currently applied in dev
  bool dev_valid = false;
  std::map<std::string, double> scalar_parameters() const;
  // Value at the given parameters, written to dev (internal order).
  const double* get_v(const std::map<std::string, double>& p, hipStream_t s);
  // dV/dp (get_dvdp): a constant (1 if `name` is a constant field's
  // parameter, else 0) and, for explicit values and "beta", the per-vertex
  // values (internal order; null otherwise).
  double dvdp_constant(const std::string& name) const { return (constant && name == param) ? 1.0 : 0.0; }
  const double* dvdp_values(const std::string& name, hipStream_t s);
};

// explicit_values magnetic potential: per-edge projection a_e / mu.
struct VectorField {
  MeshCache* mesh = nullptr;
  double mu = 0.0;
  DBuf<double> atilde;  // E, reference edge order: 0.5 (A_lo + A_hi).(x_lo - x_hi)
};

// Per-edge weights alpha_e = covolume/length * mean thickness, built once.
struct EdgeWeights {
  DBuf<double> alpha;  // E
};

template <class T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

// Assembled off-diagonal values of K (and of dK/dmu) in the split layout,
// with the row sums of alpha (the diagonal of K) and of |alpha|.
struct KeoValues {
  DBuf<double2> val;   // nnz_u + 1 (upper slots of the split layout, then 0)
  DBuf<double2> dval;  // nnz_u + 1 (dK/dp), allocated on demand
  DBuf<float2> val32;  // FP32 copy for the mixed-precision inner solver
  DBuf<double> diag;   // n: sum of alpha over the row
  DBuf<double> rowabs; // n: sum of |alpha| over the row
  double mu = 0.0;
  bool valid = false, dvalid = false, valid32 = false;
};

// Launchers (operators.hip).
void launch_atilde(const MeshCache& m, const double* A_dev, double* atilde, hipStream_t s);
void launch_alpha(const MeshCache& m, const double* thickness_ref, double* alpha, hipStream_t s);
void launch_keo_fill(const MeshCache& m, const double* alpha, const double* atilde, double mu,
                     bool dkeo, bool deriv_is_mu, KeoValues& kv, hipStream_t s);
void launch_keo_to_fp32(const MeshCache& m, KeoValues& kv, hipStream_t s);
void launch_permute_in(int64_t n, const int32_t* perm, const double* x_ref, double* x_int,
                       hipStream_t s);
void launch_permute_out(int64_t n, const int32_t* perm, const double* x_int, double* x_ref,
                        hipStream_t s);
void launch_gather_ref(int64_t n, const int32_t* perm, const double* v_ref, double* v_int,
                       hipStream_t s);

// y = K x (internal order).
void launch_keo_apply(const MeshCache& m, const KeoValues& kv, const double* x, double* y,
                      hipStream_t s);
// y = dK x + (is_g ? ct |x|^2 x : ct dvdp x)   (dF/dp column).
// y = dK/dp x + ct (is_g ? |x|^2 : dvdp + dvdp_vec[k]) x; dvdp_vec may be null.
void launch_dfdp(const MeshCache& m, const KeoValues& kv, const double* ct, const double* x,
                 bool is_g, double dvdp, const double* dvdp_vec, double* y, hipStream_t s);
// f = K x + ct (V + g |x|^2) x; optional ||f||^2 into *nrm2 (device scalar).
void launch_compute_f(const MeshCache& m, const KeoValues& kv, const double* ct,
                      const double* V, double g, const double* x, double* f, double* nrm2,
                      hipStream_t s);
// Per-vertex 2x2 blocks: J (with V) and P (keo_regularized), and the
// Chebyshev bound into *bound (device).
void launch_rebuild_blocks(const MeshCache& m, const KeoValues& kv, const double* ct,
                           const double* V, double g, const double* psi, double* jblk,
                           double* pblk, double* bound, hipStream_t s);
// y = Koff x + B x with 3-component blocks B (J or P). With t0/t1 set, the
// runtime stamps the kernel's own start and end into those events.
void launch_block_apply(const MeshCache& m, const KeoValues& kv, const double* blk,
                        const double* x, double* y, hipStream_t s, hipEvent_t t0 = nullptr,
                        hipEvent_t t1 = nullptr);

// Elementwise helpers.
void launch_fill(double* a, int64_t n, double v, hipStream_t s);
void launch_mul(const double* a, const double* b, double* c, int64_t n, hipStream_t s);
void launch_scale(const double* a, double v, double* c, int64_t n, hipStream_t s);
// y = K x (internal order), FP32 values and vectors.
void launch_block_apply32(const MeshCache& m, const KeoValues& kv, const float* blk,
                          const float* x, float* y, hipStream_t s, hipEvent_t t0 = nullptr,
                          hipEvent_t t1 = nullptr);
// One pass of ordinary (cache-allocating) 16-byte loads over `bytes` bytes of
// `buf`, one float per workgroup into out[grid]: replaces the contents of the
// L2s and the Infinity Cache with clean lines.
void launch_cache_flush(const void* buf, int64_t bytes, float* out, unsigned grid,
                        hipStream_t s);

}  // namespace nosh

namespace nosh {
// Control-volume weighted functionals of the model (internal order):
// b == nullptr: -sum c |a|^4 / sum c (Gibbs energy); else sum c Re(conj(a) b) / sum c.
// partials: >= 2 * ceil(n / 256) doubles.
void launch_cv_weighted(const MeshCache& m, const double* a, const double* b, double* partials,
                        double* out, hipStream_t s);
}  // namespace nosh
