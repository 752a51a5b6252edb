// capi.cpp -- the C ABI (include/nosh_amd.h) over device-resident C++
// objects that mirror the reference classes:
//   nosh_mesh_s   <- nosh::mesh (src/mesh.hpp), built on the device
//   nosh_vfield_s <- vector_field::explicit_values (src/vector_field_explicit_values.cpp)
//   nosh_sfield_s <- scalar_field::constant / explicit_values
//   nosh_keo_s    <- parameter_matrix::keo / DkeoDP
//   nosh_jac_s    <- jacobian_operator
//   nosh_prec_s   <- keo_regularized
//   nosh_nls_s    <- model_evaluator::nls
// plus the Newton-MINRES driver that replaces NOX/Belos as the caller.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "handles.hpp"

namespace nosh {

hipStream_t default_stream() {
  static hipStream_t s = [] {
    hipStream_t st = nullptr;
    NOSH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
  }();
  return s;
}

std::map<std::string, double> ScalarField::scalar_parameters() const {
  std::map<std::string, double> m;
  if (!param.empty()) m[param] = param_init;
  return m;
}

const double* ScalarField::get_v(const std::map<std::string, double>& p, hipStream_t s) {
  if (dev.n != size_t(mesh->n)) {
    dev.alloc(mesh->n);
    dev_valid = false;
  }
  if (constant) {
    double v = c;
    if (!param.empty()) {
      auto it = p.find(param);
      if (it != p.end()) v = c + it->second;
    }
    if (!dev_valid || v != dev_value) {
      launch_fill(dev.p, mesh->n, v, s);
      dev_value = v;
      dev_valid = true;
    }
  } else {
    // vals.scale(params.at("beta")) (src/scalar_field_explicit_values.cpp:36-43).
    const double beta = need(p, "beta");
    dvdp_values("beta", s);
    if (!dev_valid || beta != dev_value) {
      launch_scale(base.p, beta, dev.p, mesh->n, s);
      dev_value = beta;
      dev_valid = true;
    }
  }
  return dev.p;
}

const double* ScalarField::dvdp_values(const std::string& name, hipStream_t s) {
  if (constant || name != "beta") return nullptr;
  if (base.n != size_t(mesh->n)) {
    base.alloc(mesh->n);
    DBuf<double> tmp(mesh->n);
    tmp.upload(values.data(), mesh->n, s);
    launch_gather_ref(mesh->n, mesh->perm.p, tmp.p, base.p, s);
    NOSH_HIP(hipStreamSynchronize(s));
  }
  return base.p;
}

}  // namespace nosh

using namespace nosh;

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return NOSH_OK;
  } catch (const nosh::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::out_of_range& e) {
    g_err = e.what();
    return NOSH_ERR_MISSING_PARAM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NOSH_ERR_INVALID;
  } catch (...) {
    g_err = "unknown error";
    return NOSH_ERR_INVALID;
  }
}

std::map<std::string, double> to_map(int np, const char* const* names, const double* values) {
  NOSH_CHECK(np >= 0, kInvalid, "negative parameter count");
  NOSH_CHECK(np == 0 || (names && values), kInvalid, "null parameter arrays");
  std::map<std::string, double> m;
  for (int i = 0; i < np; ++i) {
    NOSH_CHECK(names[i], kInvalid, "null parameter name");
    m[names[i]] = values[i];
  }
  return m;
}

void check_apply_mode(int mode, double alpha, double beta) {
  NOSH_CHECK(mode == NOSH_NO_TRANS, kUnsupported, "Only untransposed applies supported.");
  NOSH_CHECK(alpha == 1.0, kUnsupported, "Only alpha==1.0 supported.");
  NOSH_CHECK(beta == 0.0, kUnsupported, "Only beta==0.0 supported.");
}

hipStream_t stream_of(void* s) { return s ? static_cast<hipStream_t>(s) : default_stream(); }

// Host columns -> device (internal order) -> op -> host columns.
template <class Op>
void host_apply(MeshCache& m, Staging& st, const double* x, double* y, int ncols, int64_t ld,
                Op op) {
  const int64_t len = 2 * m.n;
  NOSH_CHECK(x && y, kInvalid, "null vector");
  NOSH_CHECK(ncols >= 1, kInvalid, "ncols must be >= 1");
  NOSH_CHECK(ncols == 1 || ld >= len, kInvalid, "leading dimension smaller than 2n");
  hipStream_t s = default_stream();
  st.ensure(len);
  for (int c = 0; c < ncols; ++c) {
    st.a.upload(x + c * ld, len, s);
    launch_permute_in(m.n, m.perm.p, st.a.p, st.b.p, s);
    op(st.b.p, st.c.p, s);
    launch_permute_out(m.n, m.perm.p, st.c.p, st.d.p, s);
    st.d.download(y + c * ld, len, s);
  }
  NOSH_HIP(hipStreamSynchronize(s));
}

template <class Op>
void device_apply(MeshCache& m, Staging& st, const double* x, double* y, hipStream_t s, Op op) {
  NOSH_CHECK(x && y, kInvalid, "null vector");
  st.ensure(2 * m.n);
  launch_permute_in(m.n, m.perm.p, x, st.b.p, s);
  op(st.b.p, st.c.p, s);
  launch_permute_out(m.n, m.perm.p, st.c.p, y, s);
}

void upload_internal(MeshCache& m, Staging& st, const double* x_host, double* dst,
                     hipStream_t s) {
  st.ensure(2 * m.n);
  st.a.upload(x_host, 2 * m.n, s);
  launch_permute_in(m.n, m.perm.p, st.a.p, dst, s);
}

void download_internal(MeshCache& m, Staging& st, const double* src, double* y_host,
                       hipStream_t s) {
  st.ensure(2 * m.n);
  launch_permute_out(m.n, m.perm.p, src, st.d.p, s);
  st.d.download(y_host, 2 * m.n, s);
  NOSH_HIP(hipStreamSynchronize(s));
}

// thickness_->get_v(dummy) in reference order (for the alpha cache).
// The reference asks with an empty parameter map, so a constant field gives
// c and an explicit-values field throws (params.at("beta"),
// src/parameter_matrix_keo.cpp:202-203).
const double* thickness_ref(nosh_sfield_s* t, hipStream_t s) {
  MeshCache& m = *t->sf.mesh;
  if (!t->sf.constant) need({}, "beta");
  if (t->ref.n != size_t(m.n)) {
    t->ref.alloc(m.n);
    launch_fill(t->ref.p, m.n, t->sf.c, s);
  }
  return t->ref.p;
}

}  // namespace

namespace nosh {

void set_last_error(const std::string& msg) { g_err = msg; }

// params.at(key) of the reference: std::out_of_range when absent.
double need(const std::map<std::string, double>& p, const char* key) {
  auto it = p.find(key);
  if (it == p.end())
    throw nosh::Error(kMissingParam, std::string("parameter '") + key + "' not found (map::at)");
  return it->second;
}

void compute_ct(MeshCache& m, nosh_sfield_s* t, DBuf<double>& ct,
                const std::map<std::string, double>& p, hipStream_t s) {
  if (ct.n != size_t(m.n)) ct.alloc(m.n);
  launch_mul(m.cv_int.p, t->sf.get_v(p, s), ct.p, m.n, s);
}

}  // namespace nosh

void nosh_keo_s::set_parameters(const std::map<std::string, double>& p, hipStream_t s) {
  MeshCache& m = *mesh->m;
  // mvp_->set_parameters(params): mu = params.at("mu")
  // (src/vector_field_explicit_values.cpp:56-62).
  const double mu = need(p, "mu");
  mvp->vf.mu = mu;
  if (!alpha_valid) {
    alpha.alloc(m.E);
    launch_alpha(m, thickness_ref(thickness, s), alpha.p, s);
    alpha_valid = true;
  }
  if (kv.valid && kv.mu == mu && (!dkeo || kv.dvalid)) return;
  launch_keo_fill(m, alpha.p, mvp->vf.atilde.p, mu, dkeo, dparam == "mu", kv, s);
}

extern "C" {

const char* nosh_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------ mesh
int nosh_mesh_create(const double* coords, int64_t num_vertices, const int32_t* cells,
                     int64_t num_cells, int vertices_per_cell, int reorder, nosh_mesh_t* out) {
  return guard([&] {
    NOSH_CHECK(coords && cells && out, kInvalid, "null argument");
    auto h = std::make_unique<nosh_mesh_s>();
    h->m = build_mesh_cache(coords, num_vertices, cells, num_cells, vertices_per_cell,
                            reorder != 0, default_stream());
    *out = h.release();
  });
}

int nosh_mesh_destroy(nosh_mesh_t mesh) {
  return guard([&] { delete mesh; });
}

int nosh_mesh_sizes(nosh_mesh_t mesh, int64_t* n, int64_t* E, int64_t* T) {
  return guard([&] {
    NOSH_CHECK(mesh, kInvalid, "null mesh");
    if (n) *n = mesh->m->n;
    if (E) *E = mesh->m->E;
    if (T) *T = mesh->m->T;
  });
}

int nosh_mesh_edges(nosh_mesh_t mesh, int32_t* edges) {
  return guard([&] {
    NOSH_CHECK(mesh && edges, kInvalid, "null argument");
    mesh->m->edges.download(edges, 2 * mesh->m->E, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mesh_cell_edges(nosh_mesh_t mesh, int32_t* ce) {
  return guard([&] {
    NOSH_CHECK(mesh && ce, kInvalid, "null argument");
    mesh->m->cell_edges.download(ce, mesh->m->T * mesh->m->ne, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mesh_edge_data(nosh_mesh_t mesh, double* length, double* covolume) {
  return guard([&] {
    NOSH_CHECK(mesh && length && covolume, kInvalid, "null argument");
    mesh->m->len.download(length, mesh->m->E, default_stream());
    mesh->m->covol.download(covolume, mesh->m->E, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mesh_control_volumes(nosh_mesh_t mesh, double* cv) {
  return guard([&] {
    NOSH_CHECK(mesh && cv, kInvalid, "null argument");
    mesh->m->cv.download(cv, mesh->m->n, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mesh_permutation(nosh_mesh_t mesh, int32_t* perm) {
  return guard([&] {
    NOSH_CHECK(mesh && perm, kInvalid, "null argument");
    mesh->m->perm.download(perm, mesh->m->n, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mesh_layout(nosh_mesh_t mesh, int64_t* nnz_pad, int64_t* nslices, int* max_width) {
  return guard([&] {
    NOSH_CHECK(mesh, kInvalid, "null mesh");
    if (nnz_pad) *nnz_pad = mesh->m->nnz_pad;
    if (nslices) *nslices = mesh->m->nslices;
    if (max_width) *max_width = mesh->m->max_width;
  });
}

int nosh_mesh_layout_split(nosh_mesh_t mesh, int64_t* nnz_upper, int64_t* nnz_lower) {
  return guard([&] {
    NOSH_CHECK(mesh, kInvalid, "null mesh");
    if (nnz_upper) *nnz_upper = mesh->m->nnz_u;
    if (nnz_lower) *nnz_lower = mesh->m->nnz_l;
  });
}

int nosh_mesh_layout_column_bytes(nosh_mesh_t mesh, int* bytes) {
  return guard([&] {
    NOSH_CHECK(mesh && bytes, kInvalid, "null argument");
    *bytes = mesh->m->uoff.p ? 2 : 4;
  });
}

int nosh_mesh_to_internal(nosh_mesh_t mesh, const double* x_ref, double* x_int, void* stream) {
  return guard([&] {
    NOSH_CHECK(mesh && x_ref && x_int, kInvalid, "null argument");
    launch_permute_in(mesh->m->n, mesh->m->perm.p, x_ref, x_int, stream_of(stream));
  });
}

int nosh_mesh_to_reference(nosh_mesh_t mesh, const double* x_int, double* x_ref, void* stream) {
  return guard([&] {
    NOSH_CHECK(mesh && x_ref && x_int, kInvalid, "null argument");
    launch_permute_out(mesh->m->n, mesh->m->perm.p, x_int, x_ref, stream_of(stream));
  });
}

// ---------------------------------------------------------------- fields
int nosh_vfield_explicit_values_create(nosh_mesh_t mesh, const double* A, double mu,
                                       nosh_vfield_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && A && out, kInvalid, "null argument");
    MeshCache& m = *mesh->m;
    auto h = std::make_unique<nosh_vfield_s>();
    h->mesh = mesh;
    h->vf.mesh = &m;
    h->vf.mu = mu;
    h->vf.atilde.alloc(m.E);
    DBuf<double> Ad(3 * m.n);
    Ad.upload(A, 3 * m.n, default_stream());
    launch_atilde(m, Ad.p, h->vf.atilde.p, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
    *out = h.release();
  });
}

int nosh_vfield_edge_projections(nosh_vfield_t field, double* out) {
  return guard([&] {
    NOSH_CHECK(field && out, kInvalid, "null argument");
    const int64_t E = field->mesh->m->E;
    std::vector<double> at(E);
    field->vf.atilde.download(at.data(), E, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
    for (int64_t k = 0; k < E; ++k) out[k] = field->vf.mu * at[k];
  });
}

int nosh_vfield_destroy(nosh_vfield_t field) {
  return guard([&] { delete field; });
}

int nosh_sfield_constant_create(nosh_mesh_t mesh, double c, const char* param_name,
                                double param_init, nosh_sfield_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && out, kInvalid, "null argument");
    auto h = std::make_unique<nosh_sfield_s>();
    h->sf.mesh = mesh->m.get();
    h->sf.constant = true;
    h->sf.c = c;
    h->sf.param = param_name ? param_name : "";
    h->sf.param_init = param_init;
    *out = h.release();
  });
}

int nosh_sfield_explicit_values_create(nosh_mesh_t mesh, const double* values,
                                       nosh_sfield_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && values && out, kInvalid, "null argument");
    auto h = std::make_unique<nosh_sfield_s>();
    h->sf.mesh = mesh->m.get();
    h->sf.constant = false;
    h->sf.param = "beta";
    h->sf.param_init = 1.0;
    h->sf.values.assign(values, values + mesh->m->n);
    *out = h.release();
  });
}

int nosh_sfield_destroy(nosh_sfield_t field) {
  return guard([&] { delete field; });
}

// ------------------------------------------------------------------- keo
static nosh_keo_s* new_keo(nosh_mesh_t mesh, nosh_sfield_t t, nosh_vfield_t mvp, bool dkeo,
                           const char* dparam) {
  NOSH_CHECK(mesh && t && mvp, kInvalid, "null argument");
  auto* k = new nosh_keo_s;
  k->mesh = mesh;
  k->thickness = t;
  k->mvp = mvp;
  k->dkeo = dkeo;
  k->dparam = dparam ? dparam : "";
  return k;
}

int nosh_keo_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                    nosh_keo_t* out) {
  return guard([&] {
    NOSH_CHECK(out, kInvalid, "null argument");
    *out = new_keo(mesh, thickness, mvp, false, nullptr);
  });
}

int nosh_dkeo_dp_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                        const char* param_name, nosh_keo_t* out) {
  return guard([&] {
    NOSH_CHECK(out && param_name, kInvalid, "null argument");
    *out = new_keo(mesh, thickness, mvp, true, param_name);
  });
}

int nosh_keo_set_parameters(nosh_keo_t keo, int np, const char* const* names,
                            const double* values) {
  return guard([&] {
    NOSH_CHECK(keo, kInvalid, "null keo");
    keo->set_parameters(to_map(np, names, values), default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

static void keo_op(nosh_keo_s* keo, const double* x, double* y, hipStream_t s) {
  MeshCache& m = *keo->mesh->m;
  if (keo->dkeo) {
    // dK/dp alone: the dF/dp kernel with a zero vertex term.
    static thread_local DBuf<double>& zero = *new DBuf<double>;
    if (zero.n != size_t(m.n)) {
      zero.alloc(m.n);
      zero.zero(s);
    }
    launch_dfdp(m, keo->kv, zero.p, x, false, 0.0, nullptr, y, s);
  } else {
    launch_keo_apply(m, keo->kv, x, y, s);
  }
}

int nosh_keo_apply(nosh_keo_t keo, const double* x, double* y, int ncols, int64_t ld, int mode,
                   double alpha, double beta) {
  return guard([&] {
    NOSH_CHECK(keo, kInvalid, "null keo");
    check_apply_mode(mode, alpha, beta);
    NOSH_CHECK(keo->kv.valid, kNotBuilt, "keo: set_parameters has not been called");
    host_apply(*keo->mesh->m, keo->st, x, y, ncols, ld,
               [&](const double* a, double* b, hipStream_t s) { keo_op(keo, a, b, s); });
  });
}

int nosh_keo_apply_device(nosh_keo_t keo, const double* x, double* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(keo, kInvalid, "null keo");
    NOSH_CHECK(keo->kv.valid, kNotBuilt, "keo: set_parameters has not been called");
    device_apply(*keo->mesh->m, keo->st, x, y, stream_of(stream),
                 [&](const double* a, double* b, hipStream_t s) { keo_op(keo, a, b, s); });
  });
}

int nosh_keo_apply_internal(nosh_keo_t keo, const double* x, double* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(keo && x && y, kInvalid, "null argument");
    NOSH_CHECK(keo->kv.valid, kNotBuilt, "keo: set_parameters has not been called");
    keo_op(keo, x, y, stream_of(stream));
  });
}

int nosh_keo_destroy(nosh_keo_t keo) {
  return guard([&] { delete keo; });
}

// -------------------------------------------------------------- jacobian
int nosh_jac_create(nosh_mesh_t mesh, nosh_sfield_t V, nosh_sfield_t thickness, nosh_keo_t keo,
                    nosh_jac_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && V && thickness && keo && out, kInvalid, "null argument");
    auto* j = new nosh_jac_s;
    j->mesh = mesh;
    j->V = V;
    j->thickness = thickness;
    j->keo = keo;
    *out = j;
  });
}

static void jac_rebuild_dev(nosh_jac_s* jac, const std::map<std::string, double>& p,
                            const double* psi_int, hipStream_t s) {
  MeshCache& m = *jac->mesh->m;
  // keo_->set_parameters(params) then rebuild_diags_ (src/jacobian_operator.cpp:135-138).
  jac->keo->set_parameters(p, s);
  const double g = need(p, "g");
  compute_ct(m, jac->thickness, jac->ct, p, s);
  const double* V = jac->V->sf.get_v(p, s);
  if (jac->blk.n != size_t(3 * m.n)) jac->blk.alloc(3 * m.n);
  launch_rebuild_blocks(m, jac->keo->kv, jac->ct.p, V, g, psi_int, jac->blk.p, nullptr, nullptr,
                        s);
  jac->built = true;
  jac->blk32_valid = false;
}

int nosh_jac_rebuild(nosh_jac_t jac, int np, const char* const* names, const double* values,
                     const double* psi) {
  return guard([&] {
    NOSH_CHECK(jac && psi, kInvalid, "null argument");
    MeshCache& m = *jac->mesh->m;
    hipStream_t s = default_stream();
    if (jac->psi.n != size_t(2 * m.n)) jac->psi.alloc(2 * m.n);
    upload_internal(m, jac->st, psi, jac->psi.p, s);
    jac_rebuild_dev(jac, to_map(np, names, values), jac->psi.p, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_jac_apply(nosh_jac_t jac, const double* x, double* y, int ncols, int64_t ld, int mode,
                   double alpha, double beta) {
  return guard([&] {
    NOSH_CHECK(jac, kInvalid, "null jacobian");
    check_apply_mode(mode, alpha, beta);
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    MeshCache& m = *jac->mesh->m;
    host_apply(m, jac->st, x, y, ncols, ld, [&](const double* a, double* b, hipStream_t s) {
      launch_block_apply(m, jac->keo->kv, jac->blk.p, a, b, s);
    });
  });
}

int nosh_jac_apply_device(nosh_jac_t jac, const double* x, double* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(jac, kInvalid, "null jacobian");
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    MeshCache& m = *jac->mesh->m;
    device_apply(m, jac->st, x, y, stream_of(stream),
                 [&](const double* a, double* b, hipStream_t s) {
                   launch_block_apply(m, jac->keo->kv, jac->blk.p, a, b, s);
                 });
  });
}

int nosh_jac_apply_internal(nosh_jac_t jac, const double* x, double* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(jac && x && y, kInvalid, "null argument");
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    launch_block_apply(*jac->mesh->m, jac->keo->kv, jac->blk.p, x, y, stream_of(stream));
  });
}

namespace {
// The FP32 values of K and the FP32 copy of the blocks, made once per rebuild.
void jac_prepare_fp32(nosh_jac_t jac, hipStream_t s) {
  MeshCache& m = *jac->mesh->m;
  if (!jac->keo->kv.valid32) launch_keo_to_fp32(m, jac->keo->kv, s);
  if (!jac->blk32_valid || jac->blk32.n != size_t(3 * m.n)) {
    if (jac->blk32.n != size_t(3 * m.n)) jac->blk32.alloc(3 * m.n);
    convert_to<float>(jac->blk.p, jac->blk32.p, 3 * m.n, s);
    jac->blk32_valid = true;
  }
}
}  // namespace

int nosh_jac_apply_internal_fp32(nosh_jac_t jac, const float* x, float* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(jac && x && y, kInvalid, "null argument");
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    hipStream_t s = stream_of(stream);
    jac_prepare_fp32(jac, s);
    launch_block_apply32(*jac->mesh->m, jac->keo->kv, jac->blk32.p, x, y, s);
  });
}

int nosh_jac_bench_apply(nosh_jac_t jac, const void* x, void* y, int fp32, int cold, int reps,
                         void* stream, double* seconds_per_launch) {
  hipEvent_t flush_done = nullptr;
  std::vector<hipEvent_t> ev;
  int rc = guard([&] {
    NOSH_CHECK(jac && x && y && seconds_per_launch && reps >= 1, kInvalid,
               "invalid J.phi benchmark arguments");
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    MeshCache& m = *jac->mesh->m;
    hipStream_t s = stream_of(stream);
    if (fp32) jac_prepare_fp32(jac, s);
    auto apply = [&](hipEvent_t t0, hipEvent_t t1) {
      if (fp32)
        launch_block_apply32(m, jac->keo->kv, jac->blk32.p, static_cast<const float*>(x),
                             static_cast<float*>(y), s, t0, t1);
      else
        launch_block_apply(m, jac->keo->kv, jac->blk.p, static_cast<const double*>(x),
                           static_cast<double*>(y), s, t0, t1);
    };
    // 1 GiB, four times the Infinity Cache, read before each cold launch.
    const int64_t flush_bytes = int64_t(1) << 30;
    DBuf<unsigned char> flush;
    DBuf<float> part;
    int dev = 0, cus = 0;
    NOSH_HIP(hipGetDevice(&dev));
    NOSH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = unsigned(std::max(1, cus) * 8);
    if (cold) {
      flush.alloc(size_t(flush_bytes));
      part.alloc(grid);
      NOSH_HIP(hipMemsetAsync(flush.p, 0, size_t(flush_bytes), s));
    }
    ev.resize(size_t(2 * reps), nullptr);
    for (auto& e : ev) NOSH_HIP(hipEventCreate(&e));
    for (int r = 0; r < 2; ++r) {
      if (cold) launch_cache_flush(flush.p, flush_bytes, part.p, grid, s);
      apply(nullptr, nullptr);
    }
    for (int r = 0; r < reps; ++r) {
      if (cold) launch_cache_flush(flush.p, flush_bytes, part.p, grid, s);
      apply(ev[size_t(2 * r)], ev[size_t(2 * r + 1)]);
    }
    NOSH_HIP(hipEventCreate(&flush_done));
    NOSH_HIP(hipEventRecord(flush_done, s));
    NOSH_HIP(hipEventSynchronize(flush_done));
    double sum = 0.0;
    for (int r = 0; r < reps; ++r) {
      float ms = 0.f;
      NOSH_HIP(hipEventElapsedTime(&ms, ev[size_t(2 * r)], ev[size_t(2 * r + 1)]));
      sum += double(ms);
    }
    *seconds_per_launch = sum * 1e-3 / reps;
  });
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (flush_done) (void)hipEventDestroy(flush_done);
  return rc;
}

int nosh_jac_destroy(nosh_jac_t jac) {
  return guard([&] { delete jac; });
}

// ------------------------------------------------------- keo_regularized
int nosh_prec_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                     nosh_prec_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && thickness && mvp && out, kInvalid, "null argument");
    auto* p = new nosh_prec_s;
    p->mesh = mesh;
    p->thickness = thickness;
    p->keo.reset(new_keo(mesh, thickness, mvp, false, nullptr));
    p->bound.alloc(1);
    *out = p;
  });
}

int nosh_prec_set_chebyshev(nosh_prec_t prec, int iters, double ratio, double upper) {
  return guard([&] {
    NOSH_CHECK(prec, kInvalid, "null preconditioner");
    NOSH_CHECK(iters >= 0 && iters <= 64, kInvalid, "Chebyshev iterations must be in [0, 64]");
    NOSH_CHECK(ratio > 1.0, kInvalid, "Chebyshev ratio must be > 1");
    prec->iters = iters;
    prec->ratio = ratio;
    prec->upper = upper;
  });
}

static void prec_rebuild_dev(nosh_prec_s* prec, const std::map<std::string, double>& p,
                             const double* psi_int, hipStream_t s) {
  MeshCache& m = *prec->mesh->m;
  prec->keo->set_parameters(p, s);
  const double g = need(p, "g");
  compute_ct(m, prec->thickness, prec->ct, p, s);
  if (prec->blk.n != size_t(3 * m.n)) prec->blk.alloc(3 * m.n);
  launch_rebuild_blocks(m, prec->keo->kv, prec->ct.p, prec->ct.p, g, psi_int, nullptr,
                        prec->blk.p, prec->bound.p, s);
  prec->built = true;
}

int nosh_prec_rebuild(nosh_prec_t prec, int np, const char* const* names, const double* values,
                      const double* psi) {
  return guard([&] {
    NOSH_CHECK(prec && psi, kInvalid, "null argument");
    MeshCache& m = *prec->mesh->m;
    hipStream_t s = default_stream();
    if (prec->psi.n != size_t(2 * m.n)) prec->psi.alloc(2 * m.n);
    upload_internal(m, prec->st, psi, prec->psi.p, s);
    prec_rebuild_dev(prec, to_map(np, names, values), prec->psi.p, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

static PrecView<double> prec_view(nosh_prec_s* p) {
  PrecView<double> v;
  v.val = p->keo->kv.val.p;
  v.blk = p->blk.p;
  v.bound = p->bound.p;
  v.upper = p->upper;
  v.ratio = p->ratio;
  v.iters = p->iters;
  return v;
}

static void prec_op(nosh_prec_s* prec, const double* x, double* y, hipStream_t s) {
  MeshCache& m = *prec->mesh->m;
  prec->work.ensure(m.n);
  prec_apply<double>(m, prec_view(prec), prec->work, reinterpret_cast<const double2*>(x),
                     reinterpret_cast<double2*>(y), s);
}

int nosh_prec_apply(nosh_prec_t prec, const double* x, double* y, int ncols, int64_t ld,
                    int mode, double alpha, double beta) {
  return guard([&] {
    NOSH_CHECK(prec, kInvalid, "null preconditioner");
    check_apply_mode(mode, alpha, beta);
    NOSH_CHECK(prec->built, kNotBuilt, "keo_regularized: rebuild has not been called");
    host_apply(*prec->mesh->m, prec->st, x, y, ncols, ld,
               [&](const double* a, double* b, hipStream_t s) { prec_op(prec, a, b, s); });
  });
}

int nosh_prec_apply_device(nosh_prec_t prec, const double* x, double* y, void* stream) {
  return guard([&] {
    NOSH_CHECK(prec, kInvalid, "null preconditioner");
    NOSH_CHECK(prec->built, kNotBuilt, "keo_regularized: rebuild has not been called");
    device_apply(*prec->mesh->m, prec->st, x, y, stream_of(stream),
                 [&](const double* a, double* b, hipStream_t s) { prec_op(prec, a, b, s); });
  });
}

int nosh_prec_apply_operator(nosh_prec_t prec, const double* x, double* y) {
  return guard([&] {
    NOSH_CHECK(prec, kInvalid, "null preconditioner");
    NOSH_CHECK(prec->built, kNotBuilt, "keo_regularized: rebuild has not been called");
    MeshCache& m = *prec->mesh->m;
    host_apply(m, prec->st, x, y, 1, 2 * m.n, [&](const double* a, double* b, hipStream_t s) {
      launch_block_apply(m, prec->keo->kv, prec->blk.p, a, b, s);
    });
  });
}

int nosh_prec_upper_bound(nosh_prec_t prec, double* upper) {
  return guard([&] {
    NOSH_CHECK(prec && upper, kInvalid, "null argument");
    NOSH_CHECK(prec->built, kNotBuilt, "keo_regularized: rebuild has not been called");
    if (prec->upper > 0.0) {
      *upper = prec->upper;
      return;
    }
    double r = 0.0;
    NOSH_HIP(hipMemcpy(&r, prec->bound.p, sizeof(double), hipMemcpyDeviceToHost));
    *upper = 1.0 + r;
  });
}

int nosh_prec_destroy(nosh_prec_t prec) {
  return guard([&] { delete prec; });
}

// ------------------------------------------------------------------- nls
int nosh_nls_create(nosh_mesh_t mesh, nosh_vfield_t mvp, nosh_sfield_t V, double g,
                    nosh_sfield_t thickness, const double* initial_x, const char* deriv,
                    nosh_nls_t* out) {
  return guard([&] {
    NOSH_CHECK(mesh && mvp && V && thickness && out, kInvalid, "null argument");
    auto h = std::make_unique<nosh_nls_s>();
    h->mesh = mesh;
    h->mvp = mvp;
    h->V = V;
    h->thickness = thickness;
    h->g = g;
    h->deriv = deriv ? deriv : "mu";
    const int64_t n = mesh->m->n;
    if (initial_x)
      h->initial_x.assign(initial_x, initial_x + 2 * n);
    else {
      h->initial_x.assign(2 * n, 0.0);
      for (int64_t k = 0; k < n; ++k) h->initial_x[2 * k] = 1.0;
    }
    h->keo.reset(new_keo(mesh, thickness, mvp, false, nullptr));
    h->dkeo.reset(new_keo(mesh, thickness, mvp, true, h->deriv.c_str()));
    // Sorted union: {"g"} + scalar potential params + {"mu"}; earlier keys win
    // (src/model_evaluator_nls.cpp:102-111).
    std::map<std::string, double> params;
    params["g"] = g;
    for (auto& kv : V->sf.scalar_parameters()) params.insert(kv);
    params.insert({"mu", mvp->vf.mu});
    for (auto& kv : params) {
      h->p_names.push_back(kv.first);
      h->p_nominal.push_back(kv.second);
    }
    h->nrm2.alloc(1);
    h->bound.alloc(1);
    *out = h.release();
  });
}

int nosh_nls_num_params(nosh_nls_t nls, int* count) {
  return guard([&] {
    NOSH_CHECK(nls && count, kInvalid, "null argument");
    *count = int(nls->p_names.size());
  });
}

int nosh_nls_param_name(nosh_nls_t nls, int index, char* buf, int buflen) {
  return guard([&] {
    NOSH_CHECK(nls && buf && buflen > 0, kInvalid, "null argument");
    NOSH_CHECK(index >= 0 && index < int(nls->p_names.size()), kInvalid, "index out of range");
    const std::string& s = nls->p_names[index];
    NOSH_CHECK(int(s.size()) < buflen, kInvalid, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int nosh_nls_param_nominal(nosh_nls_t nls, int index, double* value) {
  return guard([&] {
    NOSH_CHECK(nls && value, kInvalid, "null argument");
    NOSH_CHECK(index >= 0 && index < int(nls->p_names.size()), kInvalid, "index out of range");
    *value = nls->p_nominal[index];
  });
}

static void nls_f_dev(nosh_nls_s* nls, const std::map<std::string, double>& p, const double* x,
                      double* f, double* nrm2, hipStream_t s) {
  MeshCache& m = *nls->mesh->m;
  nls->keo->set_parameters(p, s);
  const double g = need(p, "g");
  compute_ct(m, nls->thickness, nls->ct, p, s);
  const double* V = nls->V->sf.get_v(p, s);
  launch_compute_f(m, nls->keo->kv, nls->ct.p, V, g, x, f, nrm2, s);
}

}  // extern "C"

namespace nosh {
double residual_norm(nosh_nls_s* nls, const double* x_int,
                     const std::map<std::string, double>& p, hipStream_t s) {
  const int64_t n = nls->mesh->m->n;
  if (nls->f.n != size_t(2 * n)) nls->f.alloc(2 * n);
  if (nls->nrm2.n < 1) nls->nrm2.alloc(1);
  nls_f_dev(nls, p, x_int, nls->f.p, nls->nrm2.p, s);
  double nf2 = 0.0;
  NOSH_HIP(hipMemcpyAsync(&nf2, nls->nrm2.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NOSH_HIP(hipStreamSynchronize(s));
  return std::sqrt(nf2);
}
}  // namespace nosh

extern "C" {

int nosh_nls_eval_f(nosh_nls_t nls, const double* x, int np, const char* const* names,
                    const double* values, double* f) {
  return guard([&] {
    NOSH_CHECK(nls && x && f, kInvalid, "null argument");
    const auto p = to_map(np, names, values);
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    nls->st.ensure(2 * m.n);
    upload_internal(m, nls->st, x, nls->st.b.p, s);
    nls_f_dev(nls, p, nls->st.b.p, nls->st.c.p, nullptr, s);
    download_internal(m, nls->st, nls->st.c.p, f, s);
  });
}

int nosh_nls_eval_f_device(nosh_nls_t nls, const double* x, int np, const char* const* names,
                           const double* values, double* f, void* stream) {
  return guard([&] {
    NOSH_CHECK(nls && x && f, kInvalid, "null argument");
    const auto p = to_map(np, names, values);
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = stream_of(stream);
    nls->st.ensure(2 * m.n);
    launch_permute_in(m.n, m.perm.p, x, nls->st.b.p, s);
    nls_f_dev(nls, p, nls->st.b.p, nls->st.c.p, nullptr, s);
    launch_permute_out(m.n, m.perm.p, nls->st.c.p, f, s);
  });
}

int nosh_nls_eval_f_internal(nosh_nls_t nls, const double* x, int np, const char* const* names,
                             const double* values, double* f, void* stream) {
  return guard([&] {
    NOSH_CHECK(nls && x && f, kInvalid, "null argument");
    const auto p = to_map(np, names, values);
    nls_f_dev(nls, p, x, f, nullptr, stream_of(stream));
  });
}

int nosh_nls_eval_dfdp(nosh_nls_t nls, const double* x, int np, const char* const* names,
                       const double* values, int col, double* dfdp) {
  return guard([&] {
    NOSH_CHECK(nls && x && dfdp, kInvalid, "null argument");
    NOSH_CHECK(col >= 0 && col < int(nls->p_names.size()), kInvalid, "column out of range");
    const auto p = to_map(np, names, values);
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    // dkeo_dp_->set_parameters(params) (src/model_evaluator_nls.cpp:640).
    nls->dkeo->set_parameters(p, s);
    const std::string& name = nls->p_names[col];
    compute_ct(m, nls->thickness, nls->ct, p, s);
    nls->st.ensure(2 * m.n);
    upload_internal(m, nls->st, x, nls->st.b.p, s);
    launch_dfdp(m, nls->dkeo->kv, nls->ct.p, nls->st.b.p, name == "g",
                nls->V->sf.dvdp_constant(name), nls->V->sf.dvdp_values(name, s), nls->st.c.p,
                s);
    download_internal(m, nls->st, nls->st.c.p, dfdp, s);
  });
}

int nosh_nls_create_w_op(nosh_nls_t nls, nosh_jac_t* out) {
  return nls ? nosh_jac_create(nls->mesh, nls->V, nls->thickness, nls->keo.get(), out)
             : (g_err = "null model", NOSH_ERR_INVALID);
}

int nosh_nls_create_w_prec(nosh_nls_t nls, nosh_prec_t* out) {
  return nls ? nosh_prec_create(nls->mesh, nls->thickness, nls->mvp, out)
             : (g_err = "null model", NOSH_ERR_INVALID);
}

int nosh_nls_destroy(nosh_nls_t nls) {
  return guard([&] { delete nls; });
}

// --------------------------------------------------------------- solvers
int nosh_newton_default_options(nosh_newton_options* o) {
  return guard([&] {
    NOSH_CHECK(o, kInvalid, "null options");
    o->tol = 1e-10;
    o->max_iters = 30;
    o->minres_rtol = 1e-10;
    o->minres_max_iters = 1000;
    o->cheb_iters = 6;
    o->cheb_ratio = 300.0;
    o->precision = 0;
    o->forcing = 0;
    o->forcing_min = 1e-6;
    o->forcing_max = 0.1;
  });
}

}  // extern "C"

void nosh::newton_run(nosh_nls_s* nls, double* psi_int, const std::map<std::string, double>& p,
                       const nosh_newton_options& o, nosh_newton_stats* st, hipStream_t s) {
  MeshCache& m = *nls->mesh->m;
  const int64_t n = m.n;
  const auto t0 = std::chrono::steady_clock::now();
  const double g = need(p, "g");
  need(p, "mu");
  nls->keo->set_parameters(p, s);
  compute_ct(m, nls->thickness, nls->ct, p, s);
  const double* V = nls->V->sf.get_v(p, s);
  if (nls->f.n != size_t(2 * n)) nls->f.alloc(2 * n);
  if (nls->jblk.n != size_t(3 * n)) {
    nls->jblk.alloc(3 * n);
    nls->pblk.alloc(3 * n);
  }
  const bool mixed = o.precision == 1;
  if (mixed) {
    if (!nls->keo->kv.valid32) launch_keo_to_fp32(m, nls->keo->kv, s);
    if (nls->jblk32.n != size_t(3 * n)) {
      nls->jblk32.alloc(3 * n);
      nls->pblk32.alloc(3 * n);
    }
  }
  nosh_newton_stats local{};
  double nf_prev = -1.0, eta_prev = o.forcing_max;
  int it = 0;
  for (;; ++it) {
    launch_compute_f(m, nls->keo->kv, nls->ct.p, V, g, psi_int, nls->f.p, nls->nrm2.p, s);
    double nf2 = 0.0;
    NOSH_HIP(hipMemcpyAsync(&nf2, nls->nrm2.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NOSH_HIP(hipStreamSynchronize(s));
    const double nf = std::sqrt(nf2);
    if (it < 64) local.norm_f_history[it] = nf;
    local.final_norm_f = nf;
    if (!(nf > o.tol)) {
      local.converged = std::isfinite(nf) ? 1 : 0;
      break;
    }
    if (it >= o.max_iters || !std::isfinite(nf)) break;
    // Forcing term.
    double eta = o.minres_rtol;
    if (o.forcing == 1) {
      if (nf_prev < 0.0) {
        eta = o.forcing_max;
      } else {
        eta = 0.9 * (nf / nf_prev) * (nf / nf_prev);
        const double safe = 0.9 * eta_prev * eta_prev;
        if (safe > 0.1) eta = std::max(eta, safe);
      }
      eta = std::min(eta, o.forcing_max);
      eta = std::max(eta, 0.5 * o.tol / nf);
      eta = std::max(eta, o.forcing_min);
    }
    launch_rebuild_blocks(m, nls->keo->kv, nls->ct.p, V, g, psi_int, nls->jblk.p, nls->pblk.p,
                          nls->bound.p, s);
    double rr = 0.0;
    int k;
    if (!mixed) {
      JacView<double> J{nls->keo->kv.val.p, nls->jblk.p};
      PrecView<double> P{nls->keo->kv.val.p, nls->pblk.p, nls->bound.p, 0.0, o.cheb_ratio,
                         o.cheb_iters};
      k = minres_solve<double>(m, J, P, nls->w64, nls->f.p, -1.0, eta, o.minres_max_iters,
                               psi_int, s, &rr, 8);
    } else {
      convert_to<float>(nls->jblk.p, nls->jblk32.p, 3 * n, s);
      convert_to<float>(nls->pblk.p, nls->pblk32.p, 3 * n, s);
      JacView<float> J{nls->keo->kv.val32.p, nls->jblk32.p};
      PrecView<float> P{nls->keo->kv.val32.p, nls->pblk32.p, nls->bound.p, 0.0, o.cheb_ratio,
                        o.cheb_iters};
      k = minres_solve<float>(m, J, P, nls->w32, nls->f.p, -1.0, eta, o.minres_max_iters,
                              psi_int, s, &rr, 8);
    }
    if (it < 64) local.minres_history[it] = k;
    local.minres_iters_total += k;
    nf_prev = nf;
    eta_prev = eta;
  }
  local.newton_iters = it;
  local.seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = local;
}

extern "C" {

int nosh_newton_solve(nosh_nls_t nls, double* psi, int np, const char* const* names,
                      const double* values, const nosh_newton_options* opts,
                      nosh_newton_stats* stats) {
  return guard([&] {
    NOSH_CHECK(nls && psi, kInvalid, "null argument");
    nosh_newton_options o;
    if (opts)
      o = *opts;
    else
      nosh_newton_default_options(&o);
    const auto p = to_map(np, names, values);
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = default_stream();
    if (nls->x.n != size_t(2 * m.n)) nls->x.alloc(2 * m.n);
    upload_internal(m, nls->st, psi, nls->x.p, s);
    newton_run(nls, nls->x.p, p, o, stats, s);
    download_internal(m, nls->st, nls->x.p, psi, s);
  });
}

int nosh_newton_solve_device(nosh_nls_t nls, double* psi, int np, const char* const* names,
                             const double* values, const nosh_newton_options* opts,
                             nosh_newton_stats* stats, void* stream) {
  return guard([&] {
    NOSH_CHECK(nls && psi, kInvalid, "null argument");
    nosh_newton_options o;
    if (opts)
      o = *opts;
    else
      nosh_newton_default_options(&o);
    const auto p = to_map(np, names, values);
    MeshCache& m = *nls->mesh->m;
    hipStream_t s = stream_of(stream);
    if (nls->x.n != size_t(2 * m.n)) nls->x.alloc(2 * m.n);
    launch_permute_in(m.n, m.perm.p, psi, nls->x.p, s);
    newton_run(nls, nls->x.p, p, o, stats, s);
    launch_permute_out(m.n, m.perm.p, nls->x.p, psi, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_minres_solve(nosh_jac_t jac, nosh_prec_t prec, const double* b, double* x,
                      double rtol, int max_iters, int* iters, double* relres) {
  return guard([&] {
    NOSH_CHECK(jac && b && x, kInvalid, "null argument");
    NOSH_CHECK(jac->built, kNotBuilt, "jacobian_operator: rebuild has not been called");
    NOSH_CHECK(!prec || prec->built, kNotBuilt, "keo_regularized: rebuild has not been called");
    MeshCache& m = *jac->mesh->m;
    hipStream_t s = default_stream();
    jac->st.ensure(2 * m.n);
    upload_internal(m, jac->st, b, jac->st.b.p, s);
    JacView<double> J{jac->keo->kv.val.p, jac->blk.p};
    PrecView<double> P{jac->keo->kv.val.p, jac->blk.p, nullptr, 2.0, 30.0, 0};
    if (prec) P = prec_view(prec);
    static MinresWork<double>& work = *new MinresWork<double>;
    // The solution accumulates into a zeroed vector.
    NOSH_HIP(hipMemsetAsync(jac->st.c.p, 0, sizeof(double) * 2 * m.n, s));
    double rr = 0.0;
    const int k = minres_solve<double>(m, J, P, work, jac->st.b.p, 1.0, rtol, max_iters,
                                       jac->st.c.p, s, &rr, 8);
    download_internal(m, jac->st, jac->st.c.p, x, s);
    if (iters) *iters = k;
    if (relres) *relres = rr;
  });
}

}  // extern "C"
