// handles.hpp -- the C++ objects behind the opaque C-ABI handles.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nosh_amd.h"
#include "internal.hpp"
#include "solver.hpp"

using nosh::DBuf;
using nosh::KeoValues;
using nosh::MeshCache;
using nosh::MinresWork;
using nosh::ScalarField;
using nosh::VectorField;

struct nosh_mesh_s {
  std::unique_ptr<MeshCache> m;
};

struct nosh_vfield_s {
  nosh_mesh_s* mesh;
  VectorField vf;
};

struct nosh_sfield_s {
  ScalarField sf;
  // Reference-order copy of the values at the default parameters (for alpha).
  DBuf<double> ref;
};

// Device staging for host-pointer applies.
struct Staging {
  DBuf<double> a, b, c, d;
  void ensure(int64_t m) {
    if (a.n != size_t(m)) {
      a.alloc(m);
      b.alloc(m);
      c.alloc(m);
      d.alloc(m);
    }
  }
};

struct nosh_keo_s {
  nosh_mesh_s* mesh;
  nosh_sfield_s* thickness;
  nosh_vfield_s* mvp;
  bool dkeo = false;
  std::string dparam;
  DBuf<double> alpha;  // alpha cache (built on first fill)
  bool alpha_valid = false;
  KeoValues kv;
  Staging st;
  void set_parameters(const std::map<std::string, double>& p, hipStream_t s);
};

struct nosh_jac_s {
  nosh_mesh_s* mesh;
  nosh_sfield_s* V;
  nosh_sfield_s* thickness;
  nosh_keo_s* keo;
  DBuf<double> ct, blk, psi;
  DBuf<float> blk32;
  bool built = false;
  bool blk32_valid = false;  // blk32 holds the FP32 copy of the current blk
  Staging st;
};

struct nosh_prec_s {
  nosh_mesh_s* mesh;
  nosh_sfield_s* thickness;
  std::unique_ptr<nosh_keo_s> keo;
  DBuf<double> ct, blk, psi, bound;
  int iters = 6;
  double ratio = 300.0, upper = 0.0;
  bool built = false;
  MinresWork<double> work;
  Staging st;
};

struct nosh_nls_s {
  nosh_mesh_s* mesh;
  nosh_vfield_s* mvp;
  nosh_sfield_s* V;
  nosh_sfield_s* thickness;
  double g;
  std::vector<double> initial_x;
  std::string deriv;
  std::unique_ptr<nosh_keo_s> keo, dkeo;
  std::vector<std::string> p_names;
  std::vector<double> p_nominal;
  DBuf<double> ct, x, f, nrm2;
  Staging st;
  // Newton work.
  DBuf<double> jblk, pblk, bound;
  DBuf<float> jblk32, pblk32;
  MinresWork<double> w64;
  MinresWork<float> w32;
};

namespace nosh {
// Sets the message nosh_last_error() returns on this thread.
void set_last_error(const std::string& msg);
// Shared helpers of the C-ABI translation units (capi.cpp).
double need(const std::map<std::string, double>& p, const char* key);
// ||F(x)|| at the parameters p, x in internal order (uses nls->f, nls->nrm2).
double residual_norm(nosh_nls_s* nls, const double* x_int,
                     const std::map<std::string, double>& p, hipStream_t s);
void compute_ct(MeshCache& m, nosh_sfield_s* t, DBuf<double>& ct,
                const std::map<std::string, double>& p, hipStream_t s);
void newton_run(nosh_nls_s* nls, double* psi_int, const std::map<std::string, double>& p,
                const nosh_newton_options& o, nosh_newton_stats* st, hipStream_t s);
}  // namespace nosh
