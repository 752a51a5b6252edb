// batch_handle.hpp -- the batch handle behind nosh_batch_t and the host
// drivers shared by the batched solve and the batched continuation.
#pragma once

#include <vector>

#include "batched.hpp"
#include "handles.hpp"

struct nosh_batch_s {
  nosh::Batch b;
  std::vector<nosh_nls_s*> models;
  std::vector<int> counts;
  std::vector<double> mu_h, g_h;  // per instance, as given at creation
  int64_t total_len = 0;          // sum over instances of 2 n
  nosh::DBuf<double> stage;
};

namespace nosh {

// Instance vectors (reference order, concatenated, device) <-> interleaved state.
void scatter_all(nosh_batch_s* h, const double* src_dev, hipStream_t s);
void gather_all(nosh_batch_s* h, double* dst_dev, hipStream_t s);

// Lockstep Newton over the batch (FP32 inner MINRES, FP64 outer). Instances
// with skip[i] != 0 are left untouched; max_iters[i], when given, replaces
// o.max_iters per instance. Returns the lockstep MINRES iterations (per
// Newton step the largest count over the active instances, summed over the
// steps).
// stop_below > 0: return as soon as at most stop_below instances are still
// iterating (they come back unconverged with newton_iters < their limit and
// their current iterate in the state). ptc0: initial pseudo-transient shifts
// per instance. stall > 0: an instance whose ||F|| has not fallen below half
// its best value for `stall` steps stops, unconverged (Newton wandering at a
// branch transition).
int newton_batch(nosh_batch_s* h, const nosh_newton_options& o, nosh_batch_instance_stats* st,
                 hipStream_t s, const int* skip = nullptr, const int* max_iters = nullptr,
                 int stop_below = 0, const double* ptc0 = nullptr, int stall = 0);

// Tangent predictor of every instance: rebuilds J at the current psi, solves
// J z = -scale dF/dmu (FP32 MINRES to relative tolerance rtol, all instances
// in lockstep) and leaves psi + z = psi + scale dpsi/dmu in the state.
// Returns the lockstep MINRES iterations.
int tangent_batch(nosh_batch_s* h, const nosh_newton_options& o, double scale, double rtol,
                  hipStream_t s);

// ||F(psi)|| of every instance at the batch's current state and mu, g.
void batch_residuals(nosh_batch_s* h, std::vector<double>& norms, hipStream_t s);

}  // namespace nosh
