// batch_capi.cpp -- C ABI of the batched solver and its autotuner.
//
// A batch groups instances by model (mesh + fields); each instance has its
// own mu, g and state. Newton runs in lockstep over the batch: one F
// evaluation, one block rebuild and one MINRES (FP32 inner, FP64 scalars and
// update) per step for all instances still above the tolerance.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "batch_handle.hpp"

using namespace nosh;

namespace {

template <class F>
int bguard(F&& f) {
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

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

void alloc_group(GroupHost& G) {
  const int64_t m = G.mesh->n * G.dev.bp;
  G.psi.alloc(m);
  G.f.alloc(m);
  G.x.alloc(m);
  G.jblk.alloc(3 * m);
  G.pblk.alloc(3 * m);
  for (auto* v : {&G.Y, &G.Z2, &G.v, &G.v1, &G.u, &G.R0, &G.R1, &G.W0, &G.W1, &G.W2}) {
    v->alloc(m);
    v->zero(default_stream());
  }
  BGroup& d = G.dev;
  d.psi = G.psi.p;
  d.f = G.f.p;
  d.x = G.x.p;
  d.jblk = G.jblk.p;
  d.pblk = G.pblk.p;
  d.Y = G.Y.p;
  d.Z2 = G.Z2.p;
  d.v = G.v.p;
  d.v1 = G.v1.p;
  d.u = G.u.p;
  d.R0 = G.R0.p;
  d.R1 = G.R1.p;
  d.W0 = G.W0.p;
  d.W1 = G.W1.p;
  d.W2 = G.W2.p;
}

// Deterministic relative perturbation of every lane (at most 1e-3), so that
// the autotuner's timed Newton step has MINRES work even from an exact
// solution.
__global__ void k_perturb(int64_t m, double2* __restrict__ psi) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double f = 1.0 + 1e-3 * (double((i * 7919) % 2001) / 1000.0 - 1.0);
  double2 p = psi[i];
  p.x *= f;
  p.y = p.y * f + 1e-4 * (f - 1.0);
  psi[i] = p;
}

}  // namespace

namespace nosh {

void scatter_all(nosh_batch_s* h, const double* src_dev, hipStream_t s) {
  int64_t off = 0;
  for (auto& G : h->b.ghost) {
    for (int b = 0; b < G.B; ++b) {
      launch_scatter_lane(G.mesh->n, G.dev.bp, b, G.mesh->perm.p, src_dev + off,
                          reinterpret_cast<double*>(G.psi.p), s);
      off += 2 * G.mesh->n;
    }
  }
}

void gather_all(nosh_batch_s* h, double* dst_dev, hipStream_t s) {
  int64_t off = 0;
  for (auto& G : h->b.ghost) {
    for (int b = 0; b < G.B; ++b) {
      launch_gather_lane(G.mesh->n, G.dev.bp, b, G.mesh->perm.p,
                         reinterpret_cast<const double*>(G.psi.p), dst_dev + off, s);
      off += 2 * G.mesh->n;
    }
  }
}

void batch_residuals(nosh_batch_s* h, std::vector<double>& norms, hipStream_t s) {
  h->b.compute_norms(s, norms);
}

int newton_batch(nosh_batch_s* h, const nosh_newton_options& o, nosh_batch_instance_stats* st,
                 hipStream_t s, const int* skip, const int* max_iters, int stop_below,
                 const double* ptc0, int stall) {
  Batch& B = h->b;
  const int ni = B.args.ninst;
  B.args.iters = o.cheb_iters;
  B.args.ratio = o.cheb_ratio;
  B.args.upper = 0.0;
  B.args.z0 = std::max(o.cheb_iters - 1, 0) % 2;
  B.drop_graph();
  std::vector<double> nf, nf_prev(ni, -1.0), eta_prev(ni, o.forcing_max), rt(ni);
  std::vector<int> done(ni, 0), newton(ni, 0), mres(ni, 0), last_gain(ni, 0);
  std::vector<double> best(ni, -1.0), f1(ni, -1.0), f2(ni, -1.0);
  if (skip)
    for (int i = 0; i < ni; ++i) done[i] = skip[i] != 0;
  std::vector<double> final_nf(ni, 0.0);
  // Pseudo-transient shifts: sigma_{k+1} = sigma_k ||F_k|| / ||F_{k-1}||
  // (switched evolution relaxation), 0 for plain Newton.
  std::vector<double> sig(ni, 0.0);
  if (ptc0)
    for (int i = 0; i < ni; ++i) sig[i] = ptc0[i];
  int lockstep = 0;
  for (int it = 0;; ++it) {
    B.compute_norms(s, nf);
    int nactive = 0;
    for (int i = 0; i < ni; ++i) {
      if (done[i]) {
        B.active_h[i] = 0;
        continue;
      }
      final_nf[i] = nf[i];
      if (!(nf[i] > o.tol) || !std::isfinite(nf[i]) ||
          it >= (max_iters ? max_iters[i] : o.max_iters)) {
        done[i] = 1;
        B.active_h[i] = 0;
        continue;
      }
      // Stagnation: ||F|| not below half its best value for `stall` steps,
      // and not halving on each of the last two steps either (a run that
      // descends steadily from a jump up goes on).
      const bool descending = f2[i] > 0.0 && nf[i] < 0.5 * f1[i] && f1[i] < 0.5 * f2[i];
      f2[i] = f1[i];
      f1[i] = nf[i];
      if (best[i] < 0.0 || nf[i] < 0.5 * best[i]) {
        best[i] = nf[i];
        last_gain[i] = it;
      } else if (stall > 0 && it - last_gain[i] >= stall && !descending) {
        done[i] = 1;
        B.active_h[i] = 0;
        continue;
      }
      double eta = o.minres_rtol;
      if (o.forcing == 1) {
        if (nf_prev[i] < 0.0) {
          eta = o.forcing_max;
        } else {
          eta = 0.9 * (nf[i] / nf_prev[i]) * (nf[i] / nf_prev[i]);
          const double safe = 0.9 * eta_prev[i] * eta_prev[i];
          if (safe > 0.1) eta = std::max(eta, safe);
        }
        eta = std::min(eta, o.forcing_max);
        eta = std::max(eta, 0.5 * o.tol / nf[i]);
        eta = std::max(eta, o.forcing_min);
      }
      rt[i] = eta;
      eta_prev[i] = eta;
      if (ptc0 && nf_prev[i] > 0.0) sig[i] *= nf[i] / nf_prev[i];
      nf_prev[i] = nf[i];
      B.active_h[i] = 1;
      newton[i] = it + 1;
      ++nactive;
    }
    if (nactive == 0) break;
    if (stop_below > 0 && nactive <= stop_below) {
      // Hand back the instances still iterating; this step is not taken.
      for (int i = 0; i < ni; ++i)
        if (B.active_h[i]) newton[i] = it;
      break;
    }
    if (std::getenv("NOSH_PROGRESS") && (it < 8 || it % 5 == 0 || nactive < 4)) {
      double fmin = 1e300, fmax = 0.0;
      for (int i = 0; i < ni; ++i)
        if (B.active_h[i]) {
          fmin = std::min(fmin, nf[i]);
          fmax = std::max(fmax, nf[i]);
        }
      std::fprintf(stderr, "[batch newton] step %d active %d of %d, ||F|| in [%.3e, %.3e]\n", it,
                   nactive, ni, fmin, fmax);
    }
    NOSH_HIP(hipMemcpyAsync(B.active.p, B.active_h.data(), sizeof(int) * ni,
                            hipMemcpyHostToDevice, s));
    NOSH_HIP(hipMemcpyAsync(B.rtol.p, rt.data(), sizeof(double) * ni, hipMemcpyHostToDevice, s));
    if (ptc0) {
      int i0 = 0;
      while (i0 + 1 < ni && !B.active_h[i0]) ++i0;
      if (std::getenv("NOSH_PROGRESS"))
        std::fprintf(stderr, "[batch newton] step %d instance %d ||F|| %.3e sigma %.3e\n", it,
                     i0, nf[i0], sig[i0]);
      NOSH_HIP(hipMemcpyAsync(B.shift.p, sig.data(), sizeof(double) * ni, hipMemcpyHostToDevice,
                              s));
    }
    lockstep += B.minres(s, o.minres_max_iters);
    for (int i = 0; i < ni; ++i)
      if (B.active_h[i]) mres[i] += B.Sh[i].it;
  }
  if (ptc0) B.shift.zero(s);
  if (st)
    for (int i = 0; i < ni; ++i) {
      if (skip && skip[i]) continue;
      st[i].newton_iters = newton[i];
      st[i].minres_iters = mres[i];
      st[i].final_norm_f = final_nf[i];
      st[i].converged = (final_nf[i] <= o.tol) ? 1 : 0;
    }
  return lockstep;
}

int tangent_batch(nosh_batch_s* h, const nosh_newton_options& o, double scale, double rtol,
                  hipStream_t s) {
  Batch& B = h->b;
  const int ni = B.args.ninst;
  B.args.iters = o.cheb_iters;
  B.args.ratio = o.cheb_ratio;
  B.args.upper = 0.0;
  B.args.z0 = std::max(o.cheb_iters - 1, 0) % 2;
  B.drop_graph();
  B.active_h.assign(ni, 1);
  const std::vector<double> rt(ni, rtol);
  NOSH_HIP(hipMemcpyAsync(B.active.p, B.active_h.data(), sizeof(int) * ni, hipMemcpyHostToDevice,
                          s));
  NOSH_HIP(hipMemcpyAsync(B.rtol.p, rt.data(), sizeof(double) * ni, hipMemcpyHostToDevice, s));
  B.dfdmu_into_f(s, scale);
  return B.minres(s, o.minres_max_iters);
}

}  // namespace nosh

namespace {

std::string tune_key(nosh_batch_s* h, const nosh_newton_options& o) {
  hipDeviceProp_t prop{};
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipGetDeviceProperties(&prop, dev);
  std::ostringstream k;
  k << "v2|" << prop.gcnArchName << "|cheb" << o.cheb_iters;
  for (auto& G : h->b.ghost) k << "|" << G.mesh->n << "x" << G.dev.bp;
  return k.str();
}

std::string tune_path() {
  if (const char* p = std::getenv("NOSH_AUTOTUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/nosh_amd_autotune.tsv";
}

bool tune_lookup(const std::string& key, int* threads, int* fold, int* mode) {
  std::ifstream f(tune_path());
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string k;
    int t, d, m;
    if (ls >> k >> t >> d >> m && k == key) {
      *threads = t;
      *fold = d;
      *mode = m;
      return true;
    }
  }
  return false;
}

void tune_store(const std::string& key, int threads, int fold, int mode) {
  std::ofstream f(tune_path(), std::ios::app);
  if (f) f << key << " " << threads << " " << fold << " " << mode << "\n";
}

}  // namespace

extern "C" {

int nosh_batch_create(int ngroups, const nosh_nls_t* models, const int* counts, const double* mu,
                      const double* g, nosh_batch_t* out) {
  return bguard([&] {
    NOSH_CHECK(ngroups >= 1 && ngroups <= 16 && models && counts && mu && g && out, kInvalid,
               "invalid batch description");
    auto h = std::make_unique<nosh_batch_s>();
    Batch& B = h->b;
    hipStream_t s = default_stream();
    int inst = 0;
    for (int k = 0; k < ngroups; ++k) {
      NOSH_CHECK(models[k] && counts[k] >= 1 && counts[k] <= 64, kInvalid,
                 "each group needs a model and 1..64 instances");
      nosh_nls_s* md = models[k];
      h->models.push_back(md);
      h->counts.push_back(counts[k]);
      GroupHost G;
      G.mesh = md->mesh->m.get();
      G.B = counts[k];
      MeshCache& m = *G.mesh;
      // The model's nominal parameters (a parameterised potential keeps its
      // own value) with this group's first mu and g; alpha_e with the
      // model's thickness (keo::build_alpha_cache_).
      std::map<std::string, double> p0;
      for (size_t q = 0; q < md->p_names.size(); ++q) p0[md->p_names[q]] = md->p_nominal[q];
      p0["mu"] = mu[inst];
      p0["g"] = g[inst];
      md->keo->set_parameters(p0, s);
      G.s32.alloc(m.nnz_pad);
      G.a64.alloc(m.nnz_pad);
      G.t64.alloc(m.nnz_pad);
      G.diag.alloc(m.n);
      G.rowabs.alloc(m.n);
      launch_slot_fields(m, md->keo->alpha.p, md->mvp->vf.atilde.p, G.s32.p, G.a64.p,
                         G.t64.p, G.diag.p, G.rowabs.p, s);
      G.ct.alloc(m.n);
      compute_ct(m, md->thickness, G.ct, p0, s);
      G.V.alloc(m.n);
      NOSH_HIP(hipMemcpyAsync(G.V.p, md->V->sf.get_v(p0, s), sizeof(double) * m.n,
                              hipMemcpyDeviceToDevice, s));
      BGroup& d = G.dev;
      d.n = m.n;
      d.B = counts[k];
      d.bp = next_pow2(counts[k]);
      d.inst0 = inst;
      d.slice_ptr = m.slice_ptr.p;
      d.scol = m.scol.p;
      d.s32 = G.s32.p;
      d.kval = nullptr;
      d.kval64 = nullptr;
      d.a64 = G.a64.p;
      d.t64 = G.t64.p;
      d.diag = G.diag.p;
      d.rowabs = G.rowabs.p;
      d.ct = G.ct.p;
      d.V = G.V.p;
      alloc_group(G);
      h->total_len += int64_t(counts[k]) * 2 * m.n;
      inst += counts[k];
      B.ghost.push_back(std::move(G));
    }
    B.args.ngroups = ngroups;
    B.args.ninst = inst;
    B.mu.alloc(inst);
    B.g.alloc(inst);
    B.set_mu(mu, s);
    B.g.upload(g, inst, s);
    h->mu_h.assign(mu, mu + inst);
    h->g_h.assign(g, g + inst);
    B.bound.alloc(inst);
    B.rtol.alloc(inst);
    B.norms.alloc(inst);
    B.active.alloc(inst);
    B.S.alloc(inst);
    B.Sh.resize(inst);
    B.active_h.assign(inst, 0);
    B.args.S = B.S.p;
    B.args.mu = B.mu.p;
    B.args.g = B.g.p;
    B.args.bound = B.bound.p;
    B.args.rtol = B.rtol.p;
    B.shift.alloc(inst);
    B.shift.zero(s);
    B.args.shift = B.shift.p;
    B.args.active = B.active.p;
    B.relayout(256, 1);
    NOSH_HIP(hipStreamSynchronize(s));
    *out = h.release();
  });
}

int nosh_batch_set_tuning(nosh_batch_t batch, int block_threads, int fold, int mode,
                          int graph_iters, int poll) {
  return bguard([&] {
    NOSH_CHECK(batch, kInvalid, "null batch");
    int maxbp = 1;
    for (auto& G : batch->b.ghost) maxbp = std::max(maxbp, G.dev.bp);
    NOSH_CHECK(block_threads >= 64 && block_threads <= 1024 && block_threads % 64 == 0 &&
                   block_threads % maxbp == 0,
               kInvalid, "block_threads must be a multiple of 64 and of the lane count, <= 1024");
    NOSH_CHECK(fold >= 1 && fold <= 64, kInvalid, "fold must be in 1..64");
    NOSH_CHECK(mode >= 0 && mode <= 2, kInvalid,
               "mode must be 0 (graph), 1 (persistent) or 2 (graph, 2 lanes per thread)");
    NOSH_CHECK(graph_iters >= 1 && poll >= graph_iters, kInvalid, "bad graph_iters / poll");
    batch->b.graph_iters = graph_iters;
    batch->b.poll = poll;
    if (block_threads != batch->b.block_threads || fold != batch->b.fold)
      batch->b.relayout(block_threads, fold);
    batch->b.drop_graph();
    batch->b.mode = 0;
    batch->b.lanes = 1;
    int minbp = 64;
    for (auto& G : batch->b.ghost) minbp = std::min(minbp, G.dev.bp);
    NOSH_CHECK(mode != 2 || minbp >= nosh::kVecLanes, kUnsupported,
               "mode 2 needs at least 2 lanes in every group; use mode 0 (hipGraph)");
    batch->b.mode = mode == 1 ? 1 : 0;
    batch->b.lanes = mode == 2 ? nosh::kVecLanes : 1;
  });
}

int nosh_batch_time_cheb_step(nosh_batch_t batch, int reps, double* us_per_launch,
                              double* bytes_per_launch) {
  return bguard([&] {
    NOSH_CHECK(batch && us_per_launch, kInvalid, "null argument");
    hipStream_t s = default_stream();
    *us_per_launch = batch->b.time_cheb_step(s, reps);
    if (bytes_per_launch) {
      // Per padded slot the packed {column, alpha, a~} (12 B, shared by the
      // lanes); per row and lane z_in, r, z_{j-2} (read and overwritten by
      // z_j; not read at j = 1) (8 B each, FP32 complex) and the three FP32
      // entries of the P block.
      const double row_bytes = batch->b.args.iters >= 3 ? 44.0 : 36.0;
      double b = 0.0;
      for (auto& G : batch->b.ghost)
        b += 12.0 * double(G.mesh->nnz_pad) + row_bytes * double(G.mesh->n) * G.dev.bp;
      *bytes_per_launch = b;
    }
  });
}

namespace {

// Batched Newton with compaction: once at most half of the instances are
// still iterating, they continue in a new batch of their own (fewer lanes per
// row), recursively (128 random starts at 16^3-48^3: 18.4 s without, 12.1 s
// compacting at a quarter, 10.0 s at half). A lockstep launch costs about as much for one
// active lane of a row as for all of them, so without compaction the last
// few slow instances of a random-start batch set the cost of every step.
// psi: the instances' reference-order vectors, concatenated (device).
// A caller's batch compacts when it has at least kCompactMin instances; the
// batches made by compaction keep compacting down to kCompactMinSub.
constexpr int kCompactMin = 16;
constexpr int kCompactMinSub = 4;

void solve_compacting(nosh_batch_s* h, const nosh_newton_options& o, const int* caps,
                      double* psi, nosh_batch_instance_stats* st, hipStream_t s, int depth) {
  const int ni = h->b.args.ninst;
  scatter_all(h, psi, s);
  std::vector<nosh_batch_instance_stats> loc(ni);
  const int stop = ni >= (depth == 0 ? kCompactMin : kCompactMinSub) ? ni / 2 : 0;
  newton_batch(h, o, loc.data(), s, nullptr, caps, stop);
  gather_all(h, psi, s);
  std::vector<int> rem;
  for (int i = 0; i < ni; ++i) {
    st[i] = loc[i];
    const int cap = caps ? caps[i] : o.max_iters;
    if (!loc[i].converged && loc[i].newton_iters < cap && std::isfinite(loc[i].final_norm_f))
      rem.push_back(i);
  }
  if (rem.empty()) return;
  // Sub-batch of the instances still iterating, group by group.
  std::vector<int> gi(ni), off(ni);
  {
    int64_t o64 = 0;
    for (size_t g = 0; g < h->b.ghost.size(); ++g)
      for (int b = 0; b < h->b.ghost[g].B; ++b) {
        const int i = h->b.ghost[g].dev.inst0 + b;
        gi[i] = int(g);
        off[i] = int(o64);
        o64 += 2 * h->b.ghost[g].mesh->n;
      }
  }
  std::vector<nosh_nls_t> models;
  std::vector<int> counts, subcaps;
  std::vector<double> mu, g;
  for (int i : rem) {
    if (models.empty() || counts.back() == 64 || models.back() != h->models[gi[i]]) {
      models.push_back(h->models[gi[i]]);
      counts.push_back(0);
    }
    ++counts.back();
    mu.push_back(h->mu_h[i]);
    g.push_back(h->g_h[i]);
    subcaps.push_back((caps ? caps[i] : o.max_iters) - loc[i].newton_iters);
  }
  nosh_batch_t sub = nullptr;
  if (models.size() > 16 ||
      nosh_batch_create(int(models.size()), models.data(), counts.data(), mu.data(), g.data(),
                        &sub) != NOSH_OK) {
    // Not expressible as a batch: finish in this one.
    std::vector<int> skip(ni, 1), cap2(ni, 0);
    for (size_t q = 0; q < rem.size(); ++q) {
      skip[rem[q]] = 0;
      cap2[rem[q]] = subcaps[q];
    }
    scatter_all(h, psi, s);
    newton_batch(h, o, loc.data(), s, skip.data(), cap2.data());
    gather_all(h, psi, s);
    for (int i : rem) {
      loc[i].newton_iters += st[i].newton_iters;
      loc[i].minres_iters += st[i].minres_iters;
      st[i] = loc[i];
    }
    return;
  }
  std::unique_ptr<nosh_batch_s> owner(sub);
  // hipGraph mode, with 2 lanes per thread when the parent uses them and the
  // sub-batch's groups have the lanes for it (else the call leaves mode 0).
  nosh_batch_set_tuning(sub, h->b.block_threads, h->b.fold, h->b.lanes > 1 ? 2 : 0,
                        h->b.graph_iters, h->b.poll);
  DBuf<double> buf(size_t(sub->total_len));
  int64_t so = 0;
  for (int i : rem) {
    const int64_t len = 2 * h->b.ghost[gi[i]].mesh->n;
    NOSH_HIP(hipMemcpyAsync(buf.p + so, psi + off[i], sizeof(double) * len,
                            hipMemcpyDeviceToDevice, s));
    so += len;
  }
  std::vector<nosh_batch_instance_stats> sst(rem.size());
  solve_compacting(sub, o, subcaps.data(), buf.p, sst.data(), s, depth + 1);
  so = 0;
  for (size_t q = 0; q < rem.size(); ++q) {
    const int i = rem[q];
    const int64_t len = 2 * h->b.ghost[gi[i]].mesh->n;
    NOSH_HIP(hipMemcpyAsync(psi + off[i], buf.p + so, sizeof(double) * len,
                            hipMemcpyDeviceToDevice, s));
    so += len;
    st[i].newton_iters += sst[q].newton_iters;
    st[i].minres_iters += sst[q].minres_iters;
    st[i].final_norm_f = sst[q].final_norm_f;
    st[i].converged = sst[q].converged;
  }
  NOSH_HIP(hipStreamSynchronize(s));
  if (std::getenv("NOSH_PROGRESS"))
    std::fprintf(stderr, "[batch newton] compacted %d -> %d instances\n", ni, int(rem.size()));
}

}  // namespace

int nosh_batch_solve_device(nosh_batch_t batch, double* psi, const nosh_newton_options* opts,
                            nosh_batch_instance_stats* stats, void* stream) {
  return bguard([&] {
    NOSH_CHECK(batch && psi, kInvalid, "null argument");
    nosh_newton_options o;
    if (opts)
      o = *opts;
    else
      nosh_newton_default_options(&o);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : default_stream();
    std::vector<nosh_batch_instance_stats> st(batch->b.args.ninst);
    solve_compacting(batch, o, nullptr, psi, st.data(), s, 0);
    scatter_all(batch, psi, s);  // the batch holds the final states
    if (stats) std::copy(st.begin(), st.end(), stats);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_batch_solve(nosh_batch_t batch, double* psi, const nosh_newton_options* opts,
                     nosh_batch_instance_stats* stats) {
  return bguard([&] {
    NOSH_CHECK(batch && psi, kInvalid, "null argument");
    hipStream_t s = default_stream();
    if (batch->stage.n != size_t(batch->total_len)) batch->stage.alloc(batch->total_len);
    batch->stage.upload(psi, batch->total_len, s);
    const int rc = nosh_batch_solve_device(batch, batch->stage.p, opts, stats, s);
    if (rc != NOSH_OK) throw nosh::Error(rc, nosh_last_error());
    batch->stage.download(psi, batch->total_len, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_batch_autotune(nosh_batch_t batch, const double* psi, const nosh_newton_options* opts,
                        int use_cache, int* best_threads, int* best_fold, int* best_mode) {
  return bguard([&] {
    NOSH_CHECK(batch && psi, kInvalid, "null argument");
    nosh_newton_options o;
    if (opts)
      o = *opts;
    else
      nosh_newton_default_options(&o);
    const std::string key = tune_key(batch, o);
    int bt = 256, bf = 1, bm = 0;
    if (use_cache && tune_lookup(key, &bt, &bf, &bm)) {
      nosh_batch_set_tuning(batch, bt, bf, bm, batch->b.graph_iters, batch->b.poll);
    } else {
      // Time one Newton step (F, rebuild, MINRES capped at 40 iterations)
      // for every candidate block size, fold and iteration mode.
      int maxbp = 1;
      for (auto& G : batch->b.ghost) maxbp = std::max(maxbp, G.dev.bp);
      hipStream_t s = default_stream();
      if (batch->stage.n != size_t(batch->total_len)) batch->stage.alloc(batch->total_len);
      nosh_newton_options t = o;
      t.max_iters = 1;
      t.minres_max_iters = 40;
      t.forcing = 0;
      t.minres_rtol = 1e-30;
      double best = 1e30;
      for (int threads : {128, 256, 512, 1024}) {
        if (threads % maxbp || threads < maxbp) continue;
        for (int fold : {1, 2, 4, 8}) {
          for (int mode : {0, 1, 2}) {
            if (nosh_batch_set_tuning(batch, threads, fold, mode, batch->b.graph_iters,
                                      batch->b.poll) != NOSH_OK)
              continue;
            double tmin = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
              batch->stage.upload(psi, batch->total_len, s);
              NOSH_HIP(hipStreamSynchronize(s));
              const auto t0 = std::chrono::steady_clock::now();
              scatter_all(batch, batch->stage.p, s);
              for (auto& G : batch->b.ghost) {
                const int64_t m = G.mesh->n * G.dev.bp;
                k_perturb<<<grid_for(m, 256), 256, 0, s>>>(m, G.psi.p);
              }
              newton_batch(batch, t, nullptr, s);
              NOSH_HIP(hipStreamSynchronize(s));
              tmin = std::min(tmin, std::chrono::duration<double>(
                                        std::chrono::steady_clock::now() - t0).count());
            }
            if (std::getenv("NOSH_AUTOTUNE_VERBOSE"))
              std::fprintf(stderr, "autotune: block_threads %d fold %d mode %d: %.4f s\n",
                           threads, fold, mode, tmin);
            if (tmin < best) {
              best = tmin;
              bt = threads;
              bf = fold;
              bm = mode;
            }
          }
        }
      }
      nosh_batch_set_tuning(batch, bt, bf, bm, batch->b.graph_iters, batch->b.poll);
      if (use_cache) tune_store(key, bt, bf, bm);
    }
    if (best_threads) *best_threads = bt;
    if (best_fold) *best_fold = bf;
    if (best_mode) *best_mode = bm;
  });
}

int nosh_batch_destroy(nosh_batch_t batch) {
  return bguard([&] { delete batch; });
}

}  // extern "C"
