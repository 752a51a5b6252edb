// noalloc.cpp -- the Gibbs energy of a state without any device allocation.
//
// nosh_nls_gibbs_energy allocates its per-block partial sums for every call.
// Callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc) use nosh_nls_gibbs_energy_noalloc
// instead: the same two kernels' arithmetic, with the partial sums in a
// file-scope device array (`scratch`) and input and output in the caller's
// device memory. The value is -sum_k c_k |psi_k|^4 / sum_k c_k, the intended
// result of nls::gibbs_energy (src/model_evaluator_nls.cpp:697-738, whose
// code returns 0).
//
// `scratch` is one array per GPU, alive until the process ends, shared by
// every call on that GPU: calls that overlap on one GPU (different streams,
// different threads) race on it. Each call overwrites every partial it
// reads, so nothing needs clearing between calls.

#include <hip/hip_runtime.h>

#include <mutex>

#include "handles.hpp"
#include "kernels.hpp"

using namespace nosh;

namespace {

constexpr int kThreads = 256;
// Largest mesh: kMaxBlocks blocks of kThreads vertices (4,194,304 vertices;
// the 160^3 Kuhn cube has 4,173,281). Two partials per block: 256 KiB.
constexpr int kMaxBlocks = 16384;
constexpr int64_t kMaxVertices = int64_t(kMaxBlocks) * kThreads;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[2 * kMaxBlocks];

// Device address of `scratch` on each GPU (hipGetSymbolAddress, once per
// device, with that device current). The array's name is not a device
// address in host code; kernels get this pointer.
double* scratch_ptrs[kMaxDevices] = {};
std::mutex scratch_mutex;

double* scratch_for_current_device() {
  int dev = 0;
  NOSH_HIP(hipGetDevice(&dev));
  NOSH_CHECK(dev >= 0 && dev < kMaxDevices, kUnsupported, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptrs[dev]) {
    void* p = nullptr;
    NOSH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NOSH_CHECK(p != nullptr, kInvalid, "no device address for the scratch array");
    scratch_ptrs[dev] = static_cast<double*>(p);
  }
  return scratch_ptrs[dev];
}

// Partials of (sum c |psi|^4, sum c) over the internal vertex order, with psi
// read in reference order through the permutation: the same values, in the
// same order and with the same block sums, as the internal-order reduction
// of nosh_nls_gibbs_energy.
__global__ void k_gibbs_ref_partials(int64_t n, const double* __restrict__ cv_int,
                                     const int32_t* __restrict__ perm,
                                     const double2* __restrict__ psi_ref,
                                     double* __restrict__ partials) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double c = cv_int[i];
    const double2 x = psi_ref[perm[i]];
    const double r2 = x.x * x.x + x.y * x.y;
    v[0] = c * r2 * r2;
    v[1] = c;
  }
  block_partials<2>(v, partials);
}

__global__ void k_gibbs_final(const double* __restrict__ partials, int nb,
                              double* __restrict__ out) {
  double v[2];
  finalize_sum<2>(v, partials, nb);
  if (threadIdx.x == 0) *out = -1.0 * v[0] / v[1];
}

}  // namespace

extern "C" {

int nosh_nls_gibbs_energy_noalloc(nosh_nls_t nls, const double* psi, double* energy,
                                  void* stream) {
  try {
    NOSH_CHECK(nls && psi && energy, kInvalid, "null argument");
    MeshCache& m = *nls->mesh->m;
    NOSH_CHECK(m.n >= 1 && m.n <= kMaxVertices, kInvalid,
               "nosh_nls_gibbs_energy_noalloc: the mesh has more vertices than its scratch "
               "array covers (4,194,304)");
    double* part = scratch_for_current_device();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = int((m.n + kThreads - 1) / kThreads);
    k_gibbs_ref_partials<<<nb, kThreads, 0, s>>>(m.n, m.cv_int.p, m.perm.p,
                                                 reinterpret_cast<const double2*>(psi), part);
    k_gibbs_final<<<1, kFinalizeThreads, 0, s>>>(part, nb, energy);
    NOSH_HIP(hipGetLastError());
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

}  // extern "C"
