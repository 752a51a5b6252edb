// stream.hip -- the attainable HBM bandwidth of this GPU, measured with the
// library's own streaming kernels (read, write, copy) on buffers far larger
// than the 256 MB Infinity Cache. The roofline figures of the operator
// kernels are reported against this ceiling beside the nominal 8 TB/s.
//
// Each thread moves 16-byte vectors; a block of 256 threads covers
// kUnroll * 4 KiB per grid-stride iteration with all kUnroll loads issued
// before any is used, and the grid is a whole multiple of the CU count.
// Loads and stores are non-temporal (the data is used once), or ordinary
// when nt = 0. No reference counterpart: a measurement helper.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "handles.hpp"

using namespace nosh;

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;
typedef float f4v __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ f4v ld(const f4v* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ __forceinline__ void st(f4v* p, f4v v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Sum of every 16-byte vector; one float per block goes out (so the loads
// are not dead).
template <bool NT>
__global__ __launch_bounds__(kThreads) void k_stream_read(const f4v* __restrict__ a, int64_t nv,
                                                          float* __restrict__ out) {
  f4v acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t step = int64_t(gridDim.x) * kThreads * kUnroll;
  for (int64_t base = int64_t(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < nv;
       base += step) {
    f4v v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + int64_t(u) * kThreads;
      v[u] = i < nv ? ld<NT>(a + i) : f4v{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += v[u];
  }
  __shared__ float red[kThreads];
  red[threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_stream_write(f4v* __restrict__ a, int64_t nv,
                                                           float val) {
  const f4v v = {val, val, val, val};
  const int64_t step = int64_t(gridDim.x) * kThreads * kUnroll;
  for (int64_t base = int64_t(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < nv;
       base += step) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + int64_t(u) * kThreads;
      if (i < nv) st<NT>(a + i, v);
    }
  }
}

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_stream_copy(const f4v* __restrict__ a,
                                                          f4v* __restrict__ b, int64_t nv) {
  const int64_t step = int64_t(gridDim.x) * kThreads * kUnroll;
  for (int64_t base = int64_t(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < nv;
       base += step) {
    f4v v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + int64_t(u) * kThreads;
      v[u] = i < nv ? ld<NT>(a + i) : f4v{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + int64_t(u) * kThreads;
      if (i < nv) st<NT>(b + i, v[u]);
    }
  }
}

template <bool NT>
void launch(int kind, unsigned grid, f4v* a, f4v* b, int64_t nv, float* out, hipStream_t s) {
  if (kind == 0)
    k_stream_read<NT><<<grid, kThreads, 0, s>>>(a, nv, out);
  else if (kind == 1)
    k_stream_write<NT><<<grid, kThreads, 0, s>>>(a, nv, 1.0f);
  else
    k_stream_copy<NT><<<grid, kThreads, 0, s>>>(a, b, nv);
}

}  // namespace

namespace nosh {
void launch_cache_flush(const void* buf, int64_t bytes, float* out, unsigned grid,
                        hipStream_t s) {
  k_stream_read<false><<<grid, kThreads, 0, s>>>(static_cast<const f4v*>(buf), bytes / 16, out);
}
}  // namespace nosh

extern "C" {

// Average seconds per launch of the streaming kernel `kind` (0 read, 1
// write, 2 copy) over `bytes` bytes per buffer (rounded down to 16-byte
// vectors), timed with events over `reps` launches after two untimed ones;
// blocks_per_cu workgroups of 256 threads per CU (0: 8); nt: non-temporal
// accesses. The algorithmic bytes per launch are `bytes` for read and
// write and 2 * bytes for copy.
int nosh_bench_stream(int kind, int64_t bytes, int reps, int blocks_per_cu, int nt,
                      double* seconds_per_launch) {
  try {
    NOSH_CHECK(kind >= 0 && kind <= 2 && bytes >= 16 && reps >= 1 && seconds_per_launch,
               kInvalid, "invalid stream benchmark arguments");
    const int64_t nv = bytes / 16;
    hipStream_t s = default_stream();
    int dev = 0, cus = 0;
    NOSH_HIP(hipGetDevice(&dev));
    NOSH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = unsigned(std::max(1, cus) * (blocks_per_cu > 0 ? blocks_per_cu : 8));
    DBuf<f4v> a, b;
    DBuf<float> out;
    a.alloc(size_t(nv));
    if (kind == 2) b.alloc(size_t(nv));
    out.alloc(grid);
    // Defined contents for the reads.
    if (kind != 1) k_stream_write<true><<<grid, kThreads, 0, s>>>(a.p, nv, 0.5f);
    hipEvent_t e0, e1;
    NOSH_HIP(hipEventCreate(&e0));
    NOSH_HIP(hipEventCreate(&e1));
    for (int r = 0; r < 2; ++r) {
      if (nt)
        launch<true>(kind, grid, a.p, b.p, nv, out.p, s);
      else
        launch<false>(kind, grid, a.p, b.p, nv, out.p, s);
    }
    NOSH_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
      if (nt)
        launch<true>(kind, grid, a.p, b.p, nv, out.p, s);
      else
        launch<false>(kind, grid, a.p, b.p, nv, out.p, s);
    }
    NOSH_HIP(hipEventRecord(e1, s));
    NOSH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    NOSH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds_per_launch = double(ms) * 1e-3 / reps;
    set_last_error("");
    return NOSH_OK;
  } catch (const nosh::Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

}  // extern "C"
