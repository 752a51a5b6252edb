// kernels.hpp -- device building blocks shared by the operator and solver
// kernels: the sliced-ELLPACK row product, XCD-aware block mapping and
// deterministic block / grid reductions.
#pragma once

#include "internal.hpp"

namespace nosh {

template <class T>
using V2 = typename Vec2<T>::type;

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.
// Remap so that XCD x processes one contiguous range of workgroup indices,
// keeping each XCD's gathers of x within its own L2.
__device__ __forceinline__ int64_t xcd_block(bool remap) {
  const int64_t b = blockIdx.x;
  if (!remap) return b;
  const int64_t nb = gridDim.x;
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t xcd = b % 8, idx = b / 8;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

__device__ __forceinline__ int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Load of a value that no kernel writes, at a wave-uniform address, through
// the constant address space (a scalar load).
__device__ __forceinline__ int32_t ld_uniform(const int32_t* p) {
#if __HIP_DEVICE_COMPILE__
  return *reinterpret_cast<const __attribute__((address_space(4))) int32_t*>(
      reinterpret_cast<uintptr_t>(p));
#else
  return *p;  // host pass only parses device functions
#endif
}

// Off-diagonal product of one internal row over the split layout: the upper
// part reads its values in place, sum val_k x[col_k]; the lower part reads
// each value at its position in the upper part (a gather that the
// neighbouring rows' waves have just brought into L2) and uses its
// conjugate, sum conj(val_pos_k) x[col_k]. Upper slots first, each part in
// slot order. The columns come as 16-bit offsets from the row when the
// layout has them (2 bytes a slot instead of 4), else as 32-bit indices;
// the same columns in the same order either way.
template <class T, bool OFF16>
__device__ __forceinline__ V2<T> sell_row_cols(int64_t row, const SellView& S,
                                               const V2<T>* __restrict__ val,
                                               const V2<T>* __restrict__ x) {
  const int32_t s = uniform(int32_t(row >> 6));
  const int lane = int(row & 63);
  const int32_t r32 = int32_t(row);
  T ar = T(0), ai = T(0);
  {
    const int32_t b = uniform(S.uptr[s]);
    const int32_t e = uniform(S.uptr[s + 1]);
#pragma unroll 4
    for (int32_t j = b + lane; j < e; j += kSliceH) {
      const int32_t c = OFF16 ? r32 + int32_t(__builtin_nontemporal_load(S.uoff + j))
                              : __builtin_nontemporal_load(S.ucol + j);
      const V2<T> v = val[j];
      const V2<T> xv = x[c];
      ar += v.x * xv.x - v.y * xv.y;
      ai += v.x * xv.y + v.y * xv.x;
    }
  }
  {
    const int32_t b = uniform(S.lptr[s]);
    const int32_t e = uniform(S.lptr[s + 1]);
#pragma unroll 4
    for (int32_t j = b + lane; j < e; j += kSliceH) {
      const int32_t c = OFF16 ? r32 + int32_t(__builtin_nontemporal_load(S.loff + j))
                              : __builtin_nontemporal_load(S.lcol + j);
      const int32_t p = __builtin_nontemporal_load(S.lpos + j);
      const V2<T> v = val[p];
      const V2<T> xv = x[c];
      ar += v.x * xv.x + v.y * xv.y;
      ai += v.x * xv.y - v.y * xv.x;
    }
  }
  V2<T> r;
  r.x = ar;
  r.y = ai;
  return r;
}

template <class T>
__device__ __forceinline__ V2<T> sell_row(int64_t row, const SellView& S,
                                          const V2<T>* __restrict__ val,
                                          const V2<T>* __restrict__ x) {
  return S.uoff ? sell_row_cols<T, true>(row, S, val, x) : sell_row_cols<T, false>(row, S, val, x);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide sums of NV values; the totals end up in every lane of wave 0.
// lds must hold 16*NV doubles (blockDim <= 1024).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    v[q] = wave_sum(v[q]);
    if (lane == 0) lds[q * 16 + wid] = v[q];
  }
  __syncthreads();
  if (wid == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(lane < nw ? lds[q * 16 + lane] : 0.0);
  }
  __syncthreads();
}

// Phase one of a grid-wide deterministic sum: each block stores its NV
// partial sums at partials[q * gridDim.x + blockIdx.x]. Phase two is a
// separate one-block kernel (finalize_sum) that adds them in block order;
// the kernel boundary makes the partials visible across XCDs without a
// per-block device-scope fence (which on gfx950 writes back the L2).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partials) {
  __shared__ double lds[16 * NV];
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[q * gridDim.x + blockIdx.x] = v[q];
  }
}

constexpr int kFinalizeThreads = 256;

// Sums nb partials per value in fixed order; totals valid in thread 0. The
// order is that of a block of kFinalizeThreads threads whatever the block
// size (a multiple of 64): virtual thread t adds partials t, t + 256, ...,
// each virtual wave of 64 is reduced by lane shuffles, and the 4 wave sums
// by a last one, so the persistent kernel (blocks of 128 to 1024 threads)
// and the separate finalize kernels (256) give the same bits.
template <int NV>
__device__ __forceinline__ void finalize_sum(double (&v)[NV], const double* __restrict__ partials,
                                             int nb) {
  __shared__ double lds[16 * NV];
  constexpr int kVw = kFinalizeThreads / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = max(1, int(blockDim.x) >> 6);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    for (int vw = wid; vw < kVw; vw += nw) {
      double s = 0.0;
      for (int i = vw * 64 + lane; i < nb; i += kFinalizeThreads) s += partials[q * nb + i];
      s = wave_sum(s);
      if (lane == 0) lds[q * 16 + vw] = s;
    }
  }
  __syncthreads();
  if (wid == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(lane < kVw ? lds[q * 16 + lane] : 0.0);
  }
  __syncthreads();
}

}  // namespace nosh
