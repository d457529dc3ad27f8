// reduce.h -- device helpers shared by the solver kernels: global-address-space views and the
// fixed reduction tree of every dot product (wave64 shuffle tree, then (w0 + w1) + (w2 + w3),
// tile partials summed by thread t over t, t + 256, ...).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace gsh {

// Pointers read from a descriptor in device memory are generic to the compiler; casting them to
// the global address space makes every access a global_load/global_store.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) {
  return (gptr<T>)p;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8). This maps
// the blocks of each XCD onto one contiguous range of the grid's nb work items, so neighbouring
// tiles, whose gathers overlap, share that XCD's L2. It is a bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

__device__ __forceinline__ int32_t load_done(gptr<PcgState> st) {
  return *(volatile __attribute__((address_space(1))) int32_t*)&st->done;
}

__device__ __forceinline__ double wave_tree(double v) {
  v += __shfl_down(v, 32, 64);
  v += __shfl_down(v, 16, 64);
  v += __shfl_down(v, 8, 64);
  v += __shfl_down(v, 4, 64);
  v += __shfl_down(v, 2, 64);
  v += __shfl_down(v, 1, 64);
  return v;
}

// Sum over the 256 threads of a workgroup: wave trees, then (w0 + w1) + (w2 + w3).
// The total is returned to every thread.
__device__ __forceinline__ double block_tree(double v, double* red) {
  v = wave_tree(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  const double t = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return t;
}

// Orders LDS writes before reads among the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of the tile partials P[0..T): thread t sums P[t], P[t+256], ... then block_tree.
__device__ __forceinline__ double final_reduce(gptr<const double> P, int T, double* red) {
  double u = 0.0;
  for (int m = threadIdx.x; m < T; m += kTpb) u += P[m];
  return block_tree(u, red);
}

}  // namespace gsh
