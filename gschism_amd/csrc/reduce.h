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

// The done flag is written by earlier launches only; a plain (cached) load sees it after the
// launch boundary.
__device__ __forceinline__ int32_t load_done(gptr<PcgState> st) {
  return st->done;
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

// block_tree over the first 256 threads of a larger workgroup (waves 4 and up add nothing and
// write nothing; every thread passes the barriers). The total is returned to every thread.
__device__ __forceinline__ double block_tree_lo(double v, double* red) {
  v = wave_tree(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0 && w < 4) red[w] = v;
  __syncthreads();
  const double t = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return t;
}

// The partial of a tile of 256 * R rows whose workgroup has 256 * R threads, one row each, in the
// order of a 256-thread workgroup with R rows per thread: lane position t adds the values of rows
// t, t + 256, ..., t + 256 (R - 1) to 0.0 in that order, then block_tree. rows holds 256 * R
// doubles of LDS; v is the thread's own row value.
template <int R>
__device__ __forceinline__ double tile_tree_wide(double v, double* rows, double* red) {
  rows[threadIdx.x] = v;
  __syncthreads();
  double acc = 0.0;
  if (threadIdx.x < kTpb) {
#pragma unroll
    for (int j = 0; j < R; ++j) acc += rows[j * kTpb + threadIdx.x];
  }
  return block_tree_lo(acc, red);
}

// Orders LDS writes before reads among the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of the tile partials P[0..T): thread t sums P[t], P[t+256], ... then block_tree.
// Thread t adds P[t], P[t + 256], ... in that order; the loads of eight consecutive terms are
// issued together before their adds (a loop of single loads waits one memory latency per term).
__device__ __forceinline__ double final_reduce(gptr<const double> P, int T, double* red) {
  double u = 0.0;
  for (int m = threadIdx.x; m < T; m += 8 * kTpb) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (m + j * kTpb < T) ? P[m + j * kTpb] : 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (m + j * kTpb < T) u += v[j];
  }
  return block_tree(u, red);
}

// final_reduce of two partial arrays of T entries each, with the loads of both issued together.
__device__ __forceinline__ void final_reduce2(gptr<const double> P, gptr<const double> Q, int T, double* red,
                                              double& a, double& b) {
  double u = 0.0, w = 0.0;
  for (int m = threadIdx.x; m < T; m += 8 * kTpb) {
    double vp[8], vq[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const bool in = m + j * kTpb < T;
      vp[j] = in ? P[m + j * kTpb] : 0.0;
      vq[j] = in ? Q[m + j * kTpb] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (m + j * kTpb < T) {
        u += vp[j];
        w += vq[j];
      }
  }
  a = block_tree(u, red);
  b = block_tree(w, red);
}

// final_reduce2 of (P, Q) over T tiles from sums formed in an earlier launch: pre[t] is thread
// t's sum of P over all T tiles and pre[256 + t] its sum of Q over the tiles below T0, both from
// 0.0 in ascending order (pre_sums); thread t then adds Q of its tiles from T0 on in ascending
// order, so a and b equal final_reduce2's bit for bit. LO: the first 256 threads of a larger
// workgroup (final_reduce2_lo).
template <bool LO>
__device__ __forceinline__ void final_reduce2_pre(gptr<const double> pre, gptr<const double> Q, int T0, int T,
                                                  double* red, double& a, double& b) {
  double u = 0.0, w = 0.0;
  if (!LO || threadIdx.x < kTpb) {
    const int t = threadIdx.x;
    u = pre[t];
    w = pre[kTpb + t];
    int m = t;
    if (m < T0) m += ((T0 - m + kTpb - 1) / kTpb) * kTpb;
    for (; m < T; m += 8 * kTpb) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (m + j * kTpb < T) ? Q[m + j * kTpb] : 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (m + j * kTpb < T) w += v[j];
    }
  }
  a = LO ? block_tree_lo(u, red) : block_tree(u, red);
  b = LO ? block_tree_lo(w, red) : block_tree(w, red);
}

// The sums final_reduce2_pre starts from, by threads 0..255 of one workgroup: thread t's sum of
// P over tiles t, t + 256, ... below T and of Q over those below T0, each from 0.0 in that order.
__device__ __forceinline__ void pre_sums(gptr<const double> P, gptr<const double> Q, int T0, int T,
                                         gptr<double> pre) {
  const int t = threadIdx.x;
  double u = 0.0, w = 0.0;
  for (int m = t; m < T; m += 8 * kTpb) {
    double vp[8], vq[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = m + j * kTpb;
      vp[j] = i < T ? P[i] : 0.0;
      vq[j] = i < T0 ? Q[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (m + j * kTpb < T) u += vp[j];
      if (m + j * kTpb < T0) w += vq[j];
    }
  }
  pre[t] = u;
  pre[kTpb + t] = w;
}

// final_reduce / final_reduce2 by the first 256 threads of a larger workgroup, in the same order.
__device__ __forceinline__ double final_reduce_lo(gptr<const double> P, int T, double* red) {
  double u = 0.0;
  if (threadIdx.x < kTpb) {
    for (int m = threadIdx.x; m < T; m += 8 * kTpb) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (m + j * kTpb < T) ? P[m + j * kTpb] : 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (m + j * kTpb < T) u += v[j];
    }
  }
  return block_tree_lo(u, red);
}

__device__ __forceinline__ void final_reduce2_lo(gptr<const double> P, gptr<const double> Q, int T, double* red,
                                                 double& a, double& b) {
  double u = 0.0, w = 0.0;
  if (threadIdx.x < kTpb) {
    for (int m = threadIdx.x; m < T; m += 8 * kTpb) {
      double vp[8], vq[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool in = m + j * kTpb < T;
        vp[j] = in ? P[m + j * kTpb] : 0.0;
        vq[j] = in ? Q[m + j * kTpb] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (m + j * kTpb < T) {
          u += vp[j];
          w += vq[j];
        }
    }
  }
  a = block_tree_lo(u, red);
  b = block_tree_lo(w, red);
}

}  // namespace gsh
