// pcg_kernels.hip -- Jacobi-preconditioned CG for gfx950 in two fused phases per iteration.
//
// The reference issues four launches per iteration (f_update_p, cuSPARSE SpMV, f_dot,
// f_update_xrz; pcg_loop.cpp:367-403) and sums dot products with one atomicAdd per block
// (pcg_solver.py:134-167), so its results are not reproducible. Here every iteration is
//   phase A  beta; p_new = z + beta p_old (own row and, on the fly, every column of the
//            stencil); q = A p_new (CSR chunks staged through LDS, rows summed in column
//            order); per-tile partial of p.q
//   phase B  alpha; x += alpha p; r -= alpha q; z = invD r; per-tile partials of r.r, r.z
// with the same scalar recurrences as the reference (f_update_p / f_update_xrz,
// pcg_solver.py:208-276). Dot products use one fixed tree: a tile of 256*R rows is summed
// per thread over its R rows, then by a wave64 shuffle tree, then across the 4 waves; the
// tile partials are summed by every consumer workgroup with the same tree. Results are
// therefore bitwise reproducible and independent of how many workgroups run concurrently.
//
// A workgroup finds its system and tile through a launch table, so one launch advances any
// number of independent systems (batched and served solves); finished systems drop out
// through their device-side `done` flag.
#include <hip/hip_runtime.h>

#include "device.h"

namespace gsh {

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

// Sum of the tile partials P[0..T): thread t sums P[t], P[t+256], ... then block_tree.
__device__ __forceinline__ double final_reduce(const double* __restrict__ P, int T, double* red) {
  double u = 0.0;
  for (int m = threadIdx.x; m < T; m += kTpb) u += P[m];
  return block_tree(u, red);
}

// ------------------------------------------------------------------ state reset
template <class VT>
__global__ void k_pcg_reset(const PcgSys<VT>* sys, int nsys, double rtol, int maxiter, int check_every) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nsys) return;
  PcgState* st = sys[sidx].st;
  st->k = 0;
  st->kk = 0;
  st->done = -1;
  st->zero = 0;
  st->maxiter = maxiter;
  st->check_every = check_every;
  st->rtol = rtol;
  st->atol2 = 0.0;
  st->rr_last = -1.0;
  st->bb = 0.0;
  st->rho[0] = 0.0;
  st->rho[1] = 0.0;
}

// ------------------------------------------------------------------ CSR row sums
// Sums row i = r0 + threadIdx.x of A*v, v(c) given by `vf`, in column order. The chunk's
// products are staged in LDS with coalesced loads of val/col when they fit.
template <class VT, class VF>
__device__ __forceinline__ VT csr_chunk(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                        const VT* __restrict__ val, int n, int r0, VT* lds, VF vf) {
  const int i = r0 + threadIdx.x;
  const int rend = min(r0 + kTpb, n);
  const int s0 = rowptr[r0], e0 = rowptr[rend];
  VT acc = (VT)0;
  if (e0 - s0 <= kSpmvLdsCap) {
    for (int k = s0 + threadIdx.x; k < e0; k += kTpb) lds[k - s0] = val[k] * vf(col[k]);
    __syncthreads();
    if (i < n) {
      const int a = rowptr[i], b = rowptr[i + 1];
      for (int k = a; k < b; ++k) acc += lds[k - s0];
    }
    __syncthreads();
  } else if (i < n) {
    const int a = rowptr[i], b = rowptr[i + 1];
    for (int k = a; k < b; ++k) acc += val[k] * vf(col[k]);
  }
  return acc;
}

// ------------------------------------------------------------------ init
// extrapolateGuess + computeResiduals + computeInvDiag + f_init (pcg_solver.py:881-938, 172-182):
// x = 3 x0 - 3 x1 + x2; r = b - A x; invD = 1/diag; z = invD r; p_old = 0; partials of r.r, r.z, b.b.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_init(const PcgSys<VT>* __restrict__ sys,
                                                  const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  __shared__ VT lds[kSpmvLdsCap];
  const TileRef tr = tiles[blockIdx.x];
  const PcgSys<VT>& S = sys[tr.sys];
  const int n = S.n;
  const double* h0 = S.h0;
  const double* h1 = S.h1;
  const double* h2 = S.h2;
  auto xe = [&](int c) -> VT {
    if (!h0) return (VT)0;
    return (VT)((3.0 * h0[c] - 3.0 * h1[c]) + h2[c]);
  };
  double rr = 0.0, rz = 0.0, bb = 0.0;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int r0 = base + j * kTpb;
    if (r0 >= n) break;
    const VT q = csr_chunk<VT>(S.rowptr, S.col, S.val, n, r0, lds, xe);
    const int i = r0 + threadIdx.x;
    if (i < n) {
      const VT xi = xe(i);
      S.x[i] = xi;
      const VT inv = (VT)(1.0 / (double)S.val[S.diag[i]]);
      S.invd[i] = inv;
      const double bi = S.b[i];
      const VT ri = (VT)(bi - (double)q);
      const VT zi = inv * ri;
      S.r[i] = ri;
      S.z[i] = zi;
      S.pb[i] = (VT)0;
      rr += (double)ri * (double)ri;
      rz += (double)ri * (double)zi;
      bb += bi * bi;
    }
  }
  rr = block_tree(rr, red);
  rz = block_tree(rz, red);
  bb = block_tree(bb, red);
  if (threadIdx.x == 0) {
    S.part_rr[tr.tile] = rr;
    S.part_rz[tr.tile] = rz;
    S.part_bb[tr.tile] = bb;
  }
}

// ------------------------------------------------------------------ phase A
// f_update_p + SpMV + f_dot (pcg_solver.py:208-249): beta = rz(k-1)/rz(k-2) (0 at k = 0),
// p = z + beta p, q = A p, p.q partial. Workgroup 0 of each system also records rho, runs the
// stopping test of the previous iteration (pcg_loop.cpp:405-412) and, at k = 0,
// atol = rtol ||b|| (pcg_solver.py:928-930, 1089-1095).
template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_phase_a(const PcgSys<VT>* __restrict__ sys,
                                                     const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  __shared__ VT lds[kSpmvLdsCap];
  const TileRef tr = tiles[blockIdx.x];
  const PcgSys<VT>& S = sys[tr.sys];
  PcgState* st = S.st;
  if (*(volatile int32_t*)&st->done >= 0) return;
  const int k = st->k;
  const int n = S.n;
  const double rz_cur = final_reduce(S.part_rz, S.ntiles, red);
  double beta = 0.0;
  if (k > 0) {
    const double rz_old = st->rho[(k - 1) & 1];
    beta = (rz_old == 0.0) ? 0.0 : rz_cur / rz_old;
  }
  if (tr.tile == 0) {
    if (k == 0) {
      const double bb = final_reduce(S.part_bb, S.ntiles, red);
      if (threadIdx.x == 0) {
        const double atol = st->rtol * sqrt(bb);
        st->bb = bb;
        st->atol2 = atol * atol;
        if (atol == 0.0) {
          st->zero = 1;
          st->done = 0;
        }
      }
    } else {
      const int it = k - 1;
      if ((it + 1) % st->check_every == 0 || it == st->maxiter - 1) {
        const double rr = final_reduce(S.part_rr, S.ntiles, red);
        if (threadIdx.x == 0) {
          st->rr_last = rr;
          if (rr <= st->atol2) st->done = k;
        }
      }
    }
    if (threadIdx.x == 0) {
      st->rho[k & 1] = rz_cur;
      st->kk = k;
    }
  }
  const VT bv = (VT)beta;
  VT* __restrict__ pn = (k & 1) ? S.pb : S.pa;
  const VT* __restrict__ po = (k & 1) ? S.pa : S.pb;
  const VT* __restrict__ z = S.z;
  auto pf = [&](int c) -> VT { return z[c] + bv * po[c]; };
  double pq = 0.0;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int r0 = base + j * kTpb;
    if (r0 >= n) break;
    const VT q = csr_chunk<VT>(S.rowptr, S.col, S.val, n, r0, lds, pf);
    const int i = r0 + threadIdx.x;
    if (i < n) {
      const VT pi = pf(i);
      pn[i] = pi;
      S.q[i] = q;
      pq += (double)pi * (double)q;
    }
  }
  pq = block_tree(pq, red);
  if (threadIdx.x == 0) S.part_pq[tr.tile] = pq;
}

// ------------------------------------------------------------------ phase B
// f_update_xrz (pcg_solver.py:253-276): alpha = rz / pq (0 if pq == 0).
template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_phase_b(const PcgSys<VT>* __restrict__ sys,
                                                     const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  const TileRef tr = tiles[blockIdx.x];
  const PcgSys<VT>& S = sys[tr.sys];
  PcgState* st = S.st;
  if (*(volatile int32_t*)&st->done >= 0) return;
  const int k = st->kk;
  const int n = S.n;
  const double pq = final_reduce(S.part_pq, S.ntiles, red);
  const double rz = st->rho[k & 1];
  const double alpha = (pq != 0.0) ? rz / pq : 0.0;
  if (tr.tile == 0 && threadIdx.x == 0) st->k = k + 1;
  const VT av = (VT)alpha;
  const VT* __restrict__ p = (k & 1) ? S.pb : S.pa;
  double rr = 0.0, rzs = 0.0;
  const int base = tr.tile * kTpb * R;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = base + j * kTpb + threadIdx.x;
    if (i < n) {
      S.x[i] += av * p[i];
      const VT ri = S.r[i] - av * S.q[i];
      const VT zi = S.invd[i] * ri;
      S.r[i] = ri;
      S.z[i] = zi;
      rr += (double)ri * (double)ri;
      rzs += (double)ri * (double)zi;
    }
  }
  rr = block_tree(rr, red);
  rzs = block_tree(rzs, red);
  if (threadIdx.x == 0) {
    S.part_rr[tr.tile] = rr;
    S.part_rz[tr.tile] = rzs;
  }
}

// ------------------------------------------------------------------ window-end stopping test
// The test of iteration k-1 normally runs at the start of phase A(k); at the end of a window of
// launches this kernel runs it (one workgroup per system) so the host sees the flag at once.
// Phase A(k) repeats the same test on the same partials, so running both is harmless.
template <class VT>
__global__ void __launch_bounds__(256) k_pcg_check(const PcgSys<VT>* __restrict__ sys) {
  __shared__ double red[4];
  const PcgSys<VT>& S = sys[blockIdx.x];
  PcgState* st = S.st;
  if (*(volatile int32_t*)&st->done >= 0) return;
  const int k = st->k;
  if (k < 1) return;
  const int it = k - 1;
  if (!((it + 1) % st->check_every == 0 || it == st->maxiter - 1)) return;
  const double rr = final_reduce(S.part_rr, S.ntiles, red);
  if (threadIdx.x == 0) {
    st->rr_last = rr;
    if (rr <= st->atol2) st->done = k;
  }
}

// ------------------------------------------------------------------ finish
// atol == 0: x = 0 (pcg_solver.py:1089-1095).
template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_finish(const PcgSys<VT>* __restrict__ sys,
                                                    const TileRef* __restrict__ tiles) {
  const TileRef tr = tiles[blockIdx.x];
  const PcgSys<VT>& S = sys[tr.sys];
  if (!S.st->zero) return;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int i = base + j * kTpb + threadIdx.x;
    if (i < S.n) S.x[i] = (VT)0;
  }
}

// ------------------------------------------------------------------ true residual
// checkSolution (pcg_solver.py:1127-1143): r = b - A x; partials of r.r and b.b.
template <int R>
__global__ void __launch_bounds__(256) k_true_residual(const PcgSys<double>* __restrict__ sys,
                                                       const TileRef* __restrict__ tiles,
                                                       const double* __restrict__ xv, double* __restrict__ rvec) {
  __shared__ double red[4];
  __shared__ double lds[kSpmvLdsCap];
  const TileRef tr = tiles[blockIdx.x];
  const PcgSys<double>& S = sys[tr.sys];
  const int n = S.n;
  auto xf = [&](int c) -> double { return xv[c]; };
  double rr = 0.0, bb = 0.0;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int r0 = base + j * kTpb;
    if (r0 >= n) break;
    const double q = csr_chunk<double>(S.rowptr, S.col, S.val, n, r0, lds, xf);
    const int i = r0 + threadIdx.x;
    if (i < n) {
      const double ri = S.b[i] - q;
      if (rvec) rvec[i] = ri;
      rr += ri * ri;
      bb += S.b[i] * S.b[i];
    }
  }
  rr = block_tree(rr, red);
  bb = block_tree(bb, red);
  if (threadIdx.x == 0) {
    S.part_rr[tr.tile] = rr;
    S.part_bb[tr.tile] = bb;
  }
}

__global__ void k_reduce_pair(const double* __restrict__ pa, const double* __restrict__ pb, int T,
                              double* __restrict__ out) {
  __shared__ double red[4];
  const double a = final_reduce(pa, T, red);
  const double b = final_reduce(pb, T, red);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// ------------------------------------------------------------------ mixed precision helpers
__global__ void k_cast_values(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (float)in[i];
}

__global__ void k_refine_update(double* __restrict__ x, const float* __restrict__ d, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + (double)d[i];
}

__global__ void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

__global__ void k_fill_f64(double* __restrict__ dst, double v, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = v;
}

// ------------------------------------------------------------------ launch wrappers
#define GSH_DISPATCH_R(rpt, KER, VT, ...)                                              \
  switch (rpt) {                                                                       \
    case 1: hipLaunchKernelGGL((KER<VT, 1>), __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL((KER<VT, 2>), __VA_ARGS__); break;                      \
    case 8: hipLaunchKernelGGL((KER<VT, 8>), __VA_ARGS__); break;                      \
    default: hipLaunchKernelGGL((KER<VT, 4>), __VA_ARGS__); break;                     \
  }

template <class VT>
void launch_pcg_reset(const PcgSys<VT>* sys, int nsys, double rtol, int maxiter, int check_every, hipStream_t s) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(k_pcg_reset<VT>, dim3((nsys + 63) / 64), dim3(64), 0, s, sys, nsys, rtol, maxiter, check_every);
}

template <class VT>
void launch_pcg_init(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  GSH_DISPATCH_R(rpt, k_pcg_init, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
}

template <class VT>
void launch_pcg_iter(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  GSH_DISPATCH_R(rpt, k_pcg_phase_a, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
  GSH_DISPATCH_R(rpt, k_pcg_phase_b, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
}

template <class VT>
void launch_pcg_phase(int which, const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  if (which == 0) {
    GSH_DISPATCH_R(rpt, k_pcg_phase_a, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
  } else {
    GSH_DISPATCH_R(rpt, k_pcg_phase_b, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
  }
}

template <class VT>
void launch_pcg_check(const PcgSys<VT>* sys, int nsys, hipStream_t s) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(k_pcg_check<VT>, dim3(nsys), dim3(kTpb), 0, s, sys);
}

template <class VT>
void launch_pcg_finish(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  GSH_DISPATCH_R(rpt, k_pcg_finish, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
}

template void launch_pcg_reset<double>(const PcgSys<double>*, int, double, int, int, hipStream_t);
template void launch_pcg_reset<float>(const PcgSys<float>*, int, double, int, int, hipStream_t);
template void launch_pcg_init<double>(const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_init<float>(const PcgSys<float>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_iter<double>(const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_iter<float>(const PcgSys<float>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_phase<double>(int, const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_phase<float>(int, const PcgSys<float>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_check<double>(const PcgSys<double>*, int, hipStream_t);
template void launch_pcg_check<float>(const PcgSys<float>*, int, hipStream_t);
template void launch_pcg_finish<double>(const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_finish<float>(const PcgSys<float>*, const TileRef*, int, int, hipStream_t);

void launch_true_residual(const PcgSys<double>* sys, const TileRef* tiles, int ntiles, int rpt, const double* xv,
                          double* rvec, hipStream_t s) {
  if (ntiles <= 0) return;
  switch (rpt) {
    case 1: hipLaunchKernelGGL(k_true_residual<1>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles, xv, rvec); break;
    case 2: hipLaunchKernelGGL(k_true_residual<2>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles, xv, rvec); break;
    case 8: hipLaunchKernelGGL(k_true_residual<8>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles, xv, rvec); break;
    default: hipLaunchKernelGGL(k_true_residual<4>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles, xv, rvec); break;
  }
}

void launch_reduce_pair(const double* pa, const double* pb, int ntiles, double* out2, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_pair, dim3(1), dim3(kTpb), 0, s, pa, pb, ntiles, out2);
}

void launch_cast_values(const double* in, float* out, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_cast_values, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, in, out, count);
}

void launch_refine_update(double* x64, const float* d32, int32_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_refine_update, dim3((n + 255) / 256), dim3(256), 0, s, x64, d32, n);
}

void launch_copy_f64(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dst, src, count);
}

void launch_fill_f64(double* dst, double v, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dst, v, count);
}

}  // namespace gsh
