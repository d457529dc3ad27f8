// block_kernels.hip -- Jacobi PCG of a block of M systems that share one CSR pattern (ensemble
// members of one mesh: per-member matrix values and right-hand sides), on gfx950.
//
// Layout: every vector is [row][member] and the values are [entry][member], so the M members of
// one row are M consecutive words. A workgroup of 256 threads takes a tile of kBlkTile rows;
// thread t works on member m = t % M for the rows tile*kBlkTile + s + S*j (s = t / M, S = 256 / M,
// j = 0 .. kBlkTile/S - 1). With M = 64 a wave is one row of 64 members: the row pointers and
// column indices are one uniform load for all members and the value and vector gathers are
// 512-byte coalesced segments. Each member keeps the reference's recurrences (pcg_solver.py:
// 172-276, 881-947, 1075-1124; pcg_loop.cpp:361-413) and its own stopping rule; a member that has
// stopped keeps its solution and drops out of the loads.
//
// Dot products (FP64) are summed in one fixed order, restated by the oracle (or_pcg_block):
//   a thread adds its rows' terms to 0.0 in row order; the tile partial of member m adds the
//   S threads' sums to 0.0 in slot order; the reduction kernel (1024 threads, RS = 1024 / M
//   slots per member) has slot r add the tile partials r, r + RS, ... to 0.0 and then adds the
//   RS slot sums to 0.0 in slot order.
// Each iteration is four launches: phase A (p update, SpMV, p.q partials), the p.q reduction
// (alpha), phase B (x, r, z updates, r.r and r.z partials) and their reduction (stopping test,
// beta). With 64 members a phase moves gigabytes, so the launch boundaries cost little.
#include <hip/hip_runtime.h>

#include "device.h"
#include "reduce.h"

namespace gsh {

__device__ __forceinline__ void blk_slot(int M, int& s, int& m, int& S) {
  m = (int)threadIdx.x % M;
  s = (int)threadIdx.x / M;
  S = kTpb / M;
}

// tile partial of member m: the S slot sums in slot order (thread slot 0 of each member writes)
__device__ __forceinline__ void blk_tile_partial(double acc, int M, int s, int m, int S, double* lds,
                                                 double* __restrict__ out) {
  lds[threadIdx.x] = acc;
  __syncthreads();
  if (s == 0) {
    double t = 0.0;
    for (int q = 0; q < S; ++q) t += lds[q * M + m];
    *out = t;
  }
}

// ---------------------------------------------------------------- init
// cold start: x = 0, r = b, invD = 1 / diag, z = invD r, p_old = 0; partials of b.b and r.z
template <class VT>
__global__ void __launch_bounds__(256) k_blk_init(BlkSys<VT> S) {
  __shared__ double lds[2][kTpb];
  int s, m, SS;
  blk_slot(S.M, s, m, SS);
  const int M = S.M;
  const int tile = xcd_block(blockIdx.x, gridDim.x);  // neighbouring tiles on one XCD (shared L2)
  double bb = 0.0, rz = 0.0;
  for (int j = 0; j < kBlkTile / SS; ++j) {
    const int i = tile * kBlkTile + s + SS * j;
    if (i >= S.n) break;
    const int64_t e = (int64_t)i * M + m;
    const VT inv = (VT)(1.0 / (double)S.val[(int64_t)S.diag[i] * M + m]);
    const double bi = S.b[e];
    const VT ri = (VT)bi;
    const VT zi = inv * ri;
    S.invd[e] = inv;
    S.x[e] = (VT)0;
    S.r[e] = ri;
    S.z[e] = zi;
    S.pb[e] = (VT)0;
    bb += bi * bi;
    rz += (double)ri * (double)zi;
  }
  blk_tile_partial(bb, M, s, m, SS, lds[0], S.part_a + (int64_t)tile * M + m);
  blk_tile_partial(rz, M, s, m, SS, lds[1], S.part_b + (int64_t)tile * M + m);
}

// ---------------------------------------------------------------- phase A
// beta (set by the previous reduction) -> p = z + beta p_old, q = A p, partials of p.q
template <class VT>
__global__ void __launch_bounds__(256) k_blk_a(BlkSys<VT> S, int it) {
  __shared__ double lds[kTpb];
  int s, m, SS;
  blk_slot(S.M, s, m, SS);
  const int M = S.M;
  const int tile = xcd_block(blockIdx.x, gridDim.x);  // neighbouring tiles on one XCD (shared L2)
  const BlkState st = S.st[m];
  double acc = 0.0;
  if (st.done < 0) {
    const VT beta = (VT)st.beta;
    const VT* __restrict__ po = (it & 1) ? S.pa : S.pb;
    VT* __restrict__ pn = (it & 1) ? S.pb : S.pa;
    for (int j = 0; j < kBlkTile / SS; ++j) {
      const int i = tile * kBlkTile + s + SS * j;
      if (i >= S.n) break;
      const int a = S.rowptr[i], b = S.rowptr[i + 1];
      VT qv = (VT)0;
      for (int kk = a; kk < b; ++kk) {
        const int64_t c = (int64_t)S.col[kk] * M + m;
        qv += S.val[(int64_t)kk * M + m] * (S.z[c] + beta * po[c]);
      }
      const int64_t e = (int64_t)i * M + m;
      const VT pi = S.z[e] + beta * po[e];
      pn[e] = pi;
      S.q[e] = qv;
      acc += (double)pi * (double)qv;
    }
  }
  blk_tile_partial(acc, M, s, m, SS, lds, S.part_a + (int64_t)tile * M + m);
}

// ---------------------------------------------------------------- phase B
// x += alpha p, r -= alpha q, z = invD r; partials of r.r and r.z
template <class VT>
__global__ void __launch_bounds__(256) k_blk_b(BlkSys<VT> S, int it) {
  __shared__ double lds[2][kTpb];
  int s, m, SS;
  blk_slot(S.M, s, m, SS);
  const int M = S.M;
  const int tile = xcd_block(blockIdx.x, gridDim.x);  // neighbouring tiles on one XCD (shared L2)
  const BlkState st = S.st[m];
  double rr = 0.0, rz = 0.0;
  if (st.done < 0) {
    const VT alpha = (VT)st.alpha;
    const VT* __restrict__ p = (it & 1) ? S.pb : S.pa;
    for (int j = 0; j < kBlkTile / SS; ++j) {
      const int i = tile * kBlkTile + s + SS * j;
      if (i >= S.n) break;
      const int64_t e = (int64_t)i * M + m;
      S.x[e] = S.x[e] + alpha * p[e];
      const VT ri = S.r[e] - alpha * S.q[e];
      const VT zi = S.invd[e] * ri;
      S.r[e] = ri;
      S.z[e] = zi;
      rr += (double)ri * (double)ri;
      rz += (double)ri * (double)zi;
    }
  }
  blk_tile_partial(rr, M, s, m, SS, lds[0], S.part_a + (int64_t)tile * M + m);
  blk_tile_partial(rz, M, s, m, SS, lds[1], S.part_b + (int64_t)tile * M + m);
}

// ---------------------------------------------------------------- reductions
// total of member m over the tile partials P[tile][M] in the block order (1024 threads)
__device__ __forceinline__ double blk_total(const double* __restrict__ P, int ntiles, int M, double* lds) {
  const int m = (int)threadIdx.x % M, r = (int)threadIdx.x / M, RS = 1024 / M;
  double u = 0.0;
  for (int t0 = r; t0 < ntiles; t0 += 8 * RS) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (t0 + q * RS < ntiles) ? P[(int64_t)(t0 + q * RS) * M + m] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (t0 + q * RS < ntiles) u += v[q];
  }
  lds[threadIdx.x] = u;
  __syncthreads();
  double t = 0.0;
  if (r == 0)
    for (int q = 0; q < RS; ++q) t += lds[q * M + m];
  __syncthreads();
  return t;
}

// which = 0: after init (atol from b.b, rz); 1: after phase A (alpha); 2: after phase B (stopping
// test of iteration it, beta of the next iteration)
__global__ void __launch_bounds__(1024) k_blk_reduce(BlkState* __restrict__ st, const double* __restrict__ pa,
                                                     const double* __restrict__ pb, int ntiles, int M, int which,
                                                     int it, double rtol, int maxiter, int check_every) {
  __shared__ double lds[1024];
  const double ta = blk_total(pa, ntiles, M, lds);
  const double tb = (which != 1) ? blk_total(pb, ntiles, M, lds) : 0.0;
  if ((int)threadIdx.x >= M) return;
  BlkState& S = st[threadIdx.x];
  if (which == 0) {
    const double atol = rtol * sqrt(ta);
    S.atol2 = atol * atol;
    S.rz_cur = tb;
    S.rz_old = 0.0;
    S.beta = 0.0;
    S.alpha = 0.0;
    S.rr_last = -1.0;
    S.zero = 0;
    S.done = -1;
    if (atol == 0.0) {  // pcg_solver.py:1089-1095: x = 0 after 0 iterations
      S.zero = 1;
      S.done = 0;
    }
    return;
  }
  if (S.done >= 0) return;
  if (which == 1) {
    S.alpha = (ta != 0.0) ? S.rz_cur / ta : 0.0;
    return;
  }
  S.rz_old = S.rz_cur;
  S.rz_cur = tb;
  if ((it + 1) % check_every == 0 || it == maxiter - 1) {
    S.rr_last = ta;
    if (ta <= S.atol2) S.done = it + 1;
  }
  S.beta = (S.rz_old == 0.0) ? 0.0 : S.rz_cur / S.rz_old;
}

// ---------------------------------------------------------------- layout helpers
// member m's CSR values and right-hand side into the block layout
__global__ void k_blk_put(double* __restrict__ vblk, double* __restrict__ bblk, const double* __restrict__ v,
                          const double* __restrict__ b, int64_t nnz, int n, int M, int m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v && i < nnz) vblk[i * M + m] = v[i];
  if (b && i < n) bblk[i * M + m] = b[i];
}

// member m's solution out of the block layout
template <class VT>
__global__ void k_blk_get(double* __restrict__ x, const VT* __restrict__ xblk, int n, int M, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (double)xblk[(int64_t)i * M + m];
}

// ---------------------------------------------------------------- launchers
template <class VT>
void launch_blk_init(const BlkSys<VT>& S, hipStream_t st) {
  hipLaunchKernelGGL(k_blk_init<VT>, dim3(S.ntiles), dim3(kTpb), 0, st, S);
}
template <class VT>
void launch_blk_iter(const BlkSys<VT>& S, int it, double rtol, int maxiter, int check_every, hipStream_t st) {
  hipLaunchKernelGGL(k_blk_a<VT>, dim3(S.ntiles), dim3(kTpb), 0, st, S, it);
  hipLaunchKernelGGL(k_blk_reduce, dim3(1), dim3(1024), 0, st, S.st, S.part_a, S.part_b, S.ntiles, S.M, 1, it, rtol,
                     maxiter, check_every);
  hipLaunchKernelGGL(k_blk_b<VT>, dim3(S.ntiles), dim3(kTpb), 0, st, S, it);
  hipLaunchKernelGGL(k_blk_reduce, dim3(1), dim3(1024), 0, st, S.st, S.part_a, S.part_b, S.ntiles, S.M, 2, it, rtol,
                     maxiter, check_every);
}
template <class VT>
void launch_blk_reduce_init(const BlkSys<VT>& S, double rtol, hipStream_t st) {
  hipLaunchKernelGGL(k_blk_reduce, dim3(1), dim3(1024), 0, st, S.st, S.part_a, S.part_b, S.ntiles, S.M, 0, 0, rtol,
                     1, 1);
}
void launch_blk_put(double* vblk, double* bblk, const double* v, const double* b, int64_t nnz, int n, int M, int m,
                    hipStream_t st) {
  const int64_t cnt = nnz > n ? nnz : n;
  hipLaunchKernelGGL(k_blk_put, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, vblk, bblk, v, b, nnz, n, M, m);
}
template <class VT>
void launch_blk_get(double* x, const VT* xblk, int n, int M, int m, hipStream_t st) {
  hipLaunchKernelGGL(k_blk_get<VT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, xblk, n, M, m);
}

template void launch_blk_init<double>(const BlkSys<double>&, hipStream_t);
template void launch_blk_iter<double>(const BlkSys<double>&, int, double, int, int, hipStream_t);
template void launch_blk_reduce_init<double>(const BlkSys<double>&, double, hipStream_t);
template void launch_blk_get<double>(double*, const double*, int, int, int, hipStream_t);

}  // namespace gsh
