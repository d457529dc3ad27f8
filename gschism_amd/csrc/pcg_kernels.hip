// pcg_kernels.hip -- Jacobi-preconditioned CG for gfx950 in two fused phases per iteration.
//
// The reference issues four launches per iteration (f_update_p, cuSPARSE SpMV, f_dot,
// f_update_xrz; pcg_loop.cpp:367-403) and sums dot products with one atomicAdd per block
// (pcg_solver.py:134-167), so its results are not reproducible. Here every iteration is
//   phase A  beta; p_new = z + beta p_old (own row and, on the fly, every column of the
//            stencil); q = A p_new (rows summed in column order); p.q
//   phase B  alpha; x += alpha p; r -= alpha q; z = invD r; r.r, r.z; stopping test
// with the same scalar recurrences as the reference (f_update_p / f_update_xrz,
// pcg_solver.py:208-276).
//
// Dot products use one fixed tree. A tile of 256*R rows is summed per thread over its R rows
// (thread tau takes rows tile*256R + 256 j + tau), then by a wave64 shuffle tree, then as
// (w0 + w1) + (w2 + w3) over the 4 waves; every consumer workgroup sums the tile partials with
// the same tree (thread t takes partials t, t+256, ...). A single-writer-per-scalar scheme with
// device-scope counters was measured slower: thousands of workgroups serialize on one atomic
// address across the XCDs. Results are bitwise reproducible and independent of how many
// workgroups run concurrently.
//
// A workgroup finds its system and tile through a launch table, so one launch advances any
// number of independent systems (batched and served solves); finished systems drop out
// through their device-side `done` flag.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "device.h"
#include "reduce.h"

namespace gsh {

template <class VT>
struct SysView {
  int n, ntiles;
  gptr<const int32_t> rowptr, col, diag;
  gptr<const int16_t> cdel;
  gptr<const VT> val;
  gptr<const double> b, h0, h1, h2;
  gptr<VT> x, r, z, pa, pb, q, invd;
  gptr<double> part_rr, part_rz, part_pq, part_bb;
  gptr<PcgState> st;
  bool has_hist, has_cdel;
};

template <class VT>
__device__ __forceinline__ SysView<VT> view(const PcgSys<VT>* __restrict__ sys, int i) {
  const PcgSys<VT>& S = sys[i];
  SysView<VT> v;
  v.n = S.n;
  v.ntiles = S.ntiles;
  v.rowptr = G(S.rowptr);
  v.col = G(S.col);
  v.diag = G(S.diag);
  v.has_cdel = S.cdel != nullptr;
  v.cdel = G(S.cdel);
  v.val = G(S.val);
  v.b = G(S.b);
  v.has_hist = S.h0 != nullptr;
  v.h0 = G(S.h0);
  v.h1 = G(S.h1);
  v.h2 = G(S.h2);
  v.x = G(S.x);
  v.r = G(S.r);
  v.z = G(S.z);
  v.pa = G(S.pa);
  v.pb = G(S.pb);
  v.q = G(S.q);
  v.invd = G(S.invd);
  v.part_rr = G(S.part_rr);
  v.part_rz = G(S.part_rz);
  v.part_pq = G(S.part_pq);
  v.part_bb = G(S.part_bb);
  v.st = G(S.st);
  return v;
}

// ------------------------------------------------------------------ state reset
// rtols (device, may be NULL): per-system relative tolerances instead of rtol.
template <class VT>
__global__ void k_pcg_reset(const PcgSys<VT>* sys, int nsys, double rtol, int maxiter, int check_every,
                            const double* __restrict__ rtols) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nsys) return;
  if (rtols) rtol = rtols[sidx];
  gptr<PcgState> st = G(sys[sidx].st);
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

// ------------------------------------------------------------------ CSR row sums (block staged)
// Sums row i = r0 + threadIdx.x of A*v, v(c) given by `vf`, in column order. The chunk's
// products are staged in LDS with coalesced loads of val/col when they fit.
template <class VT, class VF>
__device__ __forceinline__ VT csr_chunk(gptr<const int32_t> rowptr, gptr<const int32_t> col, gptr<const VT> val,
                                        int n, int r0, VT* lds, VF vf) {
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
  const TileRef tr = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const SysView<VT> S = view(sys, tr.sys);
  const int n = S.n;
  const gptr<const double> h0 = S.h0, h1 = S.h1, h2 = S.h2;
  const bool hist = S.has_hist;
  auto xe = [&](int c) -> VT {
    if (!hist) return (VT)0;
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
// p = z + beta p, q = A p, p.q.
//
// Each wave owns the 64 consecutive rows r0w..r0w+63 of a sub-round (thread tau = 64w + lane
// handles row base + 256 j + tau, the rows of the fixed reduction order). The wave loads the
// chunk's nonzeros coalesced (lane l takes entries s0 + l + 64 m) together with the z and p_old
// gathers; the products go to the wave's LDS slice and each lane sums its own row in column
// order. Columns are read as int16 offsets from the chunk's first row when they fit.
constexpr int kWaveCap = 640;   // nonzeros of a 64-row chunk held in registers (10 per lane)
constexpr int kWaveM = kWaveCap / 64;

template <class VT>
struct Chunk {
  int s0, e0;
  VT v[kWaveM], zc[kWaveM], pc[kWaveM];
};

template <class VT>
__device__ __forceinline__ int col_of(const SysView<VT>& S, int kk, int base64) {
  return S.has_cdel ? base64 + (int)S.cdel[kk] : S.col[kk];
}

// s0, e0: the chunk's first entry and the entry after its last (rowptr at r0w and r0w + 64)
template <class VT>
__device__ __forceinline__ void load_chunk(const SysView<VT>& S, gptr<const VT> po, int r0w, int lane, int s0, int e0,
                                           Chunk<VT>& c) {
  c.s0 = s0;
  c.e0 = e0;
  if (c.e0 - c.s0 > kWaveCap) return;
#pragma unroll
  for (int m = 0; m < kWaveM; ++m) {
    const int kk = c.s0 + lane + 64 * m;
    if (kk < c.e0) {
      const int col = col_of(S, kk, r0w);
      c.v[m] = S.val[kk];
      c.zc[m] = S.z[col];
      c.pc[m] = po[col];
    }
  }
}

template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_phase_a(const PcgSys<VT>* __restrict__ sys,
                                                     const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  __shared__ VT lds[4 * kWaveCap];
  const TileRef tr = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const SysView<VT> S = view(sys, tr.sys);
  gptr<PcgState> st = S.st;
  if (load_done(st) >= 0) return;
  const int k = st->k;
  const int n = S.n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const gptr<VT> pn = (k & 1) ? S.pb : S.pa;
  const gptr<const VT> po = (k & 1) ? S.pa : S.pb;
  const int base = tr.tile * kTpb * R;
  // FP64: the row pointers of every sub-round first: they depend on nothing, so each sub-round's
  // chunk then waits for two memory steps (entries, then gathers) instead of three. FP32 loads them
  // per sub-round: its registers would otherwise drop it from 7 to 5 waves per SIMD, which its
  // batched launches of small systems need more.
  constexpr bool kPre = sizeof(VT) == 8;
  int cs0[R], ce0[R], ra[R], rb[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r0w = base + j * kTpb + w * 64;
    cs0[j] = ce0[j] = ra[j] = rb[j] = 0;
    if (kPre && r0w < n) {
      cs0[j] = S.rowptr[r0w];
      ce0[j] = S.rowptr[min(r0w + 64, n)];
      const int i = r0w + lane;
      if (i < n) {
        ra[j] = S.rowptr[i];
        rb[j] = S.rowptr[i + 1];
      }
    }
  }
  // a sub-round's first and end entries and the thread's row range (FP32: loaded here)
  auto rows_of = [&](int j, int r0w) {
    if (!kPre) {
      cs0[j] = S.rowptr[r0w];
      ce0[j] = S.rowptr[min(r0w + 64, n)];
      const int i = r0w + lane;
      if (i < n) {
        ra[j] = S.rowptr[i];
        rb[j] = S.rowptr[i + 1];
      }
    }
  };
  Chunk<VT> c;
  const int r00 = base + w * 64;
  if (r00 < n) {
    rows_of(0, r00);
    load_chunk(S, po, r00, lane, cs0[0], ce0[0], c);
  }
  const double rz_cur = final_reduce(S.part_rz, S.ntiles, red);
  double beta = 0.0;
  if (k > 0) {
    const double rz_old = st->rho[(k - 1) & 1];
    beta = (rz_old == 0.0) ? 0.0 : rz_cur / rz_old;
  }
  if (tr.tile == 0) {
    // workgroup 0: atol = rtol ||b|| at k = 0 (pcg_solver.py:928-930, 1089-1095), otherwise the
    // stopping test of iteration k-1 (pcg_loop.cpp:405-412)
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
  const gptr<const VT> z = S.z;
  VT* my = lds + w * kWaveCap;
  double pq = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r0w = base + j * kTpb + w * 64;
    if (r0w >= n) break;
    if (j > 0) {
      rows_of(j, r0w);
      load_chunk(S, po, r0w, lane, cs0[j], ce0[j], c);
    }
    const int i = r0w + lane;
    VT q = (VT)0;
    if (c.e0 - c.s0 <= kWaveCap) {
#pragma unroll
      for (int m = 0; m < kWaveM; ++m) {
        const int kk = c.s0 + lane + 64 * m;
        if (kk < c.e0) my[kk - c.s0] = c.v[m] * (c.zc[m] + bv * c.pc[m]);
      }
      wave_lds_sync();
      if (i < n) {
        const int a = ra[j], b = rb[j];
        for (int kk = a; kk < b; ++kk) q += my[kk - c.s0];
      }
      wave_lds_sync();
    } else if (i < n) {
      const int a = ra[j], b = rb[j];
      for (int kk = a; kk < b; ++kk) {
        const int col = col_of(S, kk, r0w);
        q += S.val[kk] * (z[col] + bv * po[col]);
      }
    }
    if (i < n) {
      const VT pi = z[i] + bv * po[i];
      pn[i] = pi;
      S.q[i] = q;
      pq += (double)pi * (double)q;
    }
  }
  pq = block_tree(pq, red);
  if (threadIdx.x == 0) S.part_pq[tr.tile] = pq;
}

// ------------------------------------------------------------------ phase B
// f_update_xrz (pcg_solver.py:253-276): alpha = rz / pq (0 if pq == 0). The row loads are
// issued before the scalar reduction, which only the arithmetic waits for.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_pcg_phase_b(const PcgSys<VT>* __restrict__ sys,
                                                     const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  const TileRef tr = tiles[blockIdx.x];
  const SysView<VT> S = view(sys, tr.sys);
  gptr<PcgState> st = S.st;
  if (load_done(st) >= 0) return;
  const int k = st->kk;
  const int n = S.n;
  const gptr<const VT> p = (k & 1) ? S.pb : S.pa;
  const int base = tr.tile * kTpb * R;
  VT xv[R], pv[R], rv[R], qv[R], dv[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = base + j * kTpb + threadIdx.x;
    if (i < n) {
      xv[j] = S.x[i];
      pv[j] = p[i];
      rv[j] = S.r[i];
      qv[j] = S.q[i];
      dv[j] = S.invd[i];
    }
  }
  const double pq = final_reduce(S.part_pq, S.ntiles, red);
  const double rz = st->rho[k & 1];
  const double alpha = (pq != 0.0) ? rz / pq : 0.0;
  if (tr.tile == 0 && threadIdx.x == 0) st->k = k + 1;
  const VT av = (VT)alpha;
  double rr = 0.0, rzs = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = base + j * kTpb + threadIdx.x;
    if (i < n) {
      S.x[i] = xv[j] + av * pv[j];
      const VT ri = rv[j] - av * qv[j];
      const VT zi = dv[j] * ri;
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
  const SysView<VT> S = view(sys, blockIdx.x);
  gptr<PcgState> st = S.st;
  if (load_done(st) >= 0) return;
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
  const SysView<VT> S = view(sys, tr.sys);
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
                                                       const double* __restrict__ xv_, double* __restrict__ rvec_) {
  __shared__ double red[4];
  __shared__ double lds[kSpmvLdsCap];
  const TileRef tr = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const SysView<double> S = view(sys, tr.sys);
  const int n = S.n;
  const gptr<const double> xv = G(xv_);
  const gptr<double> rvec = G(rvec_);
  const bool want_r = rvec_ != nullptr;
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
      if (want_r) rvec[i] = ri;
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

// Batched r = b - A x (x = S.x, r written to S.r) over a launch table; partials of r.r and b.b.
template <int R>
__global__ void __launch_bounds__(256) k_residual_batch(const PcgSys<double>* __restrict__ sys,
                                                        const TileRef* __restrict__ tiles) {
  __shared__ double red[4];
  __shared__ double lds[kSpmvLdsCap];
  const TileRef tr = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const SysView<double> S = view(sys, tr.sys);
  const int n = S.n;
  const gptr<const double> xv = S.x;
  auto xf = [&](int c) -> double { return xv[c]; };
  double rr = 0.0, bb = 0.0;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int r0 = base + j * kTpb;
    if (r0 >= n) break;
    const double q = csr_chunk<double>(S.rowptr, S.col, S.val, n, r0, lds, xf);
    const int i = r0 + threadIdx.x;
    if (i < n) {
      const double bi = S.b[i];
      const double ri = bi - q;
      S.r[i] = ri;
      rr += ri * ri;
      bb += bi * bi;
    }
  }
  rr = block_tree(rr, red);
  bb = block_tree(bb, red);
  if (threadIdx.x == 0) {
    S.part_rr[tr.tile] = rr;
    S.part_bb[tr.tile] = bb;
  }
}

// Per-system totals of the r.r and b.b partials: out[2 i], out[2 i + 1] (one workgroup per system).
__global__ void __launch_bounds__(256) k_reduce_systems(const PcgSys<double>* __restrict__ sys,
                                                        double* __restrict__ out) {
  __shared__ double red[4];
  const PcgSys<double>& S = sys[blockIdx.x];
  double a, b;
  final_reduce2(G((const double*)S.part_rr), G((const double*)S.part_bb), S.ntiles, red, a, b);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = a;
    out[2 * blockIdx.x + 1] = b;
  }
}

// x64 += d32 for every system of a launch table (sys64[i].x, sys32[i].x).
template <int R>
__global__ void __launch_bounds__(256) k_refine_batch(const PcgSys<double>* __restrict__ sys64,
                                                      const PcgSys<float>* __restrict__ sys32,
                                                      const TileRef* __restrict__ tiles) {
  const TileRef tr = tiles[blockIdx.x];
  const gptr<double> x = G(sys64[tr.sys].x);
  const gptr<const float> d = G((const float*)sys32[tr.sys].x);
  const int n = sys64[tr.sys].n;
  const int base = tr.tile * kTpb * R;
  for (int j = 0; j < R; ++j) {
    const int i = base + j * kTpb + threadIdx.x;
    if (i < n) x[i] = x[i] + (double)d[i];
  }
}

__global__ void k_reduce_pair(const double* __restrict__ pa, const double* __restrict__ pb, int T,
                              double* __restrict__ out) {
  __shared__ double red[4];
  double a, b;
  final_reduce2(G(pa), G(pb), T, red, a, b);
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
void launch_pcg_reset(const PcgSys<VT>* sys, int nsys, double rtol, int maxiter, int check_every, hipStream_t s,
                      const double* rtols) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(k_pcg_reset<VT>, dim3((nsys + 63) / 64), dim3(64), 0, s, sys, nsys, rtol, maxiter, check_every,
                     rtols);
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

// Launch with kernel start/end stamps in e0/e1 (hipExtLaunchKernelGGL), for profiling.
#define GSH_DISPATCH_RX(rpt, KER, VT, E0, E1, GRID, STREAM, ...)                                                \
  switch (rpt) {                                                                                             \
    case 1: hipExtLaunchKernelGGL(KER<VT, 1>, dim3(GRID), dim3(kTpb), 0, STREAM, E0, E1, 0, __VA_ARGS__); break;  \
    case 2: hipExtLaunchKernelGGL(KER<VT, 2>, dim3(GRID), dim3(kTpb), 0, STREAM, E0, E1, 0, __VA_ARGS__); break;  \
    case 8: hipExtLaunchKernelGGL(KER<VT, 8>, dim3(GRID), dim3(kTpb), 0, STREAM, E0, E1, 0, __VA_ARGS__); break;  \
    default: hipExtLaunchKernelGGL(KER<VT, 4>, dim3(GRID), dim3(kTpb), 0, STREAM, E0, E1, 0, __VA_ARGS__); break; \
  }

template <class VT>
void launch_pcg_phase(int which, const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s,
                      hipEvent_t e0, hipEvent_t e1) {
  if (ntiles <= 0) return;
  if (!e0 && !e1) {
    if (which == 0) {
      GSH_DISPATCH_R(rpt, k_pcg_phase_a, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
    } else {
      GSH_DISPATCH_R(rpt, k_pcg_phase_b, VT, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles);
    }
  } else if (which == 0) {
    GSH_DISPATCH_RX(rpt, k_pcg_phase_a, VT, e0, e1, ntiles, s, sys, tiles);
  } else {
    GSH_DISPATCH_RX(rpt, k_pcg_phase_b, VT, e0, e1, ntiles, s, sys, tiles);
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

template void launch_pcg_reset<double>(const PcgSys<double>*, int, double, int, int, hipStream_t, const double*);
template void launch_pcg_reset<float>(const PcgSys<float>*, int, double, int, int, hipStream_t, const double*);
template void launch_pcg_init<double>(const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_init<float>(const PcgSys<float>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_iter<double>(const PcgSys<double>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_iter<float>(const PcgSys<float>*, const TileRef*, int, int, hipStream_t);
template void launch_pcg_phase<double>(int, const PcgSys<double>*, const TileRef*, int, int, hipStream_t, hipEvent_t,
                                       hipEvent_t);
template void launch_pcg_phase<float>(int, const PcgSys<float>*, const TileRef*, int, int, hipStream_t, hipEvent_t,
                                      hipEvent_t);
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

void launch_residual_batch(const PcgSys<double>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  switch (rpt) {
    case 1: hipLaunchKernelGGL(k_residual_batch<1>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles); break;
    case 2: hipLaunchKernelGGL(k_residual_batch<2>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles); break;
    case 8: hipLaunchKernelGGL(k_residual_batch<8>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles); break;
    default: hipLaunchKernelGGL(k_residual_batch<4>, dim3(ntiles), dim3(kTpb), 0, s, sys, tiles); break;
  }
}

void launch_reduce_systems(const PcgSys<double>* sys, int nsys, double* out, hipStream_t s) {
  if (nsys > 0) hipLaunchKernelGGL(k_reduce_systems, dim3(nsys), dim3(kTpb), 0, s, sys, out);
}

void launch_refine_batch(const PcgSys<double>* sys64, const PcgSys<float>* sys32, const TileRef* tiles, int ntiles,
                         int rpt, hipStream_t s) {
  if (ntiles <= 0) return;
  switch (rpt) {
    case 1: hipLaunchKernelGGL(k_refine_batch<1>, dim3(ntiles), dim3(kTpb), 0, s, sys64, sys32, tiles); break;
    case 2: hipLaunchKernelGGL(k_refine_batch<2>, dim3(ntiles), dim3(kTpb), 0, s, sys64, sys32, tiles); break;
    case 8: hipLaunchKernelGGL(k_refine_batch<8>, dim3(ntiles), dim3(kTpb), 0, s, sys64, sys32, tiles); break;
    default: hipLaunchKernelGGL(k_refine_batch<4>, dim3(ntiles), dim3(kTpb), 0, s, sys64, sys32, tiles); break;
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
