// ssor_kernels.hip -- PCG with the multicolour SSOR(omega) preconditioner on gfx950.
//
// The reference solves with Jacobi PCG only (pcg_solver.py:933-938). This solver keeps the
// reference's driver (extrapolated guess, atol = rtol ||b||, the stopping rule on ||b - A x||^2
// every check_every iterations; pcg_solver.py:881-947, 1075-1124, pcg_loop.cpp:361-413) and
// preconditions with
//   M = (D~ + L) D~^-1 (D~ + U),   D~ = D / omega,
// in a multicolour order, applied by Eisenstat's trick (SIAM J. Sci. Stat. Comput. 2 (1981)
// 1-4): CG runs on A^ = (D~ + L)^-1 A (D~ + U)^-1 with the diagonal preconditioner D~ and
//   A^ v = t + (D~ + L)^-1 (v - (2 D~ - D) t),   t = (D~ + U)^-1 v,
// so one iteration reads every off-diagonal entry once. x = (D~ + U)^-1 y maps the iterate
// back; the true residual b - A x is (D~ + L) r.
//
// CG runs in the Chronopoulos-Gear form (J. Comput. Appl. Math. 25 (1989) 153-168): with
// z = D~ r and s = A^ z, gamma = r.z and delta = z.s come from one reduction,
//   beta = gamma / gamma_prev,  alpha = gamma / (delta - beta gamma / alpha_prev),
//   p = z + beta p,  q = s + beta q,  y += alpha p,  r -= alpha q,
// so every vector update is row-local and folds into the next backward sweep. Rows are in
// colour-major order; every colour is a set of independent rows, so each triangular sweep is
// one launch per colour and an iteration is 2C - 1 launches:
//   bwd(C-1) ... bwd(0)   alpha, beta (first); row updates; z = D~ r; t = (z - U t) / D~;
//                         gamma partials; colour 0 (no lower colours) also u and s below
//   fwd(1) ... fwd(C-1)   u = (z - K t - L u) / D~, s = t + u (K = 2 D~ - D); delta partials
// Colour blocks are padded to whole reduction tiles, so each tile's partial is written by
// exactly one launch and every dot product uses the fixed tile tree of the Jacobi solver.
// Row sums run over ascending columns from 0.0, in the oracle's order. The two triangular parts
// are ELL per colour: a wave reads entry m of its 64 rows with one coalesced load, so a row sum
// is two dependent memory steps (columns, then values and gathers together).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "device.h"
#include "reduce.h"

namespace gsh {

template <class VT>
struct SsorView {
  int np, ntiles;
  double omega;
  gptr<const SsorColour> colours;
  gptr<const int32_t> tile_colour, lcol, ucol, lsrc, usrc, dsrc, perm;
  gptr<const VT> aval;
  gptr<const double> b, h0, h1, h2;
  gptr<VT> xout, lval, uval, d, dt, kd, bp, y, r, p, q, t, u, s, w;
  gptr<double> part_rr, part_rz, part_pq, part_bb, pre;
  gptr<PcgState> st;
  bool has_hist, unit_omega;
  int pre_t0;
};

template <class VT>
__device__ __forceinline__ SsorView<VT> sview(const SsorSysT<VT>* __restrict__ sys) {
  const SsorSysT<VT>& S = *sys;
  SsorView<VT> v;
  v.np = S.np;
  v.ntiles = S.ntiles;
  v.omega = S.omega;
  v.unit_omega = S.unit_omega != 0;
  v.colours = G(S.colours);
  v.tile_colour = G(S.tile_colour);
  v.lcol = G(S.lcol);
  v.ucol = G(S.ucol);
  v.lsrc = G(S.lsrc);
  v.usrc = G(S.usrc);
  v.dsrc = G(S.dsrc);
  v.perm = G(S.perm);
  v.aval = G(S.aval);
  v.b = G(S.b);
  v.has_hist = S.h0 != nullptr;
  v.h0 = G(S.h0);
  v.h1 = G(S.h1);
  v.h2 = G(S.h2);
  v.xout = G(S.xout);
  v.lval = G(S.lval);
  v.uval = G(S.uval);
  v.d = G(S.d);
  v.dt = G(S.dt);
  v.kd = G(S.kd);
  v.bp = G(S.bp);
  v.y = G(S.y);
  v.r = G(S.r);
  v.p = G(S.p);
  v.t = G(S.t);
  v.u = G(S.u);
  v.s = G(S.s);
  v.q = G(S.q);
  v.w = G(S.w);
  v.part_rr = G(S.part_rr);
  v.part_rz = G(S.part_rz);
  v.part_pq = G(S.part_pq);
  v.part_bb = G(S.part_bb);
  v.pre = G(S.pre);
  v.pre_t0 = S.pre_t0;
  v.st = G(S.st);
  return v;
}

template <class T>
__device__ __forceinline__ gptr<const T> C(gptr<T> p) {
  return (gptr<const T>)p;
}

// row of thread threadIdx.x in sub-round j of a tile (the reduction order of every partial)
template <int R>
__device__ __forceinline__ int tile_row(int tile, int j) {
  return tile * kTpb * R + j * kTpb + threadIdx.x;
}

// Wide workgroups (W): 256 * R threads with one row each, thread i taking row i of the tile, i.e.
// row tile_row<R>(tile, i / 256) of lane i % 256; every partial is summed in the R-rows-per-
// thread order (tile_tree_wide), so results equal the 256-thread kernels' bit for bit.
template <int R, bool W>
__device__ __forceinline__ int srow(int tile, int j) {
  return W ? tile * kTpb * R + (int)threadIdx.x : tile_row<R>(tile, j);
}
// a thread's share of a partial: its one row (W) or the running sum over its R rows
template <bool W>
__device__ __forceinline__ void sacc(double& a, double x) {
  if (W)
    a = x;
  else
    a += x;
}
template <int R, bool W>
__device__ __forceinline__ double stree(double v, double* lrows, double* red) {
  if (W) return tile_tree_wide<R>(v, lrows, red);
  return block_tree(v, red);
}

// Row-vector stores and loads of the sweeps. Plain (cached) accesses: nontemporal stores, with or
// without nontemporal loads of the streamed rows, measured slower on the 1M-node system (83.6 and
// 83.2 against 88.5 solves/s at one row per thread).
template <class T>
__device__ __forceinline__ void vst(gptr<T> p, int i, T v) {
  p[i] = v;
}
template <class T>
__device__ __forceinline__ T vld(gptr<T> p, int i) {
  return p[i];
}

struct ColourInfo {
  int start, rows, wl, wu, offl, offu;
};

template <class VT>
__device__ __forceinline__ ColourInfo colour_info(const SsorView<VT>& S, int c) {
  const gptr<const int32_t> k = (gptr<const int32_t>)(S.colours + c);
  return {k[0], k[1], k[2], k[3], k[4], k[5]};
}

// sum over one ELL row part of a_vj x_j in entry (= ascending column) order, from 0.0; the
// columns of 8 entries are loaded together, then their values and gathers together
constexpr int kEllChunk = 8;

template <class VT>
__device__ __forceinline__ VT ell_sum(gptr<const int32_t> ec, gptr<const VT> ev, gptr<const VT> x, int base,
                                      int stride, int width) {
  VT acc = 0;
  for (int m0 = 0; m0 < width; m0 += kEllChunk) {
    int c[kEllChunk];
    VT vv[kEllChunk], xx[kEllChunk];
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m) c[m] = (m0 + m < width) ? ec[base + (m0 + m) * stride] : -1;
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m)
      if (c[m] >= 0) {
        vv[m] = ev[base + (m0 + m) * stride];
        xx[m] = x[c[m]];
      }
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m)
      if (c[m] >= 0) acc += vv[m] * xx[m];
  }
  return acc;
}

// ell_sum of one row part against two vectors at once (each sum in entry order, from 0.0): the
// columns and values are loaded once for both.
template <class VT>
__device__ __forceinline__ void ell_sum2(gptr<const int32_t> ec, gptr<const VT> ev, gptr<const VT> x, gptr<const VT> y,
                                         int base, int stride, int width, VT& ax, VT& ay) {
  VT sx = 0, sy = 0;
  for (int m0 = 0; m0 < width; m0 += kEllChunk) {
    int c[kEllChunk];
    VT vv[kEllChunk], xx[kEllChunk], yy[kEllChunk];
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m) c[m] = (m0 + m < width) ? ec[base + (m0 + m) * stride] : -1;
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m)
      if (c[m] >= 0) {
        vv[m] = ev[base + (m0 + m) * stride];
        xx[m] = x[c[m]];
        yy[m] = y[c[m]];
      }
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m)
      if (c[m] >= 0) {
        sx += vv[m] * xx[m];
        sy += vv[m] * yy[m];
      }
  }
  ax = sx;
  ay = sy;
}

template <class VT>
__device__ __forceinline__ VT lower_sum(const SsorView<VT>& S, const ColourInfo& ci, int v, gptr<const VT> x) {
  return ell_sum(S.lcol, C(S.lval), x, ci.offl + (v - ci.start), ci.rows, ci.wl);
}

template <class VT>
__device__ __forceinline__ VT upper_sum(const SsorView<VT>& S, const ColourInfo& ci, int v, gptr<const VT> x) {
  return ell_sum(S.ucol, C(S.uval), x, ci.offu + (v - ci.start), ci.rows, ci.wu);
}

// ------------------------------------------------------------------ setup
// Gathers the ELL values and the diagonal from the CSR values, b and the extrapolated guess
// x0 = 3 x0 - 3 x1 + x2 (pcg_solver.py:881-886) into the colour-major order; b.b partials.
// ELL values of one row part from the CSR values: the slots of 8 entries are loaded together,
// then their values, then the stores (a store between the loads would order every load after it).
template <class VT>
__device__ __forceinline__ void ell_gather(gptr<const int32_t> src, gptr<const VT> aval, gptr<VT> out, int base,
                                           int stride, int width) {
  for (int m0 = 0; m0 < width; m0 += kEllChunk) {
    int k[kEllChunk], sl[kEllChunk];
    VT v[kEllChunk];
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m) {
      k[m] = base + (m0 + m) * stride;
      sl[m] = (m0 + m < width) ? src[k[m]] : -1;
    }
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m) v[m] = sl[m] >= 0 ? aval[sl[m]] : (VT)0;
#pragma unroll
    for (int m = 0; m < kEllChunk; ++m)
      if (m0 + m < width) out[k[m]] = v[m];
  }
}

// gather = 0 keeps the ELL values and diagonal of the previous set-up (the inner solves of one
// mixed-precision solve share their matrix) and only loads b and the guess.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_setup(const SsorSysT<VT>* __restrict__ sys, int gather) {
  __shared__ double red[4];
  const SsorView<VT> S = sview(sys);
  const int tile = blockIdx.x;
  const ColourInfo ci = colour_info(S, S.tile_colour[tile]);
  double bb = 0.0;
  for (int j = 0; j < R; ++j) {
    const int v = tile_row<R>(tile, j);
    if (gather) {
      const int ds = S.dsrc[v];
      const VT dv = ds >= 0 ? S.aval[ds] : (VT)1;
      const VT dtv = dv / (VT)S.omega;
      S.d[v] = dv;
      S.dt[v] = dtv;
      S.kd[v] = (VT)2 * dtv - dv;
    }
    const int ro = v - ci.start;
    if (gather) {
      ell_gather(S.lsrc, S.aval, S.lval, ci.offl + ro, ci.rows, ci.wl);
      ell_gather(S.usrc, S.aval, S.uval, ci.offu + ro, ci.rows, ci.wu);
    }
    const int i = S.perm[v];
    const double bv = i >= 0 ? S.b[i] : 0.0;
    S.bp[v] = (VT)bv;
    S.w[v] = (VT)((i >= 0 && S.has_hist) ? (3.0 * S.h0[i] - 3.0 * S.h1[i]) + S.h2[i] : 0.0);
    bb += bv * bv;
  }
  bb = block_tree(bb, red);
  if (threadIdx.x == 0) S.part_bb[tile] = bb;
}

// r_true = b - A x0 (computeResiduals, pcg_solver.py:889-895) into t; y0 = (D~ + U) x0; p = 0.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_residual(const SsorSysT<VT>* __restrict__ sys) {
  const SsorView<VT> S = sview(sys);
  const ColourInfo ci = colour_info(S, S.tile_colour[blockIdx.x]);
  const gptr<const VT> x0 = C(S.w);
  for (int j = 0; j < R; ++j) {
    const int v = tile_row<R>(blockIdx.x, j);
    const VT xv = x0[v];
    const VT au = upper_sum(S, ci, v, x0);
    const VT ax = (lower_sum(S, ci, v, x0) + S.d[v] * xv) + au;
    S.t[v] = S.bp[v] - ax;
    S.y[v] = S.dt[v] * xv + au;
    S.p[v] = 0;
    S.q[v] = 0;
  }
}

// r0 = (D~ + L)^-1 r_true for the rows of one colour.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_r0(const SsorSysT<VT>* __restrict__ sys, int tile0, int c) {
  const SsorView<VT> S = sview(sys);
  const ColourInfo ci = colour_info(S, c);
  const int tile = tile0 + xcd_block(blockIdx.x, gridDim.x);
  const gptr<const VT> r = C(S.r);
  for (int j = 0; j < R; ++j) {
    const int v = tile_row<R>(tile, j);
    S.r[v] = (S.t[v] - lower_sum(S, ci, v, r)) / S.dt[v];
  }
}

// ------------------------------------------------------------------ backward sweep
// kPre: the first sweep starts from the sums of pre_sums (S.pre) instead of all tile partials;
// kPreWG: workgroup 0 of this launch (the last sweep of an iteration) forms those sums for the
// next iteration and the sweep's tiles go to workgroups 1 and up.
// kChk: the iteration closes a window; its sweeps also write the partials of ||(D~ + L) r||^2
// of their rows (colour 0 in the backward sweep, the others in the forward sweeps): per tile,
// thread t's rows' (D~ r + L r)^2 summed in row order, then the block tree.
constexpr int kFirst = 1, kMid = 2, kLast = 4, kInit = 8, kLead = 16, kPre = 32, kPreWG = 64, kChk = 128;

// One colour of the backward sweep. Outside the set-up pass the first sweep of an iteration
// (workgroup 0) turns the gamma and delta partials into alpha and beta and keeps them with
// gamma and the iteration index; every sweep then applies p = z + beta p, q = s + beta q,
// y += alpha p, r -= alpha q to its rows before t = (z - U t) / D~ with the new z = D~ r. The
// sweep of colour 0 (no lower colours) also forms u and s = t + u for its rows.
// lead: the one workgroup of the launch (per system) that writes the system's scalar state.
template <class VT, int R, bool W = false>
__device__ __forceinline__ void ssor_bwd_body(const SsorView<VT>& S, int tile, int c, int flags, bool lead,
                                              double* red, double* lrows) {
  constexpr int J = W ? 1 : R;  // rows of this thread
  gptr<PcgState> st = S.st;
  const bool init = flags & kInit, mid = flags & kMid;
  const ColourInfo ci = colour_info(S, c);
  const gptr<const VT> t = C(S.t);
  VT usum[J], dtv[J], rv[J], pv[J], qv[J], sv[J], yv[J], rsv[J];
  // the row loads are issued before the done test so that their latency overlaps it
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int v = srow<R, W>(tile, j);
    dtv[j] = vld(S.dt, v);
    rv[j] = vld(S.r, v);
    if (!init) {
      pv[j] = vld(S.p, v);
      qv[j] = vld(S.q, v);
      sv[j] = vld(S.s, v);
    }
  }
  if (load_done(st) >= 0) return;
  // y takes its updates in pairs: an odd iteration k adds alpha_(k-1) p_(k-1) (p as loaded)
  // and then alpha_k p_k, the same two roundings as one update per iteration, so y is read
  // and written every other iteration; k_ssor_xsolve adds a last update still pending
  const int kit = init ? 0 : ((flags & kFirst) ? st->k : st->kk);
  const bool ypair = !init && (kit & 1);
  if (ypair) {
#pragma unroll
    for (int j = 0; j < J; ++j) yv[j] = vld(S.y, srow<R, W>(tile, j));
  }
  // the gamma / delta partials of iteration k live in slot k & 1, so the first sweep can read
  // slot k & 1 while the sweeps write the partials of iteration k + 1 into the other slot
  double alpha = 0.0, beta = 0.0, aprev = 0.0;
  int wslot = 0;
  if (!init) {
    if (flags & kFirst) {
      const int k = kit;
      wslot = (k + 1) & 1;
      double gamma, delta;
      if (flags & kPre)
        final_reduce2_pre<W>(C(S.pre) + (k & 1) * 2 * kTpb, C(S.part_pq) + (k & 1) * S.ntiles, S.pre_t0, S.ntiles,
                             red, gamma, delta);
      else if (W)
        final_reduce2_lo(C(S.part_rz) + (k & 1) * S.ntiles, C(S.part_pq) + (k & 1) * S.ntiles, S.ntiles, red, gamma,
                         delta);
      else
        final_reduce2(C(S.part_rz) + (k & 1) * S.ntiles, C(S.part_pq) + (k & 1) * S.ntiles, S.ntiles, red, gamma,
                      delta);
      if (k == 0) {
        alpha = (delta != 0.0) ? gamma / delta : 0.0;
        if (lead) {
          const double bb = W ? final_reduce_lo(C(S.part_bb), S.ntiles, red) : final_reduce(C(S.part_bb), S.ntiles, red);
          if (threadIdx.x == 0) {
            const double atol = st->rtol * sqrt(bb);
            st->bb = bb;
            st->atol2 = atol * atol;
            if (atol == 0.0) {
              st->zero = 1;
              st->done = 0;
            }
          }
        }
      } else {
        const double gprev = st->rho[(k - 1) & 1];
        aprev = st->alpha[(k - 1) & 1];
        beta = (gprev != 0.0) ? gamma / gprev : 0.0;
        const double den = delta - beta * gamma / aprev;
        alpha = (aprev != 0.0 && den != 0.0) ? gamma / den : 0.0;
      }
      if (lead && threadIdx.x == 0) {
        st->rho[k & 1] = gamma;
        st->alpha[k & 1] = alpha;
        st->beta = beta;
        st->kk = k;
      }
    } else {
      const int kk = kit;
      alpha = st->alpha[kk & 1];
      if (ypair) aprev = st->alpha[(kk - 1) & 1];
      beta = st->beta;
      wslot = (kk + 1) & 1;
    }
  }
#pragma unroll
  for (int j = 0; j < J; ++j) usum[j] = upper_sum(S, ci, srow<R, W>(tile, j), t);
  const VT av = (VT)alpha, bv = (VT)beta, apv = (VT)aprev;
  double g = 0.0, dl = 0.0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int v = srow<R, W>(tile, j);
    VT r1 = rv[j];
    if (!init) {
      const VT zr = dtv[j] * rv[j];
      const VT pn = zr + bv * pv[j];
      const VT qn = sv[j] + bv * qv[j];
      vst(S.p, v, pn);
      vst(S.q, v, qn);
      if (ypair) vst(S.y, v, (VT)((VT)(yv[j] + apv * pv[j]) + av * pn));
      r1 = rv[j] - av * qn;
      vst(S.r, v, r1);
    }
    rsv[j] = r1;
    const VT zn = dtv[j] * r1;
    sacc<W>(g, (double)r1 * (double)zn);
    const VT tv = (zn - usum[j]) / dtv[j];
    // t of colour 0 is used only here: no lower colour gathers it and no forward sweep reads it
    if (!mid) vst(S.t, v, tv);
    if (mid) {
      const VT kdv = S.unit_omega ? dtv[j] : S.kd[v];
      const VT uv = ((zn - kdv * tv) - (VT)0) / dtv[j];
      vst(S.u, v, uv);
      const VT sn = tv + uv;
      vst(S.s, v, sn);
      sacc<W>(dl, (double)zn * (double)sn);
    }
  }
  g = stree<R, W>(g, lrows, red);
  if (threadIdx.x == 0) S.part_rz[wslot * S.ntiles + tile] = g;
  if (mid && (flags & kChk)) {
    // colour 0 has no lower part: (D~ + L) r is D~ r
    double rr = 0.0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const VT wv = dtv[j] * rsv[j] + (VT)0;
      sacc<W>(rr, (double)wv * (double)wv);
    }
    rr = stree<R, W>(rr, lrows, red);
    if (threadIdx.x == 0) S.part_rr[tile] = rr;
  }
  if (mid) {
    dl = stree<R, W>(dl, lrows, red);
    if (threadIdx.x == 0) S.part_pq[wslot * S.ntiles + tile] = dl;
    if ((flags & kLast) && !init && lead && threadIdx.x == 0) st->k = st->kk + 1;
  }
}

template <class VT, int R, bool W = false>
__global__ void __launch_bounds__(W ? kTpb * R : kTpb)
    k_ssor_bwd(const SsorSysT<VT>* __restrict__ sys, int tile0, int c, int flags) {
  __shared__ double red[4];
  __shared__ double lrows[W ? kTpb * R : 1];
  ssor_bwd_body<VT, R, W>(sview(sys), tile0 + xcd_block(blockIdx.x, gridDim.x), c, flags, blockIdx.x == 0, red,
                          lrows);
}

// Several systems in one launch: entry blockIdx of the table names the system, tile, colour and
// flags (kLead marks each system's lead workgroup); or_flags is added to every entry.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_bwd_batch(const SsorSysT<VT>* __restrict__ sys,
                                                        const SsorRef* __restrict__ refs, int or_flags) {
  __shared__ double red[4];
  const SsorRef ref = refs[xcd_block(blockIdx.x, gridDim.x)];
  const int f = ref.flags | or_flags;
  ssor_bwd_body<VT, R>(sview(sys + ref.sys), ref.tile, ref.c, f, (f & kLead) != 0, red, nullptr);
}

// ------------------------------------------------------------------ forward sweep
// u = (D~ + L)^-1 (z - K t) for one colour, s = t + u = A^ z, partials of z.s. The last sweep of
// an iteration advances the iteration counter.
// CK: the check iteration's sweep (kChk), a separate instantiation so that the other sweeps keep
// their register count.
template <class VT, int R, bool W = false, bool CK = false>
__device__ __forceinline__ void ssor_fwd_body(const SsorView<VT>& S, int tile, int c, int flags, bool lead,
                                              double* red, double* lrows) {
  constexpr int J = W ? 1 : R;  // rows of this thread
  const ColourInfo ci = colour_info(S, c);
  const gptr<const VT> u = C(S.u);
  VT lsum[J], dtv[J], tv[J], rv[J], kdv[J];
  // the row loads are issued before the done test so that their latency overlaps it
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int v = srow<R, W>(tile, j);
    dtv[j] = vld(S.dt, v);
    tv[j] = vld(S.t, v);
    rv[j] = vld(S.r, v);
    kdv[j] = S.unit_omega ? dtv[j] : vld(S.kd, v);
  }
  if (load_done(S.st) >= 0) return;
  const int wslot = (flags & kInit) ? 0 : ((S.st->kk + 1) & 1);
  VT rlo[J];  // (L r) of the rows (CK)
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int v = srow<R, W>(tile, j);
    if (CK)
      ell_sum2(S.lcol, C(S.lval), u, C(S.r), ci.offl + (v - ci.start), ci.rows, ci.wl, lsum[j], rlo[j]);
    else
      lsum[j] = lower_sum(S, ci, v, u);
  }
  double dl = 0.0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int v = srow<R, W>(tile, j);
    const VT zn = dtv[j] * rv[j];
    const VT uv = ((zn - kdv[j] * tv[j]) - lsum[j]) / dtv[j];
    // u of the last colour is gathered by no higher colour
    if (!(flags & kLast)) vst(S.u, v, uv);
    const VT sn = tv[j] + uv;
    vst(S.s, v, sn);
    sacc<W>(dl, (double)zn * (double)sn);
  }
  dl = stree<R, W>(dl, lrows, red);
  if (threadIdx.x == 0) S.part_pq[wslot * S.ntiles + tile] = dl;
  if (CK) {
    double rr = 0.0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const VT wv = dtv[j] * rv[j] + rlo[j];
      sacc<W>(rr, (double)wv * (double)wv);
    }
    rr = stree<R, W>(rr, lrows, red);
    if (threadIdx.x == 0) S.part_rr[tile] = rr;
  }
  if ((flags & kLast) && !(flags & kInit) && lead && threadIdx.x == 0) S.st->k = S.st->kk + 1;
}

template <class VT, int R, bool W, bool CK>
__device__ __forceinline__ void ssor_fwd_kernel(const SsorSysT<VT>* __restrict__ sys, int tile0, int c, int flags,
                                                double* red, double* lrows) {
  if (flags & kPreWG) {
    // every gamma partial and the delta partials of the colours before this one are final (the
    // sweeps of earlier launches wrote them): workgroup 0 forms the next first sweep's sums
    const SsorView<VT> S = sview(sys);
    if (blockIdx.x == 0) {
      if (load_done(S.st) >= 0) return;
      const int slot = (flags & kInit) ? 0 : ((S.st->kk + 1) & 1);
      if (threadIdx.x < kTpb)
        pre_sums(C(S.part_rz) + slot * S.ntiles, C(S.part_pq) + slot * S.ntiles, S.pre_t0, S.ntiles,
                 S.pre + slot * 2 * kTpb);
      return;
    }
    ssor_fwd_body<VT, R, W, CK>(S, tile0 + xcd_block(blockIdx.x - 1, gridDim.x - 1), c, flags, blockIdx.x == 1, red,
                                lrows);
    return;
  }
  ssor_fwd_body<VT, R, W, CK>(sview(sys), tile0 + xcd_block(blockIdx.x, gridDim.x), c, flags, blockIdx.x == 0, red,
                              lrows);
}

template <class VT, int R, bool W = false>
__global__ void __launch_bounds__(W ? kTpb * R : kTpb)
    k_ssor_fwd(const SsorSysT<VT>* __restrict__ sys, int tile0, int c, int flags) {
  __shared__ double red[4];
  __shared__ double lrows[W ? kTpb * R : 1];
  ssor_fwd_kernel<VT, R, W, false>(sys, tile0, c, flags, red, lrows);
}

// the check iteration's forward sweeps (kChk)
template <class VT, int R, bool W = false>
__global__ void __launch_bounds__(W ? kTpb * R : kTpb)
    k_ssor_fwd_chk(const SsorSysT<VT>* __restrict__ sys, int tile0, int c, int flags) {
  __shared__ double red[4];
  __shared__ double lrows[W ? kTpb * R : 1];
  ssor_fwd_kernel<VT, R, W, true>(sys, tile0, c, flags, red, lrows);
}

// An entry with tile -1 (kPreWG) forms its system's pre_sums in the step of the system's last
// forward sweep (whose other partials are final by then).
template <class VT, int R, bool CK>
__device__ __forceinline__ void ssor_fwd_batch_entry(const SsorSysT<VT>* __restrict__ sys,
                                                     const SsorRef* __restrict__ refs, int or_flags, double* red) {
  const SsorRef ref = refs[xcd_block(blockIdx.x, gridDim.x)];
  const int f = ref.flags | or_flags;
  const SsorView<VT> S = sview(sys + ref.sys);
  if (ref.tile < 0) {
    if (load_done(S.st) >= 0) return;
    const int slot = (f & kInit) ? 0 : ((S.st->kk + 1) & 1);
    pre_sums(C(S.part_rz) + slot * S.ntiles, C(S.part_pq) + slot * S.ntiles, S.pre_t0, S.ntiles,
             S.pre + slot * 2 * kTpb);
    return;
  }
  ssor_fwd_body<VT, R, false, CK>(S, ref.tile, ref.c, f, (f & kLead) != 0, red, nullptr);
}

template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_fwd_batch(const SsorSysT<VT>* __restrict__ sys,
                                                        const SsorRef* __restrict__ refs, int or_flags) {
  __shared__ double red[4];
  ssor_fwd_batch_entry<VT, R, false>(sys, refs, or_flags, red);
}

// the check iteration's forward steps (kChk in or_flags)
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_fwd_batch_chk(const SsorSysT<VT>* __restrict__ sys,
                                                            const SsorRef* __restrict__ refs, int or_flags) {
  __shared__ double red[4];
  ssor_fwd_batch_entry<VT, R, true>(sys, refs, or_flags, red);
}

// ------------------------------------------------------------------ stopping test
// ||b - A x||^2 = ||(D~ + L) r||^2 from the tile partials that the sweeps of a window's last
// iteration write (kChk), then the test of iteration k-1 (pcg_loop.cpp:405-412) in one workgroup.
// workgroup b tests system b (one system: b = 0)
template <class VT>
__global__ void __launch_bounds__(256) k_ssor_check(const SsorSysT<VT>* __restrict__ sys) {
  __shared__ double red[4];
  const SsorView<VT> S = sview(sys + blockIdx.x);
  gptr<PcgState> st = S.st;
  if (load_done(st) >= 0) return;
  const int k = st->k;
  if (k < 1) return;
  const int it = k - 1;
  if (!((it + 1) % st->check_every == 0 || it == st->maxiter - 1)) return;
  const double rr = final_reduce(C(S.part_rr), S.ntiles, red);
  if (threadIdx.x == 0) {
    st->rr_last = rr;
    if (rr <= st->atol2) st->done = k;
  }
}

// ------------------------------------------------------------------ back to x
// x = (D~ + U)^-1 y for one colour (into w), then scattered to the original order; x = 0 when
// atol == 0 (pcg_solver.py:1089-1095). When the last iteration run, K = k - 1, is even, its
// update y += alpha_K p_K is still pending (the sweeps update y in pairs) and is added here.
template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_xsolve(const SsorSysT<VT>* __restrict__ sys, int tile0, int c) {
  const SsorView<VT> S = sview(sys);
  if (S.st->zero) return;
  const int K = S.st->k - 1;
  const bool pend = K >= 0 && !(K & 1);
  const VT aK = pend ? (VT)S.st->alpha[K & 1] : (VT)0;
  const ColourInfo ci = colour_info(S, c);
  const int tile = tile0 + xcd_block(blockIdx.x, gridDim.x);
  const gptr<const VT> x = C(S.w);
  for (int j = 0; j < R; ++j) {
    const int v = tile_row<R>(tile, j);
    VT yv = S.y[v];
    if (pend) yv = (VT)(yv + aK * S.p[v]);
    S.w[v] = (yv - upper_sum(S, ci, v, x)) / S.dt[v];
  }
}

template <class VT, int R>
__global__ void __launch_bounds__(256) k_ssor_scatter(const SsorSysT<VT>* __restrict__ sys) {
  const SsorView<VT> S = sview(sys);
  const bool zero = S.st->zero != 0;
  for (int j = 0; j < R; ++j) {
    const int v = tile_row<R>(blockIdx.x, j);
    const int i = S.perm[v];
    if (i >= 0) S.xout[i] = zero ? (VT)0 : S.w[v];
  }
}

template <class VT>
__global__ void k_ssor_reset(const SsorSysT<VT>* __restrict__ sys, double rtol, int maxiter, int check_every) {
  if (threadIdx.x != 0) return;
  gptr<PcgState> st = G(sys->st);
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
  st->beta = 0.0;
  st->alpha[0] = 0.0;
  st->alpha[1] = 0.0;
}

// ------------------------------------------------------------------ launch wrappers
#define GSH_SSOR_R(rpt, KER, GRID, STREAM, ...)                                          \
  switch (rpt) {                                                                       \
    case 1: hipLaunchKernelGGL((KER<VT, 1>), dim3(GRID), dim3(kTpb), 0, STREAM, __VA_ARGS__); break;   \
    case 2: hipLaunchKernelGGL((KER<VT, 2>), dim3(GRID), dim3(kTpb), 0, STREAM, __VA_ARGS__); break;   \
    case 8: hipLaunchKernelGGL((KER<VT, 8>), dim3(GRID), dim3(kTpb), 0, STREAM, __VA_ARGS__); break;   \
    default: hipLaunchKernelGGL((KER<VT, 4>), dim3(GRID), dim3(kTpb), 0, STREAM, __VA_ARGS__); break;  \
  }

template <class VT>
void launch_ssor_reset(const SsorSysT<VT>* sys, double rtol, int maxiter, int check_every, hipStream_t s) {
  hipLaunchKernelGGL(k_ssor_reset<VT>, dim3(1), dim3(64), 0, s, sys, rtol, maxiter, check_every);
}
template <class VT>
void launch_ssor_setup(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s, bool gather) {
  if (ntiles > 0) GSH_SSOR_R(rpt, k_ssor_setup, ntiles, s, sys, gather ? 1 : 0);
}
template <class VT>
void launch_ssor_residual(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s) {
  if (ntiles > 0) GSH_SSOR_R(rpt, k_ssor_residual, ntiles, s, sys);
}
template <class VT>
void launch_ssor_r0(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int rpt, hipStream_t s) {
  if (ntc > 0) GSH_SSOR_R(rpt, k_ssor_r0, ntc, s, sys, tile0, c);
}
// The single-system sweeps run in wide workgroups (256 * rpt threads, one row each) for rpt 2 and
// 4 (GSH_SSOR_WIDE = 0 builds the 256-thread form everywhere; both give the same bits). With
// events e0/e1 the runtime stamps the kernel's own start and end (hipExtLaunchKernelGGL), which
// the profiling path reads; without them a plain launch, which stream capture records into
// window graphs. At 1M nodes: 86.4 against 84.1 solves/s at rpt 2, 72.2 against 69.7 at rpt 4.
#ifndef GSH_SSOR_WIDE
#define GSH_SSOR_WIDE 1
#endif
#define GSH_SSOR_WX(R, KER, GRID, E0, E1, STREAM, ...)                                                       \
  do {                                                                                                      \
    constexpr bool W_ = GSH_SSOR_WIDE && (R == 2 || R == 4);                                                \
    const dim3 blk(W_ ? kTpb * R : kTpb);                                                                   \
    if (E0 || E1)                                                                                           \
      hipExtLaunchKernelGGL((KER<VT, R, W_>), dim3(GRID), blk, 0, STREAM, E0, E1, 0, __VA_ARGS__);          \
    else                                                                                                    \
      hipLaunchKernelGGL((KER<VT, R, W_>), dim3(GRID), blk, 0, STREAM, __VA_ARGS__);                        \
  } while (0)
#define GSH_SSOR_WRX(rpt, KER, GRID, E0, E1, STREAM, ...)                        \
  switch (rpt) {                                                               \
    case 1: GSH_SSOR_WX(1, KER, GRID, E0, E1, STREAM, __VA_ARGS__); break;     \
    case 2: GSH_SSOR_WX(2, KER, GRID, E0, E1, STREAM, __VA_ARGS__); break;     \
    case 8: GSH_SSOR_WX(8, KER, GRID, E0, E1, STREAM, __VA_ARGS__); break;     \
    default: GSH_SSOR_WX(4, KER, GRID, E0, E1, STREAM, __VA_ARGS__); break;    \
  }

template <class VT>
void launch_ssor_bwd(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int flags, int rpt, hipStream_t s,
                     hipEvent_t e0, hipEvent_t e1) {
  if (ntc > 0) GSH_SSOR_WRX(rpt, k_ssor_bwd, ntc, e0, e1, s, sys, tile0, c, flags);
}
template <class VT>
void launch_ssor_fwd(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int flags, int rpt, hipStream_t s,
                     hipEvent_t e0, hipEvent_t e1) {
  if (ntc <= 0) return;
  const int grid = ntc + ((flags & kPreWG) ? 1 : 0);
  if (flags & kChk) {
    GSH_SSOR_WRX(rpt, k_ssor_fwd_chk, grid, e0, e1, s, sys, tile0, c, flags);
  } else {
    GSH_SSOR_WRX(rpt, k_ssor_fwd, grid, e0, e1, s, sys, tile0, c, flags);
  }
}
template <class VT>
void launch_ssor_check(const SsorSysT<VT>* sys, hipStream_t s) {
  hipLaunchKernelGGL(k_ssor_check<VT>, dim3(1), dim3(kTpb), 0, s, sys);
}
template <class VT>
void launch_ssor_xsolve(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int rpt, hipStream_t s) {
  if (ntc > 0) GSH_SSOR_R(rpt, k_ssor_xsolve, ntc, s, sys, tile0, c);
}
template <class VT>
void launch_ssor_scatter(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s) {
  if (ntiles > 0) GSH_SSOR_R(rpt, k_ssor_scatter, ntiles, s, sys);
}


template <class VT>
__global__ void k_ssor_gather_states(const SsorSysT<VT>* __restrict__ sys, int count, PcgState* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = *sys[i].st;
}

template <class VT>
void launch_ssor_bwd_batch(const SsorSysT<VT>* sys, const SsorRef* refs, int nrefs, int or_flags, int rpt,
                           hipStream_t s) {
  if (nrefs > 0) GSH_SSOR_R(rpt, k_ssor_bwd_batch, nrefs, s, sys, refs, or_flags);
}
template <class VT>
void launch_ssor_fwd_batch(const SsorSysT<VT>* sys, const SsorRef* refs, int nrefs, int or_flags, int rpt,
                           hipStream_t s) {
  if (nrefs <= 0) return;
  if (or_flags & kChk) {
    GSH_SSOR_R(rpt, k_ssor_fwd_batch_chk, nrefs, s, sys, refs, or_flags);
  } else {
    GSH_SSOR_R(rpt, k_ssor_fwd_batch, nrefs, s, sys, refs, or_flags);
  }
}
template <class VT>
void launch_ssor_check_batch(const SsorSysT<VT>* sys, int count, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_ssor_check<VT>, dim3(count), dim3(kTpb), 0, s, sys);
}
template <class VT>
void launch_ssor_gather_states(const SsorSysT<VT>* sys, int count, PcgState* out, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_ssor_gather_states<VT>, dim3((count + 63) / 64), dim3(64), 0, s, sys, count, out);
}

#define GSH_SSOR_INSTANTIATE(VT)                                                                       \
  template void launch_ssor_reset<VT>(const SsorSysT<VT>*, double, int, int, hipStream_t);             \
  template void launch_ssor_setup<VT>(const SsorSysT<VT>*, int, int, hipStream_t, bool);               \
  template void launch_ssor_residual<VT>(const SsorSysT<VT>*, int, int, hipStream_t);                  \
  template void launch_ssor_r0<VT>(const SsorSysT<VT>*, int, int, int, int, hipStream_t);              \
  template void launch_ssor_bwd<VT>(const SsorSysT<VT>*, int, int, int, int, int, hipStream_t, hipEvent_t, \
                                    hipEvent_t);                                                       \
  template void launch_ssor_fwd<VT>(const SsorSysT<VT>*, int, int, int, int, int, hipStream_t, hipEvent_t, \
                                    hipEvent_t);                                                       \
  template void launch_ssor_check<VT>(const SsorSysT<VT>*, hipStream_t);                               \
  template void launch_ssor_xsolve<VT>(const SsorSysT<VT>*, int, int, int, int, hipStream_t);          \
  template void launch_ssor_scatter<VT>(const SsorSysT<VT>*, int, int, hipStream_t);                   \
  template void launch_ssor_bwd_batch<VT>(const SsorSysT<VT>*, const SsorRef*, int, int, int, hipStream_t); \
  template void launch_ssor_fwd_batch<VT>(const SsorSysT<VT>*, const SsorRef*, int, int, int, hipStream_t); \
  template void launch_ssor_check_batch<VT>(const SsorSysT<VT>*, int, hipStream_t);                          \
  template void launch_ssor_gather_states<VT>(const SsorSysT<VT>*, int, PcgState*, hipStream_t);
GSH_SSOR_INSTANTIATE(double)
GSH_SSOR_INSTANTIATE(float)

}  // namespace gsh
