// mesh_kernels.hip -- gfx950 kernels of the per-step mesh work: vertical levels and wet/dry,
// ELM backtracking, RHS producers, wave-equation assembly with essential BCs, elevation update.
//
// Arithmetic follows the reference's operation order term by term (no contraction: the library
// is built with -ffp-contract=off), so the CPU oracle reproduces every result bit for bit.
#include <hip/hip_runtime.h>

#include "device.h"
#include "reduce.h"

namespace gsh {

#define GSH_BT_WALK_CAP 1
#define GSH_BT_ITER38 2
#define GSH_BT_STUCK 4
#define GSH_BT_OPEN_EXIT 8
#define GSH_BT_SLIDE 16

constexpr double kSmall = 1.0e-7;  // meshdata.h:10

// Node::compZCoord (meshdata.cu:1744-1807) at level k of a wet node with elevation eta, depth h
// and bottom level kb; sigmaTransitionDepth (:1647) is min(h, h_s). ivcor=2 (SZ: Z levels, the
// level at -min(h, h_s), S levels) and ivcor=1 (one sigma per level, or per node and level).
__device__ __forceinline__ double z_wet(const DMesh& m, int node, double eta, double h, int kb, int k) {
  const int nv = m.nvrt;
  if (m.vmode == 2) {
    const double hmod2 = fmin(h, m.h_s);
    if (k == nv - 1) return eta;
    if (k == m.izlev) return -hmod2;
    if (k == kb) return -h;
    if (k < m.izlev) return m.ztot[k];
    const int ks = k - m.izlev;
    const double sg = m.ssig[ks];
    if (hmod2 <= m.h_c) return sg * (hmod2 + eta) + eta;
    return eta * (1.0 + sg) + m.h_c * sg + (hmod2 - m.h_c) * m.cs[ks];
  }
  if (k <= kb) return -h;
  if (k == nv - 1) return eta;
  const double sg = m.vmode == 1 ? m.sig_node[(int64_t)node * nv + k] : m.sigma[k];
  return (eta + h) * sg + eta;
}

// compZCoord (meshdata.cu:1744-1807): z of level k at a node (0 when the node is dry).
__device__ __forceinline__ double zcoord(const DMesh& m, double h0, int node, int k) {
  const double eta = m.eta2[node], h = m.depth[node];
  if (eta + h <= h0) return 0.0;  // isDry2
  return z_wet(m, node, eta, h, m.kb[node], k);
}

// ---------------------------------------------------------------- a3: levels and wet/dry
// ul0ComputeDryElements (meshdata.cu:5550-5580)
__global__ void k_elem_levels(DMesh m, double h0) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int4 r = m.elem[2 * e];
  const int nd[3] = {r.x, r.y, r.z};
  bool dry = false;
  for (int j = 0; j < 3; ++j) dry = dry || (m.eta2[nd[j]] + m.depth[nd[j]] <= h0);
  m.elem_wet[e] = dry ? 0 : 1;
  int kb = 0;
  if (!dry) kb = min(min(m.kb[nd[0]], m.kb[nd[1]]), m.kb[nd[2]]);
  m.elem_kb[e] = kb;
}

// ul0ComputeDryNodes (meshdata.cu:5596-5612)
__global__ void k_node_wet(DMesh m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  int8_t wet = 0;
  for (int j = m.ball_ptr[i]; j < m.ball_ptr[i + 1]; ++j)
    if (m.elem_wet[m.ball_elem[j]]) {
      wet = 1;
      break;
    }
  m.node_wet[i] = wet;
  m.eta_lev[i] = m.eta2[i];
}

// ul0UpdateSides / ul0UpdateSides2 (meshdata.cu:5679-5760): side kb, zc, wet flag. One thread per
// side (g sides per workgroup) computes its levels into LDS (the node values are read once per side); the workgroup then
// stores its sides' zc rows, one contiguous span of HBM, with flat coalesced stores. Dry sides
// keep their zc, as in the reference.
constexpr int kLevelThreads = 128;

__global__ void __launch_bounds__(kLevelThreads) k_side_levels(DMesh m, double h0, int g) {
  extern __shared__ double lz[];  // [g][nv | 1]
  __shared__ int8_t lwet[kLevelThreads];
  const int nv = m.nvrt, zs = nv | 1;
  const int s0 = blockIdx.x * g;
  const int gc = min(g, m.ns - s0);
  const int t = threadIdx.x;
  if (t < gc) {
    const int s = s0 + t;
    const int4 r = m.side[s];
    const bool wet = m.elem_wet[r.z] || (r.w >= 0 && m.elem_wet[r.w]);
    const int kb1 = m.kb[r.x], kb2 = m.kb[r.y];
    const int kb = wet ? min(kb1, kb2) : 0;
    if (wet) {
      double* z = lz + t * zs;
      for (int k = 0; k < kb; ++k) z[k] = 0.0;
      for (int k = kb; k < nv; ++k)
        z[k] = 0.5 * (zcoord(m, h0, r.x, max(k, kb1)) + zcoord(m, h0, r.y, max(k, kb2)));
    }
    lwet[t] = wet ? 1 : 0;
    m.side_kb[s] = kb;
    m.side_wet[s] = wet ? 1 : 0;
  }
  __syncthreads();
  const int total = gc * nv;
  const float inv = 1.0f / (float)nv;
  double* zc = m.zc + (int64_t)s0 * nv;
  for (int i = t; i < total; i += kLevelThreads) {
    // side j = i / nv from a float reciprocal (exact for i < 2^17), corrected either way
    int j = (int)(((float)i + 0.5f) * inv);
    int k = i - j * nv;
    if (k < 0) {
      --j;
      k += nv;
    }
    if (k >= nv) {
      ++j;
      k -= nv;
    }
    if (lwet[j]) zc[i] = lz[j * zs + k];
  }
}

void launch_levels(const DMesh& m, const StepArgs& a, hipStream_t st) {
  if (m.ne > 0) hipLaunchKernelGGL(k_elem_levels, dim3((m.ne + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, a.h0);
  if (m.n > 0) hipLaunchKernelGGL(k_node_wet, dim3((m.n + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m);
  if (m.ns > 0) {
    // sides per workgroup: one per thread, at most 32 KB of LDS
    const int g = max(1, min(kLevelThreads, 32 * 1024 / (8 * (m.nvrt | 1))));
    hipLaunchKernelGGL(k_side_levels, dim3((m.ns + g - 1) / g), dim3(kLevelThreads), (size_t)8 * g * (m.nvrt | 1), st,
                       m, a.h0, g);
  }
}

// ---------------------------------------------------------------- a4-a7: ELM backtracking
__device__ __forceinline__ double signed_area(double p1x, double p1y, double p2x, double p2y, double p3x,
                                              double p3y) {
  return 0.5 * ((p1x - p3x) * (p2y - p3y) - (p2x - p3x) * (p1y - p3y));
}

// Node ids and coordinates of one element, as the walk leaves them for the interpolation.
struct ElemGeom {
  int n0, n1, n2;
  int s0, s1, s2;  // sides opposite the corners
  double x0, y0, x1, y1, x2, y2;
};

// Ids are non-negative int32 wherever they index (every -1 is tested first), so the hot gathers
// index with unsigned 32-bit arithmetic (2 id + 1 < 2^32): one address instruction per load.
__device__ __forceinline__ unsigned ux(int i) { return (unsigned)i; }

__device__ __forceinline__ ElemGeom elem_geom(const DMesh& m, int e) {
  const int4 r = m.elem[2u * ux(e)], q = m.elem[2u * ux(e) + 1u];
  const double2 p0 = m.nrec[2u * ux(r.x)], p1 = m.nrec[2u * ux(r.y)], p2 = m.nrec[2u * ux(r.z)];
  return {r.x, r.y, r.z, r.w, q.x, q.y, p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
}

// A node's elevation, depth, total depth H = eta + h, bottom level, dry flag and id, in registers.
struct NodeZ {
  double eta, h, H;
  int kb;
  bool dry;
  int node;
};

__device__ __forceinline__ NodeZ node_z(const DMesh& m, double h0, int node) {
  const double2 q = m.nrec[2u * ux(node) + 1u];
  const double H = q.x + q.y;
  return {q.x, q.y, H, m.kb[ux(node)], H <= h0, node};
}

// LDS-staged element patch (k_backtrack with ST = true): every element the walk and the
// interpolation read, for the elements on both sides of each of the workgroup's sides and their
// neighbours, in one record, in an LDS table keyed by element id (open addressing, linear
// probing). An element outside the patch is read from global memory as before, so the values
// and results are those of the unstaged kernel.
struct StElem {
  double x[3], y[3];    // corners in node order n0, n1, n2
  double eta[3], h[3];  // eta2 and depth of the corners
  double area;
  int n[3], s[3];       // node ids; sides opposite the corners
  int ekb, kb[3];       // element bottom level; the corners' kb
};
constexpr int kStSlots = 128;  // table size (a power of two)
constexpr int kStCand = 96;    // at most this many candidates, so a free slot always ends a probe

__device__ __forceinline__ unsigned st_hash(int e) { return ((unsigned)e * 2654435761u) >> 25; }

// slot of element e, or -1 when it is not staged
__device__ __forceinline__ int st_find(const int* key, int e) {
  unsigned h = st_hash(e);
  for (int p = 0; p < kStSlots; ++p) {
    const int k = key[h];
    if (k == e) return (int)h;
    if (k < 0) return -1;
    h = (h + 1) & (kStSlots - 1);
  }
  return -1;
}

__device__ __forceinline__ ElemGeom st_geom(const StElem& r) {
  return {r.n[0], r.n[1], r.n[2], r.s[0], r.s[1], r.s[2], r.x[0], r.y[0], r.x[1], r.y[1], r.x[2], r.y[2]};
}

__device__ __forceinline__ NodeZ st_node_z(const StElem& r, int c, double h0) {
  const double H = r.eta[c] + r.h[c];
  return {r.eta[c], r.h[c], H, r.kb[c], H <= h0, r.n[c]};
}

// Fills the table for sides s_first .. s_first + nsides - 1: candidate j of a side is one of
// its two elements (j = 0, 4) or that element's neighbour across its side j & 3 - 1.
__device__ void st_build(const DMesh& m, int* key, StElem* rec, int s_first, int nsides) {
  for (int i = threadIdx.x; i < kStSlots; i += blockDim.x) key[i] = -1;
  __syncthreads();
  const int i = threadIdx.x;
  if (i < min(nsides * 8, kStCand)) {
    const int4 sr = m.side[s_first + i / 8];
    const int j = i % 8, q = j & 3;
    int e = (j & 4) ? sr.w : sr.z;
    if (e >= 0 && q > 0) {
      const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
      const int es = q == 1 ? r0.w : (q == 2 ? r1.x : r1.y);
      const int4 o = m.side[es];
      e = (o.z == e) ? o.w : o.z;
    }
    if (e >= 0) {
      unsigned h = st_hash(e);
      for (int p = 0; p < kStSlots; ++p) {
        const int old = atomicCAS(&key[h], -1, e);
        if (old == -1 || old == e) break;
        h = (h + 1) & (kStSlots - 1);
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kStSlots; t += blockDim.x) {
    const int e = key[t];
    if (e < 0) continue;
    const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
    StElem& R = rec[t];
    const int nn[3] = {r0.x, r0.y, r0.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 p = m.nrec[2 * nn[c]], z = m.nrec[2 * nn[c] + 1];
      R.x[c] = p.x;
      R.y[c] = p.y;
      R.eta[c] = z.x;
      R.h[c] = z.y;
      R.kb[c] = m.kb[nn[c]];
      R.n[c] = nn[c];
    }
    R.s[0] = r0.w;
    R.s[1] = r1.x;
    R.s[2] = r1.y;
    R.area = m.area[e];
    R.ekb = m.elem_kb[e];
  }
  __syncthreads();
}

// zcoord(node, max(M, kb)) with sM = sigma[M] (one sigma per level): for M <= kb that is the bed
// -h, above kb level M.
__device__ __forceinline__ double zlev(const NodeZ& q, double sM, int M, int nv) {
  const double zs = (M == nv - 1) ? q.eta : q.H * sM + q.eta;
  return q.dry ? 0.0 : ((M <= q.kb) ? -q.h : zs);
}

// z of level M at the point with shape coordinates s0..s2 (the vertical bracket search's probe).
// VM = 0: one sigma per level, read once per probe from sig (LDS); otherwise each node's level
// max(M, kb) from z_wet (per-node sigma or SZ), whose level kb is the bed -h in every system.
template <int VM>
__device__ __forceinline__ double zprobe(const DMesh& m, const NodeZ& q0, const NodeZ& q1, const NodeZ& q2,
                                         double s0, double s1, double s2, const double* sig, int M, int nv) {
  if (VM == 0) {
    const double sM = sig[max(M, 0)];
    return s0 * zlev(q0, sM, M, nv) + s1 * zlev(q1, sM, M, nv) + s2 * zlev(q2, sM, M, nv) + 0.0;
  }
  auto zl = [&](const NodeZ& q) {
    return q.dry ? 0.0 : ((M <= q.kb) ? -q.h : z_wet(m, q.node, q.eta, q.h, q.kb, M));
  };
  return s0 * zl(q0) + s1 * zl(q1) + s2 * zl(q2) + 0.0;
}

// Element::interpolateVelocities for triangles, ics=1 (meshdata.cu:1209-1291) with
// computeShapeCoords(fiAnywhere) (meshdata.cu:990-1043); g is element e's geometry.
// hint: the lower level of the previous vertical bracket (updated here).
// rc: element e's staged record (ST), else NULL.
template <int VM, bool ST>
__device__ void interp_velocity(const DMesh& m, const double* sig, double h0, int e, const ElemGeom& g,
                                const StElem* rc, double px, double py, double pz, double& uo, double& vo,
                                double& wo, int& hint) {
  const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
  const double x0 = g.x0, y0 = g.y0, x1 = g.x1, y1 = g.y1, x2 = g.x2, y2 = g.y2;
  const double area = (ST && rc) ? rc->area : m.area[ux(e)];
  double s0 = signed_area(px, py, x1, y1, x2, y2) / area;
  double s1 = signed_area(x0, y0, px, py, x2, y2) / area;
  double s2 = 1.0 - s0 - s1;
  const double tmin = fmin(s0, fmin(s1, s2));
  if (tmin < 0.0) {
    const double aa = -tmin + 1.0e-8;
    const double bb = 1.0 / (s0 + s1 + s2 + 3.0 * aa);
    s0 = (s0 + aa) * bb;
    s1 = (s1 + aa) * bb;
    s2 = (s2 + aa) * bb;
  }
  const int nv = m.nvrt;
  const bool st = ST && rc;
  const NodeZ q0 = st ? st_node_z(*rc, 0, h0) : node_z(m, h0, n0);
  const NodeZ q1 = st ? st_node_z(*rc, 1, h0) : node_z(m, h0, n1);
  const NodeZ q2 = st ? st_node_z(*rc, 2, h0) : node_z(m, h0, n2);
  const int kb0 = q0.kb, kb1 = q1.kb, kb2 = q2.kb;
  const int L0 = (st ? rc->ekb : m.elem_kb[ux(e)]) - 1, R0 = nv - 1;
  int L = L0, R = R0;
  double zL = 0.0, zR = 0.0, w1, w2;
  bool inside = false;
  if (VM == 0) {
    // Start from the bracket of the previous point, checked with the exact level z values. The
    // interpolated z is non-decreasing in the level (non-negative shape coordinates, each node's
    // levels non-decreasing), so a bracket z_h < pz < z_(h+1) is the unique one the bisection
    // below would end in, with the same zL and zR; the column-end cases (pz <= z_L0,
    // pz >= z_R0) are excluded by it. Otherwise, one step down or up, then the full search.
    const int h = min(max(hint, L0), R0 - 1);
    const double zh = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, h, nv);
    const double zh1 = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, h + 1, nv);
    if (zh < pz && pz < zh1) {
      L = h, R = h + 1, zL = zh, zR = zh1, inside = true;
    } else if (pz < zh && h > L0) {
      const double zd = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, h - 1, nv);
      if (zd < pz) L = h - 1, R = h, zL = zd, zR = zh, inside = true;
    } else if (pz > zh1 && h + 2 <= R0) {
      const double zu = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, h + 2, nv);
      if (pz < zu) L = h + 1, R = h + 2, zL = zh1, zR = zu, inside = true;
    }
  }
  if (inside) {
    w2 = (pz - zL) / (zR - zL);
    w1 = 1.0 - w2;
  } else {
    zL = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, L, nv);
    zR = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, R, nv);
    if (pz <= zL) {
      w1 = 1.0;
      w2 = 0.0;
    } else if (pz >= zR) {
      w1 = 0.0;
      w2 = 1.0;
    } else {
      while (R - L > 1) {
        const int M = (L + R) / 2;
        const double zM = zprobe<VM>(m, q0, q1, q2, s0, s1, s2, sig, M, nv);
        if (zM < pz) {
          L = M;
          zL = zM;
        } else {
          R = M;
          zR = zM;
        }
      }
      w2 = (pz - zL) / (zR - zL);
      w1 = 1.0 - w2;
      hint = L;
    }
  }
  if (inside) hint = L;
  // the three nodes' level rows (64-bit row offsets), then two levels of each
  const double4* U = reinterpret_cast<const double4*>(m.uvw);
  const double4* u0 = U + (uint64_t)ux(n0) * ux(nv);
  const double4* u1 = U + (uint64_t)ux(n1) * ux(nv);
  const double4* u2 = U + (uint64_t)ux(n2) * ux(nv);
  const double4 a0 = u0[ux(max(L, kb0))], a1 = u1[ux(max(L, kb1))], a2 = u2[ux(max(L, kb2))];
  const double4 b0 = u0[ux(max(R, kb0))], b1 = u1[ux(max(R, kb1))], b2 = u2[ux(max(R, kb2))];
  const double uL = s0 * a0.x + s1 * a1.x + s2 * a2.x + 0.0, vL = s0 * a0.y + s1 * a1.y + s2 * a2.y + 0.0;
  const double uR = s0 * b0.x + s1 * b1.x + s2 * b2.x + 0.0, vR = s0 * b0.y + s1 * b1.y + s2 * b2.y + 0.0;
  const double wL = s0 * a0.z + s1 * a1.z + s2 * a2.z + 0.0, wR = s0 * b0.z + s1 * b1.z + s2 * b2.z + 0.0;
  uo = w1 * uL + w2 * uR;
  vo = w1 * vL + w2 * vR;
  wo = w1 * wL + w2 * wR;
}

// Leaded celestial walk (meshdata.cu:85-127) with Element::computePredicates (meshdata.cu:651-746).
// Corner j of a triangle is getNodeInOrder(j) = nodes[(j+1)%3]; walk side j is element side j.
// geo_valid: geo already holds element cur (the previous walk's final element), so the first
// step reads nothing. On return, geo holds the geometry of the final element cur.
// ST: slot is element cur's staged slot (-1: not staged), kept with geo.
template <bool ST>
__device__ __forceinline__ void walk_geom(const DMesh& m, const int* key, const StElem* rec, int cur, ElemGeom& geo,
                                          int& slot) {
  if (ST) {
    slot = st_find(key, cur);
    geo = slot >= 0 ? st_geom(rec[slot]) : elem_geom(m, cur);
  } else {
    geo = elem_geom(m, cur);
  }
}

template <bool ST>
__device__ bool celestial_walk(const DMesh& m, const int* key, const StElem* rec, int& cur, double x0x, double x0y,
                               double xtx, double xty, int& last_cross, int& status, ElemGeom& geo, int& slot,
                               bool geo_valid) {
  double fac = 0.00001;
  for (int iter = 0; iter < 40; ++iter) {
    if (iter > 0 || !geo_valid) walk_geom<ST>(m, key, rec, cur, geo, slot);
    const double C0x = geo.x1, C0y = geo.y1;  // corner j is node (j + 1) % 3
    const double C1x = geo.x2, C1y = geo.y2;
    const double C2x = geo.x0, C2y = geo.y0;
    const double vx = xtx - x0x, vy = xty - x0y;
    const double d0 = vx * (C0x - x0x) + vy * (C0y - x0y);
    const double d1 = vx * (C1x - x0x) + vy * (C1y - x0y);
    const double d2 = vx * (C2x - x0x) + vy * (C2y - x0y);
    double maxp = fmax(d0, d1);
    maxp = fmax(maxp, d2);
    // fmin(1, (maxp / vv) 1.00001) is 1 whenever maxp > 2 vv (the rounded quotient is then at
    // least 2; vv = 0 gives +inf as well), the usual case of a step shorter than the element:
    // the division runs only otherwise (NaN and overflow included), with the same result
    const double vv = vx * vx + vy * vy;
    double lead = 1.0;
    if (!(maxp > 2.0 * vv)) lead = fmin(1.0, (maxp / vv) * 1.00001);
    fac = fmax(1.00001 * fac, lead);
    const double ix = x0x + (xtx - x0x) * fac, iy = x0y + (xty - x0y) * fac;
    const double s0x = C1x - C0x, s0y = C1y - C0y;
    const double s1x = C2x - C1x, s1y = C2y - C1y;
    const double s2x = C0x - C2x, s2y = C0y - C2y;
    const bool left0 = (s0x * (iy - C0y) - s0y * (ix - C0x)) > -kSmall;
    const bool left1 = (s1x * (iy - C1y) - s1y * (ix - C1x)) > -kSmall;
    const bool left2 = (s2x * (iy - C2y) - s2y * (ix - C2x)) > -kSmall;
    // a side is an exit candidate only if the point is not left of it: with all three left
    // tests passed the walk ends here (the usual case), without the obtuse and bisector tests
    if (left0 && left1 && left2) return true;
    const bool obt0 = (s0x * s1x + s0y * s1y) > 0.0;
    const bool obt1 = (s1x * s2x + s1y * s2y) > 0.0;
    const bool obt2 = (s2x * s0x + s2y * s0y) > 0.0;
    const bool bis0 = ((ix - C1x) * (s0x + s1x) + (iy - C1y) * (s0y + s1y)) < 0.0;
    const bool bis1 = ((ix - C2x) * (s1x + s2x) + (iy - C2y) * (s1y + s2y)) < 0.0;
    const bool bis2 = ((ix - C0x) * (s2x + s0x) + (iy - C0y) * (s2y + s0y)) < 0.0;
    const bool v0 = !left0 && (!obt0 || bis0);
    const bool v1 = !left1 && (!obt1 || bis1);
    const bool v2 = !left2 && (!obt2 || bis2);
    if (!(v0 || v1 || v2)) return true;
    // the three side ids as opaque register values: a select among the struct's fields would
    // otherwise become an indexed load from a stack copy of it
    int gs0 = geo.s0, gs1 = geo.s1, gs2 = geo.s2;
    asm volatile("" : "+v"(gs0), "+v"(gs1), "+v"(gs2));
    const int side = v0 ? gs0 : (v1 ? gs1 : gs2);
    last_cross = side;
    const int4 sr = m.side[ux(side)];
    const int nxt = (sr.z == cur) ? sr.w : sr.z;
    if (nxt < 0 || !m.elem_wet[ux(nxt)]) return false;
    cur = nxt;
  }
  status |= GSH_BT_WALK_CAP;
  walk_geom<ST>(m, key, rec, cur, geo, slot);
  return true;
}

// Side::btrack (meshdata.cu:1449-1629). Returns the final velocity and the foot element.
template <int VM, bool ST>
__device__ void btrack(const DMesh& m, const double* sig, const int* key, const StElem* rec, const StepArgs& a, int s,
                       int k, double dtbk, int cur, double& ux, double& uy, int& foot, int& status) {
  const double h0 = a.h0;
  double x0x = m.side_cx[s], x0y = m.side_cy[s], x0z = m.zc[(int64_t)s * m.nvrt + k];
  double time_rm = a.dt, t1 = 1e10, t2 = 2e10, t3 = 3e10;
  const int4 sr = m.side[s];
  const double4* U = reinterpret_cast<const double4*>(m.uvw);
  double uix = m.su2[2 * ((int64_t)s * m.nvrt + k)], uiy = m.su2[2 * ((int64_t)s * m.nvrt + k) + 1];
  double uiz = 0.5 * (U[(int64_t)sr.x * m.nvrt + k].z + U[(int64_t)sr.y * m.nvrt + k].z);
  double xtx = -1.0e25, xty = -1.0e25, xtz = -1.0e25;
  int iter = 0;
  int hint = k;  // lower level of the last vertical bracket; the start point lies at level k
  ElemGeom geo;  // set by every walk before its first use
  int slot = -1;  // geo's staged slot (ST)
  bool geo_valid = false;
  while (true) {
    const double dtb = fmin(dtbk, time_rm);
    xtx = x0x - dtb * uix;
    xty = x0y - dtb * uiy;
    xtz = x0z - dtb * uiz;
    int last = -1;
    const bool found = celestial_walk<ST>(m, key, rec, cur, x0x, x0y, xtx, xty, last, status, geo, slot, geo_valid);
    geo_valid = true;
    const StElem* rc = (ST && slot >= 0) ? rec + slot : nullptr;
    if (t3 <= time_rm) {
      xtx = x0x;
      xty = x0y;
      xtz = x0z;
      status |= GSH_BT_STUCK;
      break;
    }
    if (!found) {
      const int4 cs = m.side[last];
      const double2 P1 = m.nrec[2 * cs.x], P2 = m.nrec[2 * cs.y];
      const double p1x = P1.x, p1y = P1.y, p2x = P2.x, p2y = P2.y;
      const double v1x = xtx - x0x, v1y = xty - x0y;
      const double v2x = p2x - p1x, v2y = p2y - p1y;
      const double denom = v2x * v1y - v2y * v1x;
      double t, sp;
      if (fabs(denom) < kSmall) {
        t = fmin(fmax(((xtx - p1x) * v2x + (xty - p1y) * v2y) / (v2x * v2x + v2y * v2y), 0.0), 1.0);
        if (t < 0.0000000001) t = 0.0;
        if (t > 0.9999999999) t = 1.0;
        sp = fmin(fmax(((p1x + t * v2x - x0x) * v1x + (p1y + t * v2y - x0y) * v1y) / (v1x * v1x + v1y * v1y), 0.0),
                  1.0);
      } else {
        t = fmin(fmax(((x0x - p1x) * v1y - (x0y - p1y) * v1x) / denom, 0.0), 1.0);
        if (t < 0.0000000001) t = 0.0;
        if (t > 0.9999999999) t = 1.0;
        sp = fmin(fmax(((p1x + t * v2x - x0x) * v1x + (p1y + t * v2y - x0y) * v1y) / (v1x * v1x + v1y * v1y), 0.0),
                  1.0);
        if (sp < 0.0000000001) sp = 0.0;
      }
      time_rm -= sp * dtb;
      const double nz = x0z + sp * (xtz - x0z);
      xtx = p1x + t * v2x;
      xty = p1y + t * v2y;
      xtz = nz;
      if (m.side_type[last] == 2) {
        status |= GSH_BT_OPEN_EXIT;
        break;
      }
      const double len = sqrt(v2x * v2x + v2y * v2y);
      const double tx = v2x / len, ty = v2y / len;
      double iu, iv, iw;
      interp_velocity<VM, ST>(m, sig, h0, cur, geo, rc, xtx, xty, xtz, iu, iv, iw, hint);
      const double d = iu * tx + iv * ty;
      uix = tx * d;
      uiy = ty * d;
      uiz = iw;
      status |= GSH_BT_SLIDE;
    } else {
      time_rm -= dtb;
      interp_velocity<VM, ST>(m, sig, h0, cur, geo, rc, xtx, xty, xtz, uix, uiy, uiz, hint);
    }
    if (time_rm <= 0.0) break;
    if (uix * uix + uiy * uiy <= a.velmin2) {
      ux = uix;
      uy = uiy;
      foot = cur;
      return;
    }
    x0x = xtx;
    x0y = xty;
    x0z = xtz;
    ++iter;
    if (iter == 38) {
      status |= GSH_BT_ITER38;
      ux = uix;
      uy = uiy;
      foot = cur;
      return;
    }
    t3 = t2;
    t2 = t1;
    t1 = time_rm;
  }
  double iw;
  interp_velocity<VM, ST>(m, sig, h0, cur, geo, (ST && slot >= 0) ? rec + slot : nullptr, xtx, xty, xtz, ux, uy, iw,
                          hint);
  foot = cur;
}

// backtrackAllNodes (meshdata.cu:3547-3660), one thread per (side, level): consecutive
// lanes are consecutive levels of one side, so the nodal [node][level] gathers of a wave
// fall on a few contiguous cache lines.
// VM: the vertical coordinate system (DMesh::vmode); 0, one sigma per level, keeps it in LDS.
// ST: the workgroup first stages its element patch in LDS (all sides and levels, no list).
template <int VM, bool ST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_backtrack(DMesh m, StepArgs a, const int32_t* list_side,
                                                   const int32_t* list_lev, int32_t count) {
  __shared__ double sig[VM == 0 ? 1024 : 1];
  __shared__ int st_key[ST ? kStSlots : 1];
  __shared__ StElem st_rec[ST ? kStSlots : 1];
  if (VM == 0) {
    for (int j = threadIdx.x; j < m.nvrt; j += blockDim.x) sig[j] = m.sigma[j];
    __syncthreads();
  }
  if (ST) {  // the workgroup's sides: from its first to its last (side, level) pair
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t tl = min(t0 + (int64_t)blockDim.x, (int64_t)m.ns * m.nvrt) - 1;
    const int s_first = (int)(t0 / m.nvrt), s_last = (int)(tl / m.nvrt);
    st_build(m, st_key, st_rec, s_first, s_last - s_first + 1);
  }
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int s, k;
  if (list_side) {
    if (t >= count) return;
    s = list_side[t];
    k = list_lev[t];
  } else {
    if (t >= (int64_t)m.ns * m.nvrt) return;
    s = (t < 0xffffffffll) ? (int)((uint32_t)t / (uint32_t)m.nvrt) : (int)(t / m.nvrt);
    k = (int)(t - (int64_t)s * m.nvrt);
  }
  const int64_t idx = (int64_t)s * m.nvrt + k;
  const double ux0 = m.su2[2 * idx], uy0 = m.su2[2 * idx + 1];
  double bx = ux0, by = uy0;
  int foot = -1, status = 0;
  if (m.side_wet[s] && k >= m.side_kb[s] && !(ux0 * ux0 + uy0 * uy0 <= a.velmin2)) {
    const int4 sr = m.side[s];
    const double4* U = reinterpret_cast<const double4*>(m.uvw);
    double suma = 0.0;
    int icount = 0;
    for (int l = 0; l < 2; ++l) {
      const int e = (l == 0) ? sr.z : sr.w;
      if (e < 0) continue;
      ++icount;
      const int4 r = m.elem[2 * e];
      const double4 q0 = U[(int64_t)r.x * m.nvrt + k], q1 = U[(int64_t)r.y * m.nvrt + k],
                    q2 = U[(int64_t)r.z * m.nvrt + k];
      const double* g = m.gradphi + 6 * (int64_t)e;
      const double gux = q0.x * g[0] + q1.x * g[2] + q2.x * g[4];
      const double guy = q0.x * g[1] + q1.x * g[3] + q2.x * g[5];
      const double gvx = q0.y * g[0] + q1.y * g[2] + q2.y * g[4];
      const double gvy = q0.y * g[1] + q1.y * g[3] + q2.y * g[5];
      suma += a.dt * sqrt((gux * gux + guy * guy) + (gvx * gvx + gvy * gvy));
    }
    const int nd = (int)(suma / icount) * 4;
    const int ndelt = max(a.ndelt_min, min(nd, a.ndelt_max));
    const double dtbk = a.dt / ndelt;
    const int start = m.elem_wet[sr.z] ? sr.z : sr.w;
    btrack<VM, ST>(m, sig, st_key, st_rec, a, s, k, dtbk, start, bx, by, foot, status);
  }
  m.btu[2 * idx] = bx;
  m.btu[2 * idx + 1] = by;
  m.foot[idx] = foot;
  m.bstatus[idx] = status;
}

// Packs (x, y) and (eta2, depth) of every node next to each other for the backtracking gathers.
__global__ void k_pack_nodes(DMesh m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  m.nrec[2 * i] = make_double2(m.x[i], m.y[i]);
  m.nrec[2 * i + 1] = make_double2(m.eta2[i], m.depth[i]);
}

void launch_backtrack(const DMesh& m, const StepArgs& a, const int32_t* ls, const int32_t* ll, int32_t count,
                      hipStream_t st) {
  if (m.n > 0) hipLaunchKernelGGL(k_pack_nodes, dim3((m.n + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m);
  const int64_t total = ls ? (int64_t)count : (int64_t)m.ns * m.nvrt;
  if (total <= 0) return;
  const dim3 grid((unsigned)((total + 255) / 256));
  const bool stage = m.bt_stage != 0 && !ls;
  if (m.vmode == 0 && stage)
    hipLaunchKernelGGL((k_backtrack<0, true>), grid, dim3(256), 0, st, m, a, ls, ll, count);
  else if (m.vmode == 0)
    hipLaunchKernelGGL((k_backtrack<0, false>), grid, dim3(256), 0, st, m, a, ls, ll, count);
  else if (stage)
    hipLaunchKernelGGL((k_backtrack<1, true>), grid, dim3(256), 0, st, m, a, ls, ll, count);
  else
    hipLaunchKernelGGL((k_backtrack<1, false>), grid, dim3(256), 0, st, m, a, ls, ll, count);
}

// ---------------------------------------------------------------- nodal velocities
// zeroNodalVelocities, computeNodalVerticalVelocities1/2 and computeNodalHorizontalVelocities1/2
// (meshdata.cu:5765-5929) for triangles, ics=1, indvel <= 0. The reference scatters each
// element's contributions to its nodes with atomicAdd; here one thread per (node, level) gathers
// them over the node's ball in ball order, so the sums are deterministic:
//   w2[k] = (1 / sum area_e) * sum w_e[k] area_e                 over all ball elements, k >= kb
//   u2[k] = (sum clamp(u_s2 + u_s3 - u_s1)[k]) * (1 / count)     over the wet ball elements, k >= kb
// (s1 the element side opposite the node, s2 and s3 the following ones; count = wet elements;
// zero below kb and where no wet element touches the node). The level index is the fastest, so a
// wave reads consecutive levels of the same side and element rows.
__device__ __forceinline__ double clamp_vel(double x, double hi) { return fmin(fmax(x, -hi), hi); }

// The workgroup's threads cover the levels of a few consecutive nodes, whose balls are one
// contiguous range of the ball arrays: its entries (element, area, wet flag and the three side
// ids) are staged in LDS once, up to kNvCap of them (the rest read from global memory, the same
// values), so each (node, level) thread only gathers its level's w and u2.
constexpr int kNvCap = 512;

__global__ void __launch_bounds__(256) k_nodal_velocity(DMesh m, double rx, double ry) {
  __shared__ double Larea[kNvCap];
  __shared__ int Le[kNvCap], Ls1[kNvCap], Ls2[kNvCap], Ls3[kNvCap];
  __shared__ int8_t Lwet[kNvCap];
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x;
  const uint32_t nv = (uint32_t)m.nvrt;
  const uint64_t total = (uint64_t)m.n * nv;
  auto node_of = [&](uint64_t tt) { return (tt < 0xffffffffull) ? (int)((uint32_t)tt / nv) : (int)(tt / nv); };
  // sides (r0.w, r1.x, r1.y) are opposite corners 0, 1, 2; s1 is the side opposite the node
  auto sides_of = [&](int e, int l, int& s1, int& s2, int& s3) {
    const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
    s1 = (l == 0) ? r0.w : ((l == 1) ? r1.x : r1.y);
    s2 = (l == 0) ? r1.x : ((l == 1) ? r1.y : r0.w);
    s3 = (l == 0) ? r1.y : ((l == 1) ? r0.w : r1.x);
  };
  const int ifirst = node_of(t0), ilast = node_of(min(t0 + blockDim.x, total) - 1);
  const int jb = m.ball_ptr[ifirst];
  const int nst = min(m.ball_ptr[ilast + 1] - jb, kNvCap);
  for (int q = threadIdx.x; q < nst; q += blockDim.x) {
    const int e = m.ball_elem[jb + q];
    Le[q] = e;
    Larea[q] = m.area[e];
    Lwet[q] = m.elem_wet[e];
    sides_of(e, m.ball_loc[jb + q], Ls1[q], Ls2[q], Ls3[q]);
  }
  __syncthreads();
  const uint64_t t = t0 + threadIdx.x;
  if (t >= total) return;
  const int i = node_of(t);
  const int k = (int)(t - (uint64_t)i * nv);
  const int j0 = m.ball_ptr[i], j1 = m.ball_ptr[i + 1];
  const bool node_wet = m.node_wet[i] != 0;
  const bool above = k >= m.kb[i];
  const double2* su = reinterpret_cast<const double2*>(m.su2);
  // one pass over the ball; each sum keeps its own ball-order sequence of additions
  double asum = 0.0, ws = 0.0, sx = 0.0, sy = 0.0;
  int count = 0;
  for (int j = j0; j < j1; ++j) {
    const int q = j - jb;
    const bool staged = q < nst;
    const int e = staged ? Le[q] : m.ball_elem[j];
    const double a = staged ? Larea[q] : m.area[e];
    asum += a;
    if (above) ws += m.ew[(int64_t)e * nv + k] * a;
    if (!(staged ? Lwet[q] : m.elem_wet[e]) || !node_wet) continue;
    ++count;
    if (!above) continue;
    int s1, s2, s3;
    if (staged) {
      s1 = Ls1[q];
      s2 = Ls2[q];
      s3 = Ls3[q];
    } else {
      sides_of(e, m.ball_loc[j], s1, s2, s3);
    }
    const double2 u1 = su[(int64_t)s1 * nv + k];
    const double2 u2 = su[(int64_t)s2 * nv + k];
    const double2 u3 = su[(int64_t)s3 * nv + k];
    sx += clamp_vel(u2.x + u3.x - u1.x, rx);
    sy += clamp_vel(u2.y + u3.y - u1.y, ry);
  }
  double ux = 0.0, uy = 0.0, w = 0.0;
  if (above) {
    w = (1.0 / asum) * ws;
    if (count > 0) {
      ux = sx * (1.0 / count);
      uy = sy * (1.0 / count);
    }
  }
  reinterpret_cast<double4*>(m.uvw)[t] = make_double4(ux, uy, w, 0.0);
}

void launch_nodal_velocity(const DMesh& m, double rmaxvel, hipStream_t st) {
  const int64_t total = (int64_t)m.n * m.nvrt;
  if (total > 0)
    hipLaunchKernelGGL(k_nodal_velocity, dim3((unsigned)((total + kTpb - 1) / kTpb)), dim3(kTpb), 0, st, m, rmaxvel,
                       rmaxvel * 1.013);
}

// ---------------------------------------------------------------- a8-a10, elevation BC
// computeVerticallyIntegratedVelocity (meshdata.cu:4459-4476) and
// computeVegetationAndDragParameters for iveg=0, nchi=0 (meshdata.cu:4615-4725).
// A workgroup takes g consecutive sides: the sides' zc and velocity rows are one contiguous
// span of HBM each, which the workgroup copies into LDS with flat coalesced loads; one thread
// per side then sums its levels in the reference's order from LDS, and the velocity rows are
// replaced by the btu rows for the second integral. Row strides in LDS are odd so the
// per-side reads spread over the banks.
// Row g = i / row_len and column j from a float reciprocal (exact here: i < 2^20, row_len <=
// 2048); the two corrections keep it exact whatever the rounding.
__device__ __forceinline__ int split_row(int i, float inv, int row_len, int& j) {
  int g = (int)(((float)i + 0.5f) * inv);
  j = i - g * row_len;
  if (j < 0) {
    --g;
    j += row_len;
  }
  if (j >= row_len) {
    ++g;
    j -= row_len;
  }
  return g;
}

// Each thread issues its loads 8 at a time before the LDS writes, so a small workgroup still
// keeps enough loads in flight.
__device__ __forceinline__ void stage_rows(double* __restrict__ dst, int stride, const double* __restrict__ src,
                                           int row_len, int rows) {
  constexpr int B = 8;
  const int total = rows * row_len;
  const float inv = 1.0f / (float)row_len;
  for (int base = threadIdx.x; base < total; base += B * blockDim.x) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = base + u * blockDim.x;
      v[u] = i < total ? src[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = base + u * blockDim.x;
      if (i < total) {
        int j;
        const int g = split_row(i, inv, row_len, j);
        dst[g * stride + j] = v[u];
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_side_rhs(DMesh m, StepArgs a, int g) {
  extern __shared__ double lds[];
  const int nv = m.nvrt;
  const int zs = nv | 1, us = (2 * nv) | 1;
  double* Lz = lds;           // [g][zs]: zc
  double* Lu = lds + g * zs;  // [g][us]: side velocity u2, then btu
  const int s0 = blockIdx.x * g;
  const int gc = min(g, m.ns - s0);
  stage_rows(Lz, zs, m.zc + (int64_t)s0 * nv, nv, gc);
  stage_rows(Lu, us, m.su2 + 2 * (int64_t)s0 * nv, 2 * nv, gc);
  __syncthreads();
  const int t = threadIdx.x;
  const bool mine = t < gc;
  const int s = s0 + t;
  bool wet = false;
  int kb = 0;
  double Ux = 0.0, Uy = 0.0;
  const double* zc = Lz + t * zs;
  const double* u = Lu + t * us;
  if (mine) {
    wet = m.side_wet[s] != 0;
    if (wet) {
      kb = m.side_kb[s];
      for (int k = kb; k < nv - 1; ++k) {
        const double ucx = 0.5 * (u[2 * k] + u[2 * (k + 1)]);
        const double ucy = 0.5 * (u[2 * k + 1] + u[2 * (k + 1) + 1]);
        const double dz = zc[k + 1] - zc[k];
        Ux += dz * ucx;
        Uy += dz * ucy;
      }
    }
  }
  __syncthreads();
  stage_rows(Lu, us, m.btu + 2 * (int64_t)s0 * nv, 2 * nv, gc);
  __syncthreads();
  if (!mine) return;
  double hhat = 0.0, chi = 0.0;
  if (wet) {
    const int4 r = m.side[s];
    const double htot = 0.5 * (m.eta2[r.x] + m.eta2[r.y]) + 0.5 * (m.depth[r.x] + m.depth[r.y]);
    const double* bt = u;
    // the side's share of computeElemAverageGhat (meshdata.cu:6183-6424): vertical integral of btu,
    // the same for both elements of the side
    double usx = 0.0, usy = 0.0;
    if (kb == nv - 2) {
      const double dz = zc[nv - 1] - zc[nv - 2];
      usx += bt[2 * (nv - 1)] * dz;
      usy += bt[2 * (nv - 1) + 1] * dz;
    } else {
      for (int k = kb + 1; k < nv; ++k) {
        const double dz = zc[k] - zc[k - 1];
        usx += (bt[2 * k] + bt[2 * (k - 1)]) / 2.0 * dz;
        usy += (bt[2 * k + 1] + bt[2 * (k - 1) + 1]) / 2.0 * dz;
      }
    }
    m.side_us[2 * s] = usx;
    m.side_us[2 * s + 1] = usy;
    const double bx = bt[2 * (kb + 1)], by = bt[2 * (kb + 1) + 1];
    const double vmag1 = sqrt(bx * bx + by * by);
    const double cd = 0.5 * (m.cd[r.x] + m.cd[r.y]);
    const double chi2 = cd * vmag1;
    chi = chi2;
    if (kb == nv - 2) {
      const double fricH = htot + chi2 * a.dt;
      hhat = htot * htot / fricH;
    } else {
      hhat = htot - chi * a.dt;
      if (a.ihhat == 1) hhat = fmax(0.0, hhat);
    }
  }
  m.sU[2 * s] = Ux;
  m.sU[2 * s + 1] = Uy;
  m.hhat[s] = hhat;
  m.chi[s] = chi;
}

// computeElemAverageGhat (meshdata.cu:6183-6424) for the barotropic, ics=1, iveg=0 subset with
// no Coriolis, wind stress, horizontal viscosity, atmospheric pressure or tidal potential:
// those terms are identically zero here and are kept as explicit zero operands.
__global__ void k_elem_ghat(DMesh m, StepArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double gx = 0.0, gy = 0.0;
  if (m.elem_wet[e]) {
    const int nv = m.nvrt;
    const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
    const int sides[3] = {r0.w, r1.x, r1.y};
    const double dt = a.dt;
    for (int j = 0; j < 3; ++j) {
      const int s = sides[j];
      const int4 sr = m.side[s];
      const double chi = m.chi[s];
      const double htot = 0.5 * (m.eta2[sr.x] + m.eta2[sr.y]) + 0.5 * (m.depth[sr.x] + m.depth[sr.y]);
      const int kb = m.side_kb[s];
      double cff1, cff2, cff3;
      if (kb == nv - 2) {
        cff1 = m.hhat[s] / htot;
        cff2 = 0.0;
        cff3 = 0.0;
      } else {
        cff1 = 1.0;
        cff2 = 1.0;
        cff3 = 0.0;
      }
      const double* bt = m.btu + 2 * (int64_t)s * nv;
      const double usx = m.side_us[2 * s], usy = m.side_us[2 * s + 1];  // from k_side_rhs
      gx += cff1 * usx - cff3 * 0.0;
      gy += cff1 * usy - cff3 * 0.0;
      const double ubx = bt[2 * (kb + 1)], uby = bt[2 * (kb + 1) + 1];
      gx -= chi * dt * cff2 * ubx + dt * cff1 * 0.0;
      gy -= chi * dt * cff2 * uby + dt * cff1 * 0.0;
    }
    gx = (1.0 / 3) * gx;
    gy = (1.0 / 3) * gy;
  }
  m.ghat1[2 * e] = gx;
  m.ghat1[2 * e + 1] = gy;
}

// Elevation BC type of node i: its own (GSH_F_ETYPE) or, where never set, the step parameter.
__device__ __forceinline__ int node_etype(const DMesh& m, const StepArgs& a, int i) {
  const int t = m.etype[i];
  return t < 0 ? a.iettype : t;
}

// computeElevationBoundaryConditions (meshdata.cu:4276-4339) for node types 1-5: eth for 1, 2, 4,
// the tidal constituents for 3, both for 5; no ramp-down of an initial elevation
// (nramp_elev = 0), no loading tide, no inverse barometric term. Type 0 leaves elbc as it is.
__global__ void k_elbc(DMesh m, StepArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  if (!m.open_bnd[i]) return;
  const int iettype = node_etype(m, a, i);
  const double eic = 0.0;
  const double r_eic = (1.0 - a.ramp) * eic;
  const double r_eth = a.ramp * m.eth[i];
  double r_eta1_bar = 0.0;
  if (iettype == 3 || iettype == 5) {
    const double pi = 3.14159265358979323846;
    for (int j = 0; j < m.nbfr; ++j) {
      const double amig = m.tide[j], ff = m.tide[m.nbfr + j], face = m.tide[2 * m.nbfr + j];
      const double efa = m.efa[(int64_t)i * m.nbfr + j], emo = m.emo[(int64_t)i * m.nbfr + j];
      const int ncyc = (int)(amig * a.time / (2.0 * pi));
      const double arg = amig * a.time - (2.0 * pi) * ncyc + face - efa;
      r_eta1_bar += a.ramp * ff * emo * cos(arg);
    }
  }
  switch (iettype) {
    case 1:
    case 2:
    case 4: m.elbc[i] = r_eic + r_eth; break;
    case 3: m.elbc[i] = r_eic + r_eta1_bar; break;
    case 5: m.elbc[i] = r_eic + r_eta1_bar + r_eth; break;
    default: break;
  }
}

void launch_rhs(const DMesh& m, const StepArgs& a, hipStream_t st) {
  if (m.n > 0) hipLaunchKernelGGL(k_elbc, dim3((m.n + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, a);
  if (m.ns > 0) {
    // sides per workgroup: as many as fit 16 KB of LDS (measured best against 4-64 KB at 1M
    // sides x 31 levels), at most one per thread
    const int row = 8 * ((m.nvrt | 1) + ((2 * m.nvrt) | 1));
    const int g = max(1, min(kTpb, (16 * 1024) / row));
    hipLaunchKernelGGL(k_side_rhs, dim3((unsigned)((m.ns + g - 1) / g)), dim3(kTpb), (size_t)g * row, st, m, a, g);
  }
  if (m.ne > 0) hipLaunchKernelGGL(k_elem_ghat, dim3((m.ne + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, a);
}

// ---------------------------------------------------------------- momentum (the step after the solve)
// The inverse of stage_rows: `rows` LDS rows of `row_len` values (stride `stride`) to one
// contiguous span of HBM, with flat coalesced stores.
__device__ __forceinline__ void unstage_rows(double* __restrict__ dst, const double* __restrict__ src, int stride,
                                             int row_len, int rows) {
  const int total = rows * row_len;
  const float inv = 1.0f / (float)row_len;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int j;
    const int g = split_row(i, inv, row_len, j);
    dst[i] = src[g * stride + j];
  }
}

// Side::grad2 (meshdata.cu:1363-1398) of eta1 over the side's elements that are wet (isDry of
// the last level update) and of eta2 over those wet by eta2 (Element::isDry2, meshdata.cu:774-778:
// a node with eta2 + depth <= h0 makes its element dry), each the mean of the element gradients
// (Element::grad2, meshdata.h:1105-1114).
__device__ __forceinline__ void side_eta_grads(const DMesh& m, const int4 sr, double h0, double& g1x, double& g1y,
                                               double& g2x, double& g2y) {
  g1x = g1y = g2x = g2y = 0.0;
  int c1 = 0, c2 = 0;
  const int es[2] = {sr.z, sr.w};
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int e = es[j];
    if (e < 0) continue;
    const int4 r0 = m.elem[2 * e];
    const double* gp = m.gradphi + 6 * (int64_t)e;
    const double p0x = gp[0], p0y = gp[1], p1x = gp[2], p1y = gp[3], p2x = gp[4], p2y = gp[5];
    if (m.elem_wet[e]) {
      const double f0 = m.eta1[r0.x], f1 = m.eta1[r0.y], f2 = m.eta1[r0.z];
      g1x += f0 * p0x + f1 * p1x + f2 * p2x;
      g1y += f0 * p0y + f1 * p1y + f2 * p2y;
      ++c1;
    }
    const double f0 = m.eta2[r0.x], f1 = m.eta2[r0.y], f2 = m.eta2[r0.z];
    const bool dry2 = f0 + m.depth[r0.x] <= h0 || f1 + m.depth[r0.y] <= h0 || f2 + m.depth[r0.z] <= h0;
    if (!dry2) {
      g2x += f0 * p0x + f1 * p1x + f2 * p2x;
      g2y += f0 * p0y + f1 * p1y + f2 * p2y;
      ++c2;
    }
  }
  if (c1 > 0) {
    g1x = g1x * (1.0 / c1);
    g1y = g1y * (1.0 / c1);
  }
  if (c2 > 0) {
    g2x = g2x * (1.0 / c2);
    g2y = g2y * (1.0 / c2);
  }
}

// solveMomentumEquations (meshdata.cu:5074-5352) for one side, barotropic, itur=0 (dfv = dfv0),
// iveg=0, ics=1, imm=0 and no Coriolis, wind, horizontal viscosity, pressure or tidal potential
// (zero terms omitted), then applyMomentumBC for land sides (meshdata.cu:5356-5446, islip=0).
// zc and u are the side's rows in LDS; u holds btu on entry and u2 on exit. The 3D case is
// tridag_sch<2> (meshdata.cu:3665-3711) with the coefficients formed on the fly: the forward
// sweep keeps u_j in the btu slot k-1 and gam_j in the zc slot k-1 (both dead once row j is
// formed), the back substitution runs over those slots, and the result moves up one slot.
// Component c (0: x, 1: y) of one side's momentum solve, by one of the side's two lanes (adjacent
// lanes of one wave): both lanes compute the shared coefficients identically (the Thomas pivots,
// written to the same zc slots with the same values, in lockstep) and each solves its own
// component. The free-slip projection needs both components and runs in lane 0 after a wave
// fence. Every value has the operands and order of the one-lane form.
__device__ void momentum_side(const DMesh& m, const StepArgs& a, int s, int c, double* zc, double* u) {
  const int nv = m.nvrt;
  const int kb = m.side_kb[s];
  const int4 sr = m.side[s];
  const double rx = a.rmaxvel, ry = a.rmaxvel * 1.013;
  const double rc = c ? ry : rx;
  if (!m.side_wet[s]) {
    for (int k = 0; k < nv; ++k) u[2 * k + c] = 0.0;
  } else {
    double g1x, g1y, g2x, g2y;
    side_eta_grads(m, sr, a.h0, g1x, g1y, g2x, g2y);
    const double g1 = c ? g1y : g1x, g2 = c ? g2y : g2x;
    const double grav3 = 0.5 * (m.grav2[sr.x] + m.grav2[sr.y]);
    const double Cd = 0.5 * (m.cd[sr.x] + m.cd[sr.y]);
    const double dt = a.dt, th = a.theta;
    if (kb == nv - 2) {
      const double htot = zc[nv - 1] - zc[kb];
      const double hhat = m.hhat[s];
      const double hc = u[2 * (nv - 1) + c] + dt * (0.0 - grav3 * ((1.0 - th) * g1 + th * g2));
      const double vc = fmin(fmax(hc * hhat / htot, -rc), rc);
      for (int k = 0; k < nv; ++k) u[2 * k + c] = vc;
    } else {
      const double G = grav3 * (th * g2 + (1.0 - th) * g1);
      const double chi2 = m.chi[s], dfv0 = a.dfv0;
      double zm = zc[kb], z0 = zc[kb + 1];         // zc[k-1], zc[k]
      double bm = u[2 * kb + c], b0 = u[2 * kb + 2 + c];  // btu[k-1], btu[k] of this component
      double bet = 1.0, cprev = 0.0, up = 0.0;
      // dz/6, dz/3 and dt dfv0 / dz of layer (k, k+1); the next row reuses them for its layer
      // (k-1, k), whose dz is the same difference of the same two levels
      double d6p = 0.0, d3p = 0.0, tp = 0.0;
      for (int k = kb + 1; k < nv; ++k) {
        const bool top = k < nv - 1;
        const double zp = top ? zc[k + 1] : 0.0;
        const double bp = top ? u[2 * (k + 1) + c] : 0.0;
        const double dz_0 = z0 - zm;
        const double dz_1 = top ? zp - z0 : 0.0;
        double aj = 0.0, bj = 0.0, cj = 0.0, d6 = 0.0, d3 = 0.0, tq = 0.0;
        if (top) {
          tq = dt * dfv0 / dz_1;
          d3 = dz_1 / 3.0;
          d6 = d3 * 0.5;  // = dz_1 / 6.0 exactly: halving commutes with rounding for normal values
          cj += d6 - tq;
          bj += d3 + tq;
        }
        if (k > kb + 1) {
          aj += d6p - tp;
          bj += d3p + tp;
        } else {
          bj += dt * chi2;
        }
        double dz = 0.0;
        dz += dz_1;
        if (k > kb + 1) dz += dz_0;
        double q = -(0.5 * dz) * dt * G;
        if (top) q += d6 * (2.0 * b0 + bp);
        if (k > kb + 1) q += d6p * (2.0 * b0 + bm);
        d6p = d6;
        d3p = d3;
        tp = tq;
        double uk;
        if (k == kb + 1) {
          bet = bj;
          uk = q / bet;
        } else {
          const double gam = cprev / bet;
          bet = bj - aj * gam;
          uk = (q - aj * up) / bet;
          zc[k - 1] = gam;
        }
        u[2 * (k - 1) + c] = uk;
        up = uk;
        cprev = cj;
        zm = z0;
        z0 = zp;
        bm = b0;
        b0 = bp;
      }
      // back substitution: u_j (slot kb + j) -= gam_{j+1} (zc slot kb + j + 1) u_{j+1}
      for (int k = nv - 3; k >= kb; --k) {
        const double gam = zc[k + 1];
        up = u[2 * k + c] - gam * up;
        u[2 * k + c] = up;
      }
      for (int k = nv - 1; k > kb; --k) u[2 * k + c] = fmin(fmax(u[2 * (k - 1) + c], -rc), rc);
    }
    // bottom condition, imm = 0
    if (Cd == 0.0) {
      u[2 * kb + c] = u[2 * (kb + 1) + c];
    } else if (kb < nv - 2) {
      u[2 * kb + c] = 0.0;
    }
    for (int k = 0; k < kb; ++k) u[2 * k + c] = 0.0;
  }
}

// The free-slip projection of a land side (side_type 1), both components, by the side's lane 0
// after both lanes' momentum_side.
__device__ void momentum_free_slip(const DMesh& m, int s, double* u) {
  if (m.side_type[s] != 1) return;
  const int nv = m.nvrt;
  const int kb = m.side_kb[s];
  const int4 sr = m.side[s];
  // free slip: the tangential component, t = (-n.y, n.x), n from populateSides (meshdata.cu:2283-2287)
  const double vx = -(m.y[sr.x] - m.y[sr.y]), vy = m.x[sr.x] - m.x[sr.y];
  const double len = sqrt(vx * vx + vy * vy);
  const double nx = vx / len, ny = vy / len;
  const double tx = -ny, ty = nx;
  for (int k = kb; k < nv; ++k) {
    const double vtan = u[2 * k] * tx + u[2 * k + 1] * ty;
    u[2 * k] = vtan * tx;
    u[2 * k + 1] = vtan * ty;
  }
}

// A workgroup of 64 takes g <= 32 consecutive sides: their zc and btu rows come into LDS with flat
// coalesced loads, two lanes per side solve in LDS (one component each), the u2 rows leave with
// flat stores.
__global__ void __launch_bounds__(64) k_momentum(DMesh m, StepArgs a, int g) {
  extern __shared__ double lds[];
  const int nv = m.nvrt;
  const int zs = nv | 1, us = (2 * nv) | 1;
  double* Lz = lds;
  double* Lu = lds + g * zs;
  const int s0 = blockIdx.x * g;
  const int gc = min(g, m.ns - s0);
  stage_rows(Lz, zs, m.zc + (int64_t)s0 * nv, nv, gc);
  stage_rows(Lu, us, m.btu + 2 * (int64_t)s0 * nv, 2 * nv, gc);
  __syncthreads();
  const int t = threadIdx.x, sl = t >> 1, c = t & 1;  // two lanes per side
  if (sl < gc) momentum_side(m, a, s0 + sl, c, Lz + sl * zs, Lu + sl * us);
  wave_lds_sync();
  if (sl < gc && c == 0) momentum_free_slip(m, s0 + sl, Lu + sl * us);
  __syncthreads();
  unstage_rows(m.su2 + 2 * (int64_t)s0 * nv, Lu, us, 2 * nv, gc);
}

// shapiroFilter1 + shapiroFilter2 (meshdata.cu:5456-5518) in one pass that writes every side's
// final u2 into `out` (the caller then swaps it with u2). For a wet internal side, level k > kb
// takes bcc[k] = u2[k] + shapiro0 (0.25 sum_j u2_j[kin_j] - u2[k]) over its four neighbours
// Side::neigh(j) (meshdata.h:1409-1422: the sides before (j = 0, 1) and after (j = 2, 3) it in
// elements 0 and 1, summed in that order; kin_j = max(k, kb_j + 1) for a wet neighbour); a 2D
// side's bottom slot takes bcc[kb+1] (recomputed by its own thread, the same bits); levels below
// kb - 1 become 0; every other value, and boundary and dry sides, pass through unchanged.
// The workgroup's threads cover the levels of a few consecutive sides; each side's record (its
// filter flag and kb, its four neighbours and the lowest level each neighbour contributes,
// kb_j + 1 for a wet neighbour) is read once into LDS, so each (side, level) thread only gathers
// its five u2 values.
constexpr int kShSides = 256;  // at most 256 sides per workgroup (one level each)

__global__ void __launch_bounds__(256) k_shapiro(DMesh m, double shapiro0, double* __restrict__ out) {
  __shared__ int Lkb[kShSides], Lnb[kShSides][4], Lkin[kShSides][4];
  __shared__ int8_t Lf[kShSides];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x;
  const int nv = m.nvrt;
  const int64_t total = (int64_t)m.ns * nv;
  // a 32-bit division below 2^32 (the 64-bit one is a long instruction sequence)
  auto side_of = [&](int64_t tt) { return (tt < 0xffffffffll) ? (int)((uint32_t)tt / (uint32_t)nv) : (int)(tt / nv); };
  const int sfirst = side_of(t0), nsd = side_of(min(t0 + (int64_t)blockDim.x, total) - 1) - sfirst + 1;
  for (int q = threadIdx.x; q < nsd; q += blockDim.x) {
    const int sd = sfirst + q;
    const int4 sr = m.side[sd];
    const bool filtered = m.side_wet[sd] && sr.w >= 0;
    Lf[q] = filtered;
    Lkb[q] = m.side_kb[sd];
    if (filtered) {
      const int4 nb4 = reinterpret_cast<const int4*>(m.side_nbr)[sd];
      const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        Lnb[q][j] = nb[j];
        Lkin[q][j] = m.side_wet[nb[j]] ? m.side_kb[nb[j]] + 1 : -1;
      }
    }
  }
  __syncthreads();
  const int64_t t = t0 + threadIdx.x;
  if (t >= total) return;
  const int s = side_of(t), k = (int)(t - (int64_t)s * nv), q = s - sfirst;
  const double2* u2 = reinterpret_cast<const double2*>(m.su2);
  double2* o = reinterpret_cast<double2*>(out);
  const int kb = Lkb[q];
  const bool filtered = Lf[q] != 0;
  const int kf = (kb == nv - 2) ? kb + 1 : k;  // the level whose filtered value this slot takes
  if (!filtered || (k >= kb - 1 && (k < kb || (kb < nv - 2 && k == kb)))) {
    o[t] = u2[t];
    return;
  }
  if (k < kb - 1) {
    o[t] = make_double2(0.0, 0.0);
    return;
  }
  // u2 + shapiro0 (sum of the four neighbours' u2 / 4 - u2), neighbour j at level max(kf, kb_j + 1)
  // when wet, in Side::neigh order
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int kin = max(kf, Lkin[q][j]);
    const double2 v = u2[(int64_t)Lnb[q][j] * nv + kin];
    sx += v.x;
    sy += v.y;
  }
  const double2 u = u2[(int64_t)s * nv + kf];
  o[t] = make_double2(u.x + shapiro0 * (0.25 * sx - u.x), u.y + shapiro0 * (0.25 * sy - u.y));
}

void launch_shapiro(const DMesh& m, double shapiro0, double* out, hipStream_t st) {
  const int64_t total = (int64_t)m.ns * m.nvrt;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_shapiro, dim3((unsigned)((total + kTpb - 1) / kTpb)), dim3(kTpb), 0, st, m, shapiro0, out);
}

void launch_momentum(const DMesh& m, const StepArgs& a, hipStream_t st) {
  if (m.ns <= 0) return;
  // sides per workgroup: one per two lanes, as many as 24 KB of LDS holds (measured best against
  // 16-64 KB and 32-256 threads at 3.1M sides x 31 levels)
  const int row = 8 * ((m.nvrt | 1) + ((2 * m.nvrt) | 1));
  const int g = max(1, min(32, (24 * 1024) / row));
  hipLaunchKernelGGL(k_momentum, dim3((unsigned)((m.ns + g - 1) / g)), dim3(64), (size_t)g * row, st, m, a, g);
}

// ---------------------------------------------------------------- vertical velocity
// computeVerticalVelocity (meshdata.cu:5978-6102), triangles, ics=1, imm=0: element w from the
// flux balance of its prisms, bottom up, with the levels of the last level update (node z from
// the elevation of that update, side zc and kb) and the side velocities u2. A workgroup takes E
// elements. Phase 1, one thread per (element, level l): the side fluxes of layer (l, l+1), the
// horizontal flux through the face at level l+1 and that face's n_z A (Element::n,
// meshdata.cu:559-631); phase 2, one thread per element: the running top flux
// top_l = -side_l - bottom_l, bottom_l+1 = -top_l, in the reference's order; phase 3, one thread
// per (element, level): w = (top - horizontal) / (n_z A), stored with coalesced writes.

// Per-element values that every level of phase 1 reads: node positions and level data, side
// normals (populateSides, meshdata.cu:2283-2287), orientation signs, lengths and bottom levels.
struct VvElem {
  double px[3], py[3], eta[3], h[3];
  double snx[3], sny[3], sign[3], slen[3];
  int nkb[3], s[3], skb[3], dry[3], nd[3];
};

__global__ void __launch_bounds__(256) k_vertical_velocity(DMesh m, double h0, int E) {
  extern __shared__ double lds[];
  const int nv = m.nvrt, nvp = nv - 1;
  double* Lsf = lds;            // [E][nvp]: side flux of layer (l, l+1), then the top flux
  double* Lth = Lsf + E * nvp;  // [E][nvp]: horizontal flux through the face at level l+1
  double* Lden = Lth + E * nvp; // [E][nvp]: n_z A of that face
  double* Lbw = Lden + E * nvp; // [E]: w at the element's bottom level
  VvElem* Lel = reinterpret_cast<VvElem*>(Lbw + E);  // [E]
  const int e0 = blockIdx.x * E;
  const int ec = min(E, m.ne - e0);
  // three threads per element, thread j taking corner j and side j
  for (int t = threadIdx.x; t < 3 * ec; t += blockDim.x) {
    const int el = t / 3, j = t - 3 * el;
    const int e = e0 + el;
    VvElem& q = Lel[el];
    const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
    const int nd[3] = {r0.x, r0.y, r0.z};
    const int sd[3] = {r0.w, r1.x, r1.y};
    {
      const int i = nd[j];
      q.px[j] = m.x[i];
      q.py[j] = m.y[i];
      q.eta[j] = m.eta_lev[i];
      q.h[j] = m.depth[i];
      q.nkb[j] = m.kb[i];
      q.nd[j] = i;
      q.dry[j] = q.eta[j] + q.h[j] <= h0;  // isDry2 at the level update
      const int s = sd[j];
      const int4 sr = m.side[s];
      const double vx = -(m.y[sr.x] - m.y[sr.y]), vy = m.x[sr.x] - m.x[sr.y];
      const double ql = sqrt(vx * vx + vy * vy);
      q.snx[j] = vx / ql;
      q.sny[j] = vy / ql;
      q.sign[j] = (sr.z == e) ? 1.0 : -1.0;
      q.slen[j] = m.side_len[s];
      q.s[j] = s;
      q.skb[j] = m.side_kb[s];
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < ec * nvp; idx += blockDim.x) {
    const int el = idx / nvp, l = idx - el * nvp;
    const int e = e0 + el;
    if (!m.elem_wet[e] || l < m.elem_kb[e]) continue;
    const VvElem& q = Lel[el];
    // the face at level l+1 (node levels clamped to each node's bed), zcoord of the level update
    double pz[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int k = max(l + 1, q.nkb[j]);
      pz[j] = q.dry[j] ? 0.0 : z_wet(m, q.nd[j], q.eta[j], q.h[j], q.nkb[j], k);
    }
    const double n20x = q.px[2] - q.px[0], n20y = q.py[2] - q.py[0], n20z = pz[2] - pz[0];
    const double n10x = q.px[1] - q.px[0], n10y = q.py[1] - q.py[0], n10z = pz[1] - pz[0];
    const double nx = (n10y) * (n20z) - (n20y) * (n10z);
    const double ny = (n20x) * (n10z) - (n10x) * (n20z);
    const double nz = 2.0 * m.area[e];
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    const double ta = 0.5 * len;
    const double inv_len = 1.0 / len;
    const double tnx = nx * inv_len, tny = ny * inv_len, tnz = nz * inv_len;
    double sf = 0.0, vx[3], vy[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int s = q.s[j], kbs = q.skb[j];
      const int kt = max(l + 1, kbs), kl = max(l, kbs);
      const double2* u2 = reinterpret_cast<const double2*>(m.su2) + (int64_t)s * nv;
      const double* zc = m.zc + (int64_t)s * nv;
      const double2 ut = u2[kt], ul = u2[kl];
      const double ucx = 0.5 * (ut.x + ul.x), ucy = 0.5 * (ut.y + ul.y);
      const double sideArea = q.slen[j] * (zc[kt] - zc[kl]);
      sf += q.sign[j] * (ucx * q.snx[j] + ucy * q.sny[j]) * sideArea;
      vx[j] = ut.x;
      vy[j] = ut.y;
    }
    const double utx = (vx[0] + vx[1] + vx[2] + 0.0) / 3, uty = (vy[0] + vy[1] + vy[2] + 0.0) / 3;
    Lsf[idx] = sf;
    Lth[idx] = (utx * tnx + uty * tny) * ta;
    Lden[idx] = tnz * ta;
  }
  __syncthreads();
  if (threadIdx.x < ec) {
    const int el = threadIdx.x, e = e0 + el;
    if (m.elem_wet[e]) {
      const VvElem& q = Lel[el];
      const int kb = m.elem_kb[e];
      const double* gp = m.gradphi + 6 * (int64_t)e;
      const double ghx = q.h[0] * gp[0] + q.h[1] * gp[2] + q.h[2] * gp[4];
      const double ghy = q.h[0] * gp[1] + q.h[1] * gp[3] + q.h[2] * gp[5];
      double bx[3], by[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double2 u = reinterpret_cast<const double2*>(m.su2)[(int64_t)q.s[j] * nv + q.skb[j]];
        bx[j] = u.x;
        by[j] = u.y;
      }
      const double ubx = (bx[0] + bx[1] + bx[2] + 0.0) / 3, uby = (by[0] + by[1] + by[2] + 0.0) / 3;
      const double bottomW = -(ghx * ubx + ghy * uby);
      Lbw[el] = 0.0 + bottomW;
      // the running sums in level order; the side fluxes of 8 levels are read ahead so that
      // the dependent additions do not wait for LDS one at a time
      double bottom = 0.0;
      double* row = Lsf + el * nvp;
      for (int l0 = kb; l0 < nvp; l0 += 8) {
        double f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) f[u] = (l0 + u < nvp) ? row[l0 + u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (l0 + u < nvp) {
            const double top = -f[u] - bottom;
            row[l0 + u] = top;
            bottom = -top;
          }
      }
    }
  }
  __syncthreads();
  double* w = m.ew + (int64_t)e0 * nv;
  for (int idx = threadIdx.x; idx < ec * nv; idx += blockDim.x) {
    const int el = idx / nv, k = idx - el * nv;
    const int e = e0 + el;
    double v = 0.0;
    if (m.elem_wet[e]) {
      const int kb = m.elem_kb[e];
      if (k == kb) {
        v = Lbw[el];
      } else if (k > kb) {
        const int j = el * nvp + k - 1;
        v = (Lsf[j] - Lth[j]) / Lden[j];
      }
    }
    w[idx] = v;
  }
}

void launch_vertical_velocity(const DMesh& m, double h0, hipStream_t st) {
  if (m.ne <= 0 || m.nvrt < 2) return;
  const int nvp = m.nvrt - 1;
  // elements per workgroup: about one thread per (element, level), within 32 KB of LDS
  const size_t per_elem = 8 * (3 * (size_t)nvp + 1) + sizeof(VvElem);
  const int E = max(1, min(256 / nvp, (int)((32 * 1024) / per_elem)));
  const size_t lds = per_elem * E;
  hipLaunchKernelGGL(k_vertical_velocity, dim3((unsigned)((m.ne + E - 1) / E)), dim3(256), lds, st, m, h0, E);
}

// ---------------------------------------------------------------- a11+a12: assembly + BC
// setupCoefficientMatrixAndRHS for triangles, imm=0 (meshdata.cu:3182-3311), fused with
// setEssentialBC (meshdata.cu:4400-4446), in two passes. k_elem_assemble evaluates, once per
// element and corner id1, the terms the reference adds when node nodes[id1] visits the element
// in its ball: the three matrix entries (diagonal, next, previous corner) and the five RHS
// increments in their order (advective/ghat term, eta terms of corners 0-2, volume source).
// k_assemble then gives each node its CSR row (node-gather, no atomics) by adding those terms
// over its ball in ball order, so every sum has the reference's operands and order.
// Open nodes of elevation type 0 (no essential BC) add the boundary integral I_5; I_3 is zero
// because every open boundary here has flow type ifltype = 0 (meshdata.cu:3373-3470).
__global__ void __launch_bounds__(256) k_elem_assemble(DMesh m, StepArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int4 r0 = m.elem[2 * e], r1 = m.elem[2 * e + 1];
  const int nd[3] = {r0.x, r0.y, r0.z};
  const int sd[3] = {r0.w, r1.x, r1.y};
  const double dt = a.dt, th = a.theta;
  const double grav3 = (m.grav2[nd[0]] + m.grav2[nd[1]] + m.grav2[nd[2]] + 0.0) / 3;
  const double hhatb = (m.hhat[sd[0]] + m.hhat[sd[1]] + m.hhat[sd[2]] + 0.0) / 3;
  const double area = m.area[e];
  const double Ux = (m.sU[2 * sd[0]] + m.sU[2 * sd[1]] + m.sU[2 * sd[2]] + 0.0) / 3;
  const double Uy = (m.sU[2 * sd[0] + 1] + m.sU[2 * sd[1] + 1] + m.sU[2 * sd[2] + 1] + 0.0) / 3;
  const double* g = m.gradphi + 6 * (int64_t)e;
  double gp[6];
  for (int l = 0; l < 6; ++l) gp[l] = g[l];
  const double ghx = m.ghat1[2 * e], ghy = m.ghat1[2 * e + 1];
  const double px[3] = {m.x[nd[0]], m.x[nd[1]], m.x[nd[2]]}, py[3] = {m.y[nd[0]], m.y[nd[1]], m.y[nd[2]]};
  const double eta[3] = {m.eta2[nd[0]], m.eta2[nd[1]], m.eta2[nd[2]]};
  const double wet = m.elem_wet[e] ? 1.0 : 0.0;
  const double src = (dt / 3) * m.vsource[e];
  const double fric = grav3 * th * (1.0 - th) * dt * dt * hhatb;
  const double c = grav3 * th * th * dt * dt * hhatb;
  double2* out = reinterpret_cast<double2*>(m.econ + 24 * (int64_t)e);
  for (int id1 = 0; id1 < 3; ++id1) {
    const double g1x = gp[2 * id1], g1y = gp[2 * id1 + 1];
    const double I41 = g1x * Ux + g1y * Uy;
    const double I42 = g1x * ghx + g1y * ghy;
    const double adv = dt * area * ((1 - th) * I41 + th * I42);
    const int id2 = (id1 + 1) % 3, id3 = (id1 + 2) % 3;
    const double x0x = px[id1], x0y = py[id1];
    const double x1x = px[id2], x1y = py[id2];
    const double x2x = px[id3], x2y = py[id3];
    const double ipp_ii = (1.0 / 6.0) * area, ipp_ij = (1.0 / 12.0) * area, ipp_ik = (1.0 / 12.0) * area;
    const double ax = x2x - x1x, ay = x2y - x1y;
    const double igpgp_ii = ((1.0 / 4.0) / area) * (ax * ax + ay * ay);
    const double igpgp_ij = ((1.0 / 4.0) / area) * (ax * (x0x - x2x) + ay * (x0y - x2y));
    const double igpgp_ik = -igpgp_ii - igpgp_ij;
    double q[3];
    for (int l = 0; l < 3; ++l) {
      const double ipp = ((id1 == l) ? (2.0 / 12.0) : (1.0 / 12.0)) * area;
      const double igpgp = area * (g1x * gp[2 * l] + g1y * gp[2 * l + 1]);
      q[l] = (eta[l] + 0.0) * ipp - eta[l] * fric * igpgp * wet;
    }
    out[4 * id1 + 0] = make_double2(ipp_ii + c * igpgp_ii, ipp_ij + c * igpgp_ij);
    out[4 * id1 + 1] = make_double2(ipp_ik + c * igpgp_ik, adv);
    out[4 * id1 + 2] = make_double2(q[0], q[1]);
    out[4 * id1 + 3] = make_double2(q[2], src);
  }
}

// Row accumulation for k_assemble: in LDS for rows of up to kRowLds entries, in the CSR values
// themselves for longer rows; both add the same terms in the same order.
constexpr int kRowLds = 16;

template <class Row>
__device__ __forceinline__ double assemble_row(const DMesh& m, const StepArgs& a, int i, int len, int diag, Row rv) {
  for (int k = 0; k < len; ++k) rv(k) = 0.0;
  double qel = 0.0;
  const double2* con = reinterpret_cast<const double2*>(m.econ);
  for (int j = m.ball_ptr[i]; j < m.ball_ptr[i + 1]; ++j) {
    const int64_t base = 4 * (3 * (int64_t)m.ball_elem[j] + m.ball_loc[j]);
    const double2 c0 = con[base], c1 = con[base + 1], c2 = con[base + 2], c3 = con[base + 3];
    qel += c1.y;
    rv(diag) += c0.x;
    rv(m.ball_slot[2 * j]) += c0.y;
    rv(m.ball_slot[2 * j + 1]) += c1.x;
    qel += c2.x;
    qel += c2.y;
    qel += c3.x;
    qel += c3.y;
  }
  // I_5 of an open node without an essential BC (meshdata.cu:3373-3380, 3457-3463): its two
  // boundary sides (open or land, ascending side index), the non-dry ones.
  if (m.open_bnd[i] && node_etype(m, a, i) == 0) {
    const int2 bs = m.node_bside[i];
    const int sl[2] = {bs.x, bs.y};
    for (int l = 0; l < 2; ++l) {
      const int s = sl[l];
      if (s < 0 || !m.side_wet[s]) continue;
      const int4 sr = m.side[s];
      const double vx = -(m.y[sr.x] - m.y[sr.y]), vy = m.x[sr.x] - m.x[sr.y];
      const double ql = sqrt(vx * vx + vy * vy);
      const double Unbar = m.sU[2 * s] * (vx / ql) + m.sU[2 * s + 1] * (vy / ql);
      qel -= 0.5 * (1.0 - a.theta) * a.dt * m.side_len[s] * Unbar;
    }
  }
  if (m.open_bnd[i] && node_etype(m, a, i) != 0) {
    for (int k = 0; k < len; ++k) rv(k) = 0.0;
    rv(diag) = 1.0;
    qel = m.elbc[i];
  } else {
    for (int j = m.nbr_ptr[i]; j < m.nbr_ptr[i + 1]; ++j) {
      const int nb = m.nbr_node[j];
      if (m.open_bnd[nb] && node_etype(m, a, nb) != 0) {
        const int sl = m.csr_slot[j];
        qel -= rv(sl) * m.elbc[nb];
        rv(sl) = 0.0;
      }
    }
  }
  return qel;
}

__global__ void __launch_bounds__(256) k_assemble(DMesh m, StepArgs a, double* __restrict__ val,
                                                  double* __restrict__ b) {
  __shared__ double rows[kRowLds][256];  // entry k of thread t's row at rows[k][t]
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  const int row = m.indptr[i];
  const int len = m.indptr[i + 1] - row;
  const int diag = m.diag_off[i];
  double qel;
  if (len <= kRowLds) {
    const int t = threadIdx.x;
    qel = assemble_row(m, a, i, len, diag, [&](int k) -> double& { return rows[k][t]; });
    for (int k = 0; k < len; ++k) val[row + k] = rows[k][t];
  } else {
    double* rv = val + row;
    qel = assemble_row(m, a, i, len, diag, [&](int k) -> double& { return rv[k]; });
  }
  b[i] = qel;
}

void launch_assemble(const DMesh& m, const StepArgs& a, double* val, double* b, hipStream_t st) {
  if (m.ne > 0) hipLaunchKernelGGL(k_elem_assemble, dim3((m.ne + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, a);
  if (m.n > 0) hipLaunchKernelGGL(k_assemble, dim3((m.n + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, a, val, b);
}

// rollUpElevation + updateElevation (meshdata.cu:3538-3544, 4449-4456)
__global__ void k_update_elevation(DMesh m, const double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  m.eta1[i] = m.eta2[i];
  m.eta2[i] = x[i];
}

void launch_update_elevation(const DMesh& m, const double* x, hipStream_t st) {
  if (m.n > 0) hipLaunchKernelGGL(k_update_elevation, dim3((m.n + kTpb - 1) / kTpb), dim3(kTpb), 0, st, m, x);
}

}  // namespace gsh
