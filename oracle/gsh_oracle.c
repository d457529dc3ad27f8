/*
 * gsh_oracle.c -- CPU FP64 restatement of the reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity checker and the CPU baseline of the
 * benchmark; the product (libgsh) never links, loads or calls it.
 *
 * It restates, in the reference's own operation order, the gschism code paths
 * (paths relative to the reference checkout):
 *   src/grid/read_and_process_grid.cpp:27-174, 272-437, 496-574   topology
 *   src/kernels/meshdata.cu:1903-1958, 2094-2141                   CSR pattern
 *   src/kernels/meshdata.cu:2274-2312, 2350-2440, 749-771           geometry
 *   src/kernels/meshdata.cu:5521-5612, 5679-5760, 1676-1807, 1647    levels (SZ, LSC2), wet/dry
 *   src/grid/read_and_process_grid.cpp:1497-1502                    SZ stretching C(sigma)
 *   src/kernels/meshdata.cu:3547-3660, 1449-1629, 85-127, 635-746,
 *                           990-1043, 1209-1291                     ELM backtracking
 *   src/kernels/meshdata.cu:4459-4476, 4615-4725, 6183-6424, 4276-4339   RHS producers
 *   src/kernels/meshdata.cu:3182-3311, 3373-3380, 3457-3463, 4400-4446   assembly, I_5, essential BC
 *   src/solver/pcg_solver.py:172-276, 881-947, 1075-1124; pcg_loop.cpp:361-413   Jacobi PCG
 *                                                                   (single, and a multi-RHS block)
 *   src/kernels/meshdata.cu:5765-5929                               nodal velocities
 *   src/kernels/meshdata.cu:5045-5446, 3665-3711, 5456-5518         momentum, Shapiro filter
 *   src/kernels/meshdata.cu:5978-6102, 559-631                      element vertical velocity
 * for the option subset ics=1, ivcor=1 (one sigma per level, or per node and level: LSC2) and
 * ivcor=2 (SZ), imm=0, iveg=0, ibc=1 (barotropic),
 * itur=0, islip=0, nws=0, ncor=0, ihorcon=0, nchi=0, triangles; elevation BC types 0-5 per node.
 *
 * Deliberate deviations, shared with the device code:
 *  - Dot products are summed in the device kernels' fixed tile / wave64 tree (see
 *    or_dot_canon) instead of the reference's atomics (pcg_solver.py:134-167), whose order
 *    is not reproducible; with identical order the CPU and device iterates agree bit for bit.
 *  - The celestial walk's assert on exceeding 40 steps (meshdata.cu:126) and the ELM asserts
 *    become status flags; the walk then keeps the current element.
 *  - Terms of computeElemAverageGhat and of the momentum solve that are identically zero in
 *    this subset (Coriolis, wind stress, viscosity, pressure, tidal potential, baroclinic) are
 *    omitted; they could only change the sign of a zero.
 *  - The nodal velocities sum each node's ball elements in ball order instead of the
 *    reference's per-element atomics.
 * Compile with -ffp-contract=off and without fast-math.
 */
#include "gsh_oracle.h"

/* OR_PLANT = k (k > 0) builds the oracle with one deliberate mistake, so that the CPU tests can
 * show that their pins catch it (oracle/build.py: build_planted). The default build has none. */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BT_WALK_CAP 1
#define BT_ITER38 2
#define BT_STUCK 4
#define BT_OPEN_EXIT 8
#define BT_SLIDE 16

static const double small = 1.0e-7; /* meshdata.h:10 */

static char g_err[512];
const char* or_error(void) { return g_err; }
static void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

struct or_mesh {
  int32_t n, ne, ns, nvrt;
  double *x, *y, *depth, *sigma;
  int8_t* open;
  int32_t* kb;
  int32_t* elnode;
  /* topology */
  int32_t *ic3, *elside, *isdel, *isidenode;
  int8_t *isbnd, *side_type;
  int32_t *nne_ptr, *nne;   /* node -> elements (counter-clockwise after reordering) */
  int32_t *nnp_ptr, *nnp;   /* node -> neighbour nodes */
  int32_t *indptr, *indices;
  int16_t* csr_slot;
  int32_t* diag_slot;
  /* geometry */
  double *area, *gradphi, *ssign, *scx, *scy, *slen, *snx, *sny;
  /* state */
  double *eta2, *eta1, *uvw, *su2, *btu, *cd, *grav2, *eth, *vsource, *elbc, *ew;
  int32_t *foot, *bstatus;
  double *sU, *hhat, *chi, *ghat1, *zc, *nz;
  int8_t *node_dry, *elem_dry, *side_dry;
  int32_t *elem_kb, *side_kb;
  /* elevation BC type per node (-1: the parameter iettype) and tidal constituents */
  int8_t* etype;
  int32_t nbfr;
  double *amig, *ff, *face, *emo, *efa;
  /* vertical coordinates: vmode 0 = ivcor 1 with sigma per level, 1 = ivcor 1 with a sigma per
   * node and level (sig_node [n][nvrt]), 2 = ivcor 2 (SZ: ztot [nz], ssig and cs [nvrt-nz+1]) */
  int32_t vmode, izlev;
  double h_s, h_c;
  double *sig_node, *ztot, *ssig, *cs;
};

static int nxq(int offset, int index, int size) { return (index + offset + 16 * size) % size; }

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

void or_mesh_free(or_mesh* m) {
  if (!m) return;
  void* ptrs[] = {m->x, m->y, m->depth, m->sigma, m->open, m->kb, m->elnode, m->ic3, m->elside, m->isdel,
                  m->isidenode, m->isbnd, m->side_type, m->nne_ptr, m->nne, m->nnp_ptr, m->nnp, m->indptr,
                  m->indices, m->csr_slot, m->diag_slot, m->area, m->gradphi, m->ssign, m->scx, m->scy,
                  m->slen, m->snx, m->sny, m->eta2, m->eta1, m->uvw, m->su2, m->btu, m->cd, m->grav2,
                  m->eth, m->vsource, m->elbc, m->foot, m->bstatus, m->sU, m->hhat, m->chi, m->ghat1,
                  m->zc, m->nz, m->node_dry, m->elem_dry, m->side_dry, m->elem_kb, m->side_kb, m->ew,
                  m->etype, m->amig, m->ff, m->face, m->emo, m->efa, m->sig_node, m->ztot, m->ssig, m->cs};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(m);
}

static int local_index(const or_mesh* m, int e, int node) {
  for (int l = 0; l < 3; ++l)
    if (m->elnode[3 * e + l] == node) return l;
  return -1;
}

/* ------------------------------------------------------------------ topology */
static int build_topology(or_mesh* m) {
  const int n = m->n, ne = m->ne;
  const int32_t* en = m->elnode;
  /* buildListsOfElementNeighbors (read_and_process_grid.cpp:27-84) */
  m->nne_ptr = xcalloc(n + 1, sizeof(int32_t));
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < 3; ++j) {
      if (en[3 * e + j] < 0 || en[3 * e + j] >= n) {
        set_err("element %d has an invalid node", e);
        return -1;
      }
      m->nne_ptr[en[3 * e + j] + 1]++;
    }
  for (int i = 0; i < n; ++i)
    if (m->nne_ptr[i + 1] == 0) {
      set_err("hanging node %d", i);
      return -1;
    }
  for (int i = 0; i < n; ++i) m->nne_ptr[i + 1] += m->nne_ptr[i];
  m->nne = xcalloc(m->nne_ptr[n], sizeof(int32_t));
  {
    int32_t* cnt = xcalloc(n, sizeof(int32_t));
    for (int e = 0; e < ne; ++e)
      for (int j = 0; j < 3; ++j) {
        const int v = en[3 * e + j];
        m->nne[m->nne_ptr[v] + cnt[v]++] = e;
      }
    free(cnt);
  }
  /* buildElementSideElementTable, first part (read_and_process_grid.cpp:272-344) */
  m->ic3 = xcalloc(3 * (size_t)ne, sizeof(int32_t));
  for (int ie = 0; ie < ne; ++ie)
    for (int k = 0; k < 3; ++k) {
      m->ic3[3 * ie + k] = -1;
      const int n1 = en[3 * ie + nxq(1, k, 3)], n2 = en[3 * ie + nxq(2, k, 3)];
      for (int l = m->nne_ptr[n1]; l < m->nne_ptr[n1 + 1]; ++l) {
        const int je = m->nne[l];
        if (je != ie) {
          int found = 0;
          for (int q = 0; q < 3; ++q)
            if (en[3 * je + q] == n2) found = 1;
          if (found) {
            m->ic3[3 * ie + k] = je;
            break;
          }
        }
      }
      const int je = m->ic3[3 * ie + k];
      if (je != -1)
        for (int l = 0; l < 3; ++l)
          if (en[3 * je + nxq(0, l, 3)] == n1 && en[3 * je + nxq(1, l, 3)] == n2) {
            set_err("elements %d and %d have opposite orientation", ie, je);
            return -1;
          }
    }
  /* boundary nodes: the hgrid boundary lists name every node of a side without a neighbour */
  m->isbnd = xcalloc(n, 1);
  for (int ie = 0; ie < ne; ++ie)
    for (int k = 0; k < 3; ++k)
      if (m->ic3[3 * ie + k] == -1) {
        m->isbnd[en[3 * ie + nxq(1, k, 3)]] = 1;
        m->isbnd[en[3 * ie + nxq(2, k, 3)]] = 1;
      }
  /* reorderElementsCounterClockwise (read_and_process_grid.cpp:496-574) */
  for (int i = 0; i < n; ++i) {
    int32_t* row = m->nne + m->nne_ptr[i];
    const int rs = m->nne_ptr[i + 1] - m->nne_ptr[i];
    if (m->isbnd[i]) {
      int icount = 0;
      for (int l = 0; l < rs; ++l) {
        const int ie = row[l];
        const int il = local_index(m, ie, i);
        if (m->ic3[3 * ie + nxq(-1, il, 3)] == -1) {
          icount++;
          row[0] = ie;
        }
      }
      if (icount != 1) {
        set_err("illegal boundary node %d: %d boundary elements", i, icount);
        return -1;
      }
    }
    for (int j = 1; j < rs; ++j) {
      const int ie = row[j - 1];
      const int il = local_index(m, ie, i);
      const int ne2 = m->ic3[3 * ie + nxq(-2, il, 3)];
      if (ne2 == -1) {
        set_err("incomplete ball around node %d", i);
        return -1;
      }
      row[j] = ne2;
    }
  }
  /* buildListsOfNodeNeighbors (read_and_process_grid.cpp:87-174) */
  m->nnp_ptr = xcalloc(n + 1, sizeof(int32_t));
  for (int i = 0; i < n; ++i) {
    int c = 0;
    const int rs = m->nne_ptr[i + 1] - m->nne_ptr[i];
    for (int j = 0; j < rs; ++j) c += (j == rs - 1 && m->isbnd[i]) ? 3 - 1 : 3 - 2;
    m->nnp_ptr[i + 1] = m->nnp_ptr[i] + c;
  }
  m->nnp = xcalloc(m->nnp_ptr[n], sizeof(int32_t));
  for (int i = 0; i < n; ++i) {
    int cur = m->nnp_ptr[i];
    const int rs = m->nne_ptr[i + 1] - m->nne_ptr[i];
    for (int j = 0; j < rs; ++j) {
      const int ie = m->nne[m->nne_ptr[i] + j];
      const int ii = local_index(m, ie, i);
      if (ii < 0) {
        set_err("node %d not found in element %d", i, ie);
        return -1;
      }
      for (int k = 1; k < 3 - 1; ++k) m->nnp[cur++] = en[3 * ie + nxq(k, ii, 3)];
      if (j == rs - 1) {
        const int last = en[3 * ie + nxq(3 - 1, ii, 3)];
        if (!m->isbnd[i]) {
          if (last != m->nnp[m->nnp_ptr[i]]) {
            set_err("broken ball for node %d", i);
            return -1;
          }
        } else {
          m->nnp[cur++] = last;
        }
      }
    }
  }
  /* side numbering, isdel, isidenode (read_and_process_grid.cpp:346-437) */
  m->elside = xcalloc(3 * (size_t)ne, sizeof(int32_t));
  for (int k = 0; k < 3 * ne; ++k) m->elside[k] = -1;
  int ns = 0;
  for (int ie = 0; ie < ne; ++ie)
    for (int j = 0; j < 3; ++j) {
      if (m->ic3[3 * ie + j] != -1) {
        if (ie < m->ic3[3 * ie + j]) {
          m->elside[3 * ie + j] = ns;
          const int je = m->ic3[3 * ie + j];
          int l = -1;
          for (int k = 0; k < 3; ++k)
            if (m->ic3[3 * je + k] == ie) {
              l = k;
              break;
            }
          if (l == -1) {
            set_err("wrong ball info for elements %d", ie);
            return -1;
          }
          m->elside[3 * je + l] = ns;
          ns++;
        }
      } else {
        m->elside[3 * ie + j] = ns;
        ns++;
      }
    }
  if (ns < ne || ns < n) {
    set_err("weird grid with ns < ne or ns < np");
    return -1;
  }
  m->ns = ns;
  m->isdel = xcalloc(2 * (size_t)ns, sizeof(int32_t));
  m->isidenode = xcalloc(2 * (size_t)ns, sizeof(int32_t));
  for (int ie = 0; ie < ne; ++ie)
    for (int j = 0; j < 3; ++j) {
      const int jsj = m->elside[3 * ie + j];
      const int n1 = en[3 * ie + nxq(1, j, 3)], n2 = en[3 * ie + nxq(2, j, 3)];
      const int ic = m->ic3[3 * ie + j];
      if (ic == -1 || (ic > -1 && ie < ic)) {
        m->isdel[2 * jsj] = ie;
        m->isdel[2 * jsj + 1] = ic;
        m->isidenode[2 * jsj] = n1;
        m->isidenode[2 * jsj + 1] = n2;
      } else {
        m->isdel[2 * jsj] = ic;
        m->isdel[2 * jsj + 1] = ie;
        m->isidenode[2 * jsj] = n2;
        m->isidenode[2 * jsj + 1] = n1;
      }
    }
  /* boundary segments: a boundary side is open when both of its nodes are open nodes */
  m->side_type = xcalloc(ns, 1);
  for (int s = 0; s < ns; ++s)
    if (m->isdel[2 * s + 1] == -1)
      m->side_type[s] = (m->open[m->isidenode[2 * s]] && m->open[m->isidenode[2 * s + 1]]) ? 2 : 1;
  /* an open node links at most two boundary sides (the assert of NodeBoundary::linkSideBoundary,
   * meshdata.h:1544-1555) */
  {
    int8_t* cnt = xcalloc(n, 1);
    for (int s = 0; s < ns; ++s)
      if (m->isdel[2 * s + 1] == -1)
        for (int j = 0; j < 2; ++j) {
          const int i = m->isidenode[2 * s + j];
          if (m->open[i] && ++cnt[i] > 2) {
            free(cnt);
            set_err("open boundary node %d ends more than two boundary sides", i);
            return -1;
          }
        }
    free(cnt);
  }
  /* populateNodes / populateNodes2 (meshdata.cu:1931-1935, 2094-2141) */
  m->indptr = xcalloc(n + 1, sizeof(int32_t));
  for (int i = 0; i <= n; ++i) m->indptr[i] = m->nnp_ptr[i] + i;
  m->indices = xcalloc(m->indptr[n], sizeof(int32_t));
  m->csr_slot = xcalloc(m->nnp_ptr[n], sizeof(int16_t));
  m->diag_slot = xcalloc(n, sizeof(int32_t));
  for (int i = 0; i < n; ++i) {
    const int b = m->nnp_ptr[i], nn = m->nnp_ptr[i + 1] - b;
    for (int a = 0; a < nn; ++a) {
      const int col = m->nnp[b + a];
      int16_t idx = 0;
      if (i < col) ++idx;
      for (int j = 0; j < nn; ++j)
        if (m->nnp[b + j] < col) ++idx;
      m->csr_slot[b + a] = idx;
    }
    int16_t d = 0;
    for (int j = 0; j < nn; ++j)
      if (m->nnp[b + j] < i) ++d;
    m->diag_slot[i] = d;
    const int rs = m->indptr[i];
    m->indices[rs + d] = i;
    for (int j = 0; j < nn; ++j) m->indices[rs + m->csr_slot[b + j]] = m->nnp[b + j];
  }
  return 0;
}

/* ------------------------------------------------------------------ geometry */
static double signed_area(double p1x, double p1y, double p2x, double p2y, double p3x, double p3y) {
  /* signedArea (meshdata.h:355-358) */
  return 0.5 * ((p1x - p3x) * (p2y - p3y) - (p2x - p3x) * (p1y - p3y));
}

static int build_geometry(or_mesh* m) {
  const int ne = m->ne, ns = m->ns;
  m->area = xcalloc(ne, sizeof(double));
  m->gradphi = xcalloc(6 * (size_t)ne, sizeof(double));
  m->ssign = xcalloc(3 * (size_t)ne, sizeof(double));
  for (int e = 0; e < ne; ++e) {
    const int* nd = m->elnode + 3 * e;
    /* computeSignedArea (meshdata.cu:749-771), triangles */
    const double area = signed_area(m->x[nd[0]], m->y[nd[0]], m->x[nd[1]], m->y[nd[1]], m->x[nd[2]], m->y[nd[2]]);
    if (!(area > 0.0)) {
      set_err("element %d is not counter-clockwise", e);
      return -1;
    }
    m->area[e] = area;
    for (int j = 0; j < 3; ++j) {
      /* ssign (meshdata.cu:2405-2414) */
      const int s = m->elside[3 * e + j];
      m->ssign[3 * e + j] = (m->isdel[2 * s] == e) ? 1.0 : -1.0;
      /* gradPhi for triangles (meshdata.cu:2420-2440) */
      const int id2 = nxq(1 + j, 0, 3), id3 = nxq(2 + j, 0, 3);
      const double p2x = m->x[nd[id2]], p2y = m->y[nd[id2]];
      const double p3x = m->x[nd[id3]], p3y = m->y[nd[id3]];
      m->gradphi[6 * e + 2 * j] = (p2y - p3y) / (2 * area);
      m->gradphi[6 * e + 2 * j + 1] = (p3x - p2x) / (2 * area);
    }
  }
  m->scx = xcalloc(ns, sizeof(double));
  m->scy = xcalloc(ns, sizeof(double));
  m->slen = xcalloc(ns, sizeof(double));
  m->snx = xcalloc(ns, sizeof(double));
  m->sny = xcalloc(ns, sizeof(double));
  for (int s = 0; s < ns; ++s) {
    /* populateSides (meshdata.cu:2274-2312), ics=1, gz=0 */
    const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
    m->scx[s] = 0.5 * (m->x[n1] + m->x[n2]);
    m->scy[s] = 0.5 * (m->y[n1] + m->y[n2]);
    const double vx = -(m->y[n1] - m->y[n2]), vy = (m->x[n1] - m->x[n2]);
    const double len = sqrt(vx * vx + vy * vy);
    m->snx[s] = vx / len;
    m->sny[s] = vy / len;
    const double dx = m->x[n2] - m->x[n1], dy = m->y[n2] - m->y[n1], dz = 0.0 - 0.0;
    m->slen[s] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  return 0;
}

or_mesh* or_mesh_new(int32_t n, int32_t ne, int32_t nvrt, const double* xy, const int32_t* elnode,
                     const double* depth, const int8_t* open_bnd, const int32_t* kb, const double* sigma) {
  if (n < 3 || ne < 1 || nvrt < 2 || !xy || !elnode || !depth || !sigma) {
    set_err("bad arguments");
    return NULL;
  }
  or_mesh* m = xcalloc(1, sizeof(or_mesh));
  m->n = n;
  m->ne = ne;
  m->nvrt = nvrt;
  m->x = xcalloc(n, sizeof(double));
  m->y = xcalloc(n, sizeof(double));
  for (int i = 0; i < n; ++i) m->x[i] = xy[2 * i], m->y[i] = xy[2 * i + 1];
  m->depth = xcalloc(n, sizeof(double));
  memcpy(m->depth, depth, sizeof(double) * n);
  m->sigma = xcalloc(nvrt, sizeof(double));
  memcpy(m->sigma, sigma, sizeof(double) * nvrt);
  m->open = xcalloc(n, 1);
  if (open_bnd)
    for (int i = 0; i < n; ++i) m->open[i] = open_bnd[i] ? 1 : 0;
  m->kb = xcalloc(n, sizeof(int32_t));
  if (kb) memcpy(m->kb, kb, sizeof(int32_t) * n);
  for (int i = 0; i < n; ++i)
    if (m->kb[i] < 0 || m->kb[i] > nvrt - 2) {
      set_err("kb must be in [0, nvrt-2] (node %d)", i);
      or_mesh_free(m);
      return NULL;
    }
  m->elnode = xcalloc(3 * (size_t)ne, sizeof(int32_t));
  memcpy(m->elnode, elnode, sizeof(int32_t) * 3 * (size_t)ne);
  if (build_topology(m) || build_geometry(m)) {
    or_mesh_free(m);
    return NULL;
  }
  const size_t ns = m->ns, nv = nvrt;
  m->eta2 = xcalloc(n, sizeof(double));
  m->eta1 = xcalloc(n, sizeof(double));
  m->uvw = xcalloc(3 * (size_t)n * nv, sizeof(double));
  m->su2 = xcalloc(2 * ns * nv, sizeof(double));
  m->ew = xcalloc((size_t)m->ne * nv, sizeof(double));
  m->btu = xcalloc(2 * ns * nv, sizeof(double));
  m->foot = xcalloc(ns * nv, sizeof(int32_t));
  m->bstatus = xcalloc(ns * nv, sizeof(int32_t));
  m->cd = xcalloc(n, sizeof(double));
  m->grav2 = xcalloc(n, sizeof(double));
  for (int i = 0; i < n; ++i) m->cd[i] = 2.5e-3, m->grav2[i] = 9.81;
  m->eth = xcalloc(n, sizeof(double));
  m->etype = xcalloc(n, sizeof(int8_t));
  for (int i = 0; i < n; ++i) m->etype[i] = -1;
  m->elbc = xcalloc(n, sizeof(double));
  m->vsource = xcalloc(ne, sizeof(double));
  m->sU = xcalloc(2 * ns, sizeof(double));
  m->hhat = xcalloc(ns, sizeof(double));
  m->chi = xcalloc(ns, sizeof(double));
  m->ghat1 = xcalloc(2 * (size_t)ne, sizeof(double));
  m->zc = xcalloc(ns * nv, sizeof(double));
  m->nz = xcalloc((size_t)n * nv, sizeof(double));
  m->node_dry = xcalloc(n, 1);
  m->elem_dry = xcalloc(ne, 1);
  m->side_dry = xcalloc(ns, 1);
  m->elem_kb = xcalloc(ne, sizeof(int32_t));
  m->side_kb = xcalloc(ns, sizeof(int32_t));
  or_params p = {150.0, 0.6, 0.01, 1.0, 10.0, 30.0, 1e-4, 1, 2};
  or_levels(m, &p);
  return m;
}

int32_t or_mesh_ns(const or_mesh* m) { return m->ns; }
int64_t or_mesh_nnz(const or_mesh* m) { return m->indptr[m->n]; }
int32_t or_mesh_nbr_count(const or_mesh* m) { return m->nnp_ptr[m->n]; }

int or_mesh_get(const or_mesh* m, int what, void* out) {
  const size_t n = m->n, ne = m->ne, ns = m->ns;
  switch (what) {
    case 0: {
      double* o = out;
      for (size_t i = 0; i < n; ++i) o[2 * i] = m->x[i], o[2 * i + 1] = m->y[i];
      return 0;
    }
    case 1: memcpy(out, m->elnode, 12 * ne); return 0;
    case 2: memcpy(out, m->elside, 12 * ne); return 0;
    case 3: memcpy(out, m->isdel, 8 * ns); return 0;
    case 4: memcpy(out, m->isidenode, 8 * ns); return 0;
    case 5: memcpy(out, m->nne_ptr, 4 * (n + 1)); return 0;
    case 6: memcpy(out, m->nne, 4 * (size_t)m->nne_ptr[n]); return 0;
    case 7: memcpy(out, m->nnp, 4 * (size_t)m->nnp_ptr[n]); return 0;
    case 8: memcpy(out, m->indptr, 4 * (n + 1)); return 0;
    case 9: memcpy(out, m->indices, 4 * (size_t)m->indptr[n]); return 0;
    case 10: memcpy(out, m->csr_slot, 2 * (size_t)m->nnp_ptr[n]); return 0;
    case 11: memcpy(out, m->depth, 8 * n); return 0;
    case 12: memcpy(out, m->open, n); return 0;
    case 13: memcpy(out, m->kb, 4 * n); return 0;
    case 14: memcpy(out, m->side_type, ns); return 0;
    default: set_err("unknown topology array %d", what); return -1;
  }
}

static int field_info(const or_mesh* m, int f, void** ptr, size_t* bytes) {
  const size_t n = m->n, ne = m->ne, ns = m->ns, nv = m->nvrt;
  switch (f) {
    case 0: *ptr = m->eta2; *bytes = 8 * n; return 0;
    case 1: *ptr = m->eta1; *bytes = 8 * n; return 0;
    case 2: *ptr = m->uvw; *bytes = 24 * n * nv; return 0;
    case 3: *ptr = m->su2; *bytes = 16 * ns * nv; return 0;
    case 4: *ptr = m->btu; *bytes = 16 * ns * nv; return 0;
    case 5: *ptr = m->foot; *bytes = 4 * ns * nv; return 0;
    case 6: *ptr = m->bstatus; *bytes = 4 * ns * nv; return 0;
    case 7: *ptr = m->cd; *bytes = 8 * n; return 0;
    case 8: *ptr = m->grav2; *bytes = 8 * n; return 0;
    case 9: *ptr = m->eth; *bytes = 8 * n; return 0;
    case 10: *ptr = m->vsource; *bytes = 8 * ne; return 0;
    case 11: *ptr = m->sU; *bytes = 16 * ns; return 0;
    case 12: *ptr = m->hhat; *bytes = 8 * ns; return 0;
    case 13: *ptr = m->chi; *bytes = 8 * ns; return 0;
    case 14: *ptr = m->ghat1; *bytes = 16 * ne; return 0;
    case 16: *ptr = m->zc; *bytes = 8 * ns * nv; return 0;
    case 17: *ptr = m->elbc; *bytes = 8 * n; return 0;
    case 19: *ptr = m->depth; *bytes = 8 * n; return 0;
    case 20: *ptr = m->ew; *bytes = 8 * ne * nv; return 0;
    case 21: *ptr = m->etype; *bytes = n; return 0;
    default: set_err("unknown field %d", f); return -1;
  }
}

int or_set_field(or_mesh* m, int field, const void* data) {
  void* p;
  size_t b;
  if (field_info(m, field, &p, &b)) return -1;
  memcpy(p, data, b);
  return 0;
}

int or_get_field(const or_mesh* m, int field, void* out) {
  if (field == 18) { /* wet flags: nodes, elements, sides */
    int8_t* o = out;
    for (int i = 0; i < m->n; ++i) o[i] = !m->node_dry[i];
    for (int e = 0; e < m->ne; ++e) o[m->n + e] = !m->elem_dry[e];
    for (int s = 0; s < m->ns; ++s) o[m->n + m->ne + s] = !m->side_dry[s];
    return 0;
  }
  void* p;
  size_t b;
  if (field_info(m, field, &p, &b)) return -1;
  memcpy(out, p, b);
  return 0;
}

/* ------------------------------------------------------------------ levels */
static int is_dry2(const or_mesh* m, int i, double h0) { return m->eta2[i] + m->depth[i] <= h0; }

/* Node::compZCoord (meshdata.cu:1744-1807) with sigmaTransitionDepth (:1647): ivcor=2 (SZ) and
 * ivcor=1 (localized sigma: one sigma per level, or one per node and level) */
static double comp_z(const or_mesh* m, int i, int k, double h0) {
  if (is_dry2(m, i, h0)) return 0.0;
  const double eta2 = m->eta2[i], depth = m->depth[i];
  if (m->vmode == 2) {
    const double hmod2 = fmin(depth, m->h_s);
    const double h_c = m->h_c;
    if (k == m->nvrt - 1) return eta2;
    else if (k == m->izlev) return -hmod2;
    else if (k == m->kb[i]) return -depth;
    else if (k < m->izlev) return m->ztot[k];
#if OR_PLANT == 10
    const int ksig = k - m->izlev - 1;
#else
    const int ksig = k - m->izlev;
#endif
    if (hmod2 <= h_c) return m->ssig[ksig] * (hmod2 + eta2) + eta2;
#if OR_PLANT == 11
    return eta2 * (1.0 + m->ssig[ksig]) + hmod2 * m->ssig[ksig] + (h_c - hmod2) * m->cs[ksig];
#else
    return eta2 * (1.0 + m->ssig[ksig]) + h_c * m->ssig[ksig] + (hmod2 - h_c) * m->cs[ksig];
#endif
  }
#if OR_PLANT == 12
  const double sg = m->vmode == 1 ? m->sig_node[(size_t)((i + 1) % m->n) * m->nvrt + k] : m->sigma[k];
#else
  const double sg = m->vmode == 1 ? m->sig_node[(size_t)i * m->nvrt + k] : m->sigma[k];
#endif
  if (k <= m->kb[i]) return -depth;
  else if (k == m->nvrt - 1) return eta2;
  else return (eta2 + depth) * sg + eta2;
}

/* Node::updateKb for ivcor=2 (meshdata.cu:1676-1718): the bisection over the Z levels;
 * returns -1 where the reference's assert fails (bed below the deepest Z level). */
static int sz_kb(const or_mesh* m, double depth) {
  if (depth <= m->h_s) return m->izlev;
  int kb = -1, L = 0, R = m->izlev + 1; /* R = nzlev */
  const double* ztot = m->ztot;
  if (!(-depth >= ztot[L] && -depth < (R <= m->izlev ? ztot[R] : 0.0))) return -1;
  while (L <= R) {
    const int M = (R + L) / 2;
    const double zM = M <= m->izlev ? ztot[M] : 0.0; /* readVGrid leaves ztot[nz] = 0 */
    if (-depth >= ztot[L] && -depth < zM)
      R = M;
    else
      L = M;
    if (R - L == 1) {
      kb = L;
      break;
    }
  }
  return kb;
}

static int update_sz_kb(or_mesh* m) {
  for (int i = 0; i < m->n; ++i) {
    const int kb = sz_kb(m, m->depth[i]);
    if (kb < 0) {
      set_err("ZCOOR: invalid depth %g at node %d (below the deepest Z level)", m->depth[i], i);
      return -1;
    }
    m->kb[i] = kb;
  }
  return 0;
}

int or_set_vertical(or_mesh* m, int32_t ivcor, const double* sigma_node, int32_t nz, double h_s, double h_c,
                    double theta_b, double theta_f, const double* ztot, const double* sigma_s) {
  const int nv = m->nvrt;
  if (ivcor == 1) {
    if (!sigma_node) {
      set_err("ivcor 1 needs sigma_node");
      return -1;
    }
    free(m->sig_node);
    m->sig_node = xcalloc((size_t)m->n * nv, sizeof(double));
    memcpy(m->sig_node, sigma_node, sizeof(double) * (size_t)m->n * nv);
    m->vmode = 1;
    return 0;
  }
  if (ivcor != 2 || nz < 1 || nz > nv - 1 || !ztot || !sigma_s) {
    set_err("ivcor 2 needs 1 <= nz <= nvrt-1, ztot and sigma_s");
    return -1;
  }
  const int nsig = nv - nz + 1;
  free(m->ztot);
  free(m->ssig);
  free(m->cs);
  m->ztot = xcalloc(nz, sizeof(double));
  memcpy(m->ztot, ztot, sizeof(double) * nz);
  m->ssig = xcalloc(nsig, sizeof(double));
  memcpy(m->ssig, sigma_s, sizeof(double) * nsig);
  /* stretching function C(sigma) of readVGrid (read_and_process_grid.cpp:1497-1502) */
  m->cs = xcalloc(nsig, sizeof(double));
  for (int k = 0; k < nsig; ++k)
    m->cs[k] = (1.0 - theta_b) * sinh(theta_f * m->ssig[k]) / sinh(theta_f) +
               theta_b * (tanh(theta_f * (m->ssig[k] + 0.5)) - tanh(theta_f * 0.5)) / 2.0 / tanh(theta_f * 0.5);
  m->h_s = h_s;
  m->h_c = h_c;
  m->izlev = nz - 1;
  m->vmode = 2;
  return update_sz_kb(m);
}

void or_levels(or_mesh* m, const or_params* p) {
  const int n = m->n, ne = m->ne, ns = m->ns, nv = m->nvrt;
  /* ul0UpdateZCoords (meshdata.cu:5521-5548): Node::updateKb (ivcor=2 only), then the levels */
  if (m->vmode == 2) update_sz_kb(m);
  for (int i = 0; i < n; ++i)
    for (int k = m->kb[i]; k < nv; ++k) m->nz[(size_t)i * nv + k] = comp_z(m, i, k, p->h0);
  /* ul0ComputeDryElements (meshdata.cu:5550-5580) */
  for (int e = 0; e < ne; ++e) {
    const int* nd = m->elnode + 3 * e;
    m->elem_dry[e] = is_dry2(m, nd[0], p->h0) || is_dry2(m, nd[1], p->h0) || is_dry2(m, nd[2], p->h0);
    m->elem_kb[e] = 0;
    if (m->elem_dry[e]) continue;
    const int kb1 = m->kb[nd[0]], kb2 = m->kb[nd[1]], kb3 = m->kb[nd[2]];
    const int a = kb1 < kb2 ? kb1 : kb2;
    m->elem_kb[e] = a < kb3 ? a : kb3;
  }
  /* ul0ComputeDryNodes (meshdata.cu:5596-5612) */
  for (int i = 0; i < n; ++i) {
    m->node_dry[i] = 1;
    for (int j = m->nne_ptr[i]; j < m->nne_ptr[i + 1]; ++j)
      if (!m->elem_dry[m->nne[j]]) {
        m->node_dry[i] = 0;
        break;
      }
  }
  /* ul0UpdateSides + ul0UpdateSides2 (meshdata.cu:5679-5760) */
  for (int s = 0; s < ns; ++s) {
    const int e0 = m->isdel[2 * s], e1 = m->isdel[2 * s + 1];
    const int wet = !m->elem_dry[e0] || (e1 != -1 && !m->elem_dry[e1]);
    const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
    const int kb1 = m->kb[n1], kb2 = m->kb[n2];
    m->side_kb[s] = 0;
    if (wet) {
      m->side_kb[s] = kb1 < kb2 ? kb1 : kb2;
      for (int k = 0; k < m->side_kb[s]; ++k) m->zc[(size_t)s * nv + k] = 0.0;
      for (int k = m->side_kb[s]; k < nv; ++k) {
        const int k1 = k > kb1 ? k : kb1, k2 = k > kb2 ? k : kb2;
        m->zc[(size_t)s * nv + k] = 0.5 * (m->nz[(size_t)n1 * nv + k1] + m->nz[(size_t)n2 * nv + k2]);
      }
    }
    m->side_dry[s] = !wet;
  }
}

/* ------------------------------------------------------------------ nodal velocities */
/* clampVelocity (meshdata.cu:192-194) with clamp = fmin(fmax(x, lo), hi) (meshdata.h:155-158) */
static double clamp_v(double x, double hi) { return fmin(fmax(x, -hi), hi); }

/* zeroNodalVelocities, computeNodalVerticalVelocities1/2 and computeNodalHorizontalVelocities1/2
 * (meshdata.cu:5765-5929) for triangles, ics=1, indvel <= 0. The reference adds each element's
 * contribution to its nodes with atomicAdd; here node i adds its ball elements in ball order:
 *   w2[k]  = (1 / sum_e area_e) * sum_e w_e[k] area_e            (k >= kb_i; all ball elements)
 *   u2[k]  = (sum_e clamp(u_s2 + u_s3 - u_s1)[k]) * (1 / count)  (k >= kb_i; wet elements, s1 the
 *            side opposite node i, s2 and s3 the next sides of the element; count = their number)
 * and zero below kb_i or when no wet element touches the node. */
void or_nodal_velocity(or_mesh* m, const or_params* p, int nthreads) {
  const int n = m->n, nv = m->nvrt;
  const double rx = p->rmaxvel, ry = p->rmaxvel * 1.013;
  (void)nthreads;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int i = 0; i < n; ++i) {
    const int kbn = m->kb[i];
    double asum = 0.0;
    int count = 0;
    for (int j = m->nne_ptr[i]; j < m->nne_ptr[i + 1]; ++j) {
      const int e = m->nne[j];
      asum += m->area[e];
      if (!m->elem_dry[e] && !m->node_dry[i]) ++count;
    }
    const double inv_area = 1.0 / asum;
    for (int k = 0; k < nv; ++k) {
      double w = 0.0, ux = 0.0, uy = 0.0;
      if (k >= kbn) {
        double ws = 0.0, sx = 0.0, sy = 0.0;
        for (int j = m->nne_ptr[i]; j < m->nne_ptr[i + 1]; ++j) {
          const int e = m->nne[j];
          ws += m->ew[(size_t)e * nv + k] * m->area[e];
          if (m->elem_dry[e] || m->node_dry[i]) continue;
          const int l = local_index(m, e, i);
          const size_t s1 = (size_t)m->elside[3 * e + l], s2 = (size_t)m->elside[3 * e + nxq(1, l, 3)],
                       s3 = (size_t)m->elside[3 * e + nxq(2, l, 3)];
          const double* u1 = m->su2 + 2 * (s1 * nv + k);
          const double* u2 = m->su2 + 2 * (s2 * nv + k);
          const double* u3 = m->su2 + 2 * (s3 * nv + k);
          sx += clamp_v(u2[0] + u3[0] - u1[0], rx);
          sy += clamp_v(u2[1] + u3[1] - u1[1], ry);
        }
        w = inv_area * ws;
        if (count > 0) {
          ux = sx * (1.0 / count);
          uy = sy * (1.0 / count);
        }
      }
      double* o = m->uvw + 3 * ((size_t)i * nv + k);
      o[0] = ux;
      o[1] = uy;
      o[2] = w;
    }
  }
}

/* ------------------------------------------------------------------ momentum */
/* Element::grad2 (meshdata.h:1105-1114), triangles: sum_l f_l gradphi_l */
static void elem_grad(const or_mesh* m, int e, const double* f, double g[2]) {
  const int32_t* nd = m->elnode + 3 * e;
  const double* gp = m->gradphi + 6 * e;
  g[0] = f[nd[0]] * gp[0] + f[nd[1]] * gp[2] + f[nd[2]] * gp[4];
  g[1] = f[nd[0]] * gp[1] + f[nd[1]] * gp[3] + f[nd[2]] * gp[5];
}

/* Element::isDry2 (meshdata.cu:774-778) from Node::isDry2 (meshdata.h:1757): eta2 + depth <= h0 */
static int elem_dry2(const or_mesh* m, int e, double h0) {
  const int32_t* nd = m->elnode + 3 * e;
  for (int l = 0; l < 3; ++l)
    if (m->eta2[nd[l]] + m->depth[nd[l]] <= h0) return 1;
  return 0;
}

/* computeHorGradientAtSide (meshdata.cu:5045-5071) with Side::grad2 (meshdata.cu:1363-1398):
 * gradEta1 over the side's elements that are wet (isDry of the last level update), gradEta2 over
 * the elements wet by eta2 (isDry2), each the mean of the element gradients; gradETP and gradPr
 * are zero in the supported subset (no tidal potential, no atmospheric pressure). */
static void side_grads(const or_mesh* m, int s, double h0, double g1[2], double g2[2]) {
  g1[0] = g1[1] = g2[0] = g2[1] = 0.0;
  int c1 = 0, c2 = 0;
  for (int j = 0; j < 2; ++j) {
    const int e = m->isdel[2 * s + j];
    if (e < 0) continue;
    double g[2];
    if (!m->elem_dry[e]) {
      elem_grad(m, e, m->eta1, g);
      g1[0] += g[0];
      g1[1] += g[1];
      ++c1;
    }
    if (!elem_dry2(m, e, h0)) {
      elem_grad(m, e, m->eta2, g);
      g2[0] += g[0];
      g2[1] += g[1];
      ++c2;
    }
  }
  if (c1 > 0) {
    g1[0] = g1[0] * (1.0 / c1);
    g1[1] = g1[1] * (1.0 / c1);
  }
  if (c2 > 0) {
    g2[0] = g2[0] * (1.0 / c2);
    g2[1] = g2[1] * (1.0 / c2);
  }
}

/* computeHorGradientAtSide + solveMomentumEquations (meshdata.cu:5045-5352) + applyMomentumBC
 * (meshdata.cu:5356-5446) for the barotropic (ibc=1), itur=0 (dfv = dfv0 at every node and
 * level), iveg=0, ics=1, imm=0, islip=0 subset with no Coriolis, wind stress, horizontal
 * viscosity, atmospheric pressure or tidal potential (terms that are zero there are omitted;
 * they could only change the sign of a zero). Per side:
 *   dry: u2 = 0.
 *   2D (kb = nvrt-2): u2[k] = clamp((btu[nvrt-1] - dt grav3 ((1-th) gradEta1 + th gradEta2)) Hhat / htot1)
 *       for every k, htot1 = zc[nvrt-1] - zc[kb].
 *   3D: the Galerkin tridiagonal system over k = kb+1 .. nvrt-1 (mass dz/6, dz/3; viscosity
 *       dt dfv0 / dz; bottom drag dt chi2 on the first row; right-hand side -(dz/2) dt grad2D +
 *       the mass-weighted btu) solved by tridag_sch (meshdata.cu:3665-3711), u2[k] = clamp(soln);
 *       u2[kb] = u2[kb+1] when Cd = 0, else 0. (The reference's loop also assigns k = kb from
 *       soln[-1], which the bottom condition then overwrites; it is not restated.)
 *   k < kb: 0. Land boundary sides then keep only the tangential component for k >= kb. */
/* shapiroFilter1 + shapiroFilter2 (meshdata.cu:5456-5518), niter_shap = 1: for each wet internal
 * side, bcc[k] = u2[k] + shapiro0 (0.25 sum_j u2_j[kin_j] - u2[k]) for k > kb, over its four
 * neighbours Side::neigh(j) (meshdata.h:1409-1422: element j & 1, the side before (j < 2) or after
 * it there), kin_j = max(k, kb_j + 1) for a wet neighbour; then u2 <- bcc above kb (2D: u2[kb+1]
 * and u2[kb] both take bcc[kb+1]) and u2[k] = 0 for k < kb - 1. */
static void shapiro_filter(or_mesh* m, double shapiro0) {
  const int ns = m->ns, nv = m->nvrt;
  double* bcc = xcalloc(2 * (size_t)ns * nv, sizeof(double));
  for (int s = 0; s < ns; ++s) {
    const int e0 = m->isdel[2 * s], e1 = m->isdel[2 * s + 1];
    if (m->side_dry[s] || e1 < 0) continue;
    const int kb = m->side_kb[s];
    const int es[2] = {e0, e1};
    int nb[4];
    for (int j = 0; j < 4; ++j) {
      const int e = es[j & 1], off = 2 * (j >> 1) - 1;
      int loc = 0;
      while (m->elside[3 * e + loc] != s) ++loc;
      nb[j] = m->elside[3 * e + (loc + off + 48) % 3];
    }
    const double* u2 = m->su2 + 2 * (size_t)s * nv;
    for (int k = kb + 1; k < nv; ++k) {
      double sx = 0.0, sy = 0.0;
      for (int j = 0; j < 4; ++j) {
        const int n = nb[j];
        const int kin = m->side_dry[n] ? k : (k > m->side_kb[n] + 1 ? k : m->side_kb[n] + 1);
        sx += m->su2[2 * ((size_t)n * nv + kin)];
        sy += m->su2[2 * ((size_t)n * nv + kin) + 1];
      }
      bcc[2 * ((size_t)s * nv + k)] = u2[2 * k] + shapiro0 * (0.25 * sx - u2[2 * k]);
      bcc[2 * ((size_t)s * nv + k) + 1] = u2[2 * k + 1] + shapiro0 * (0.25 * sy - u2[2 * k + 1]);
    }
  }
  for (int s = 0; s < ns; ++s) {
    if (m->side_dry[s] || m->isdel[2 * s + 1] < 0) continue;
    const int kb = m->side_kb[s];
    double* u2 = m->su2 + 2 * (size_t)s * nv;
    const double* f = bcc + 2 * (size_t)s * nv;
    for (int k = 0; k < kb - 1; ++k) u2[2 * k] = u2[2 * k + 1] = 0.0;
    if (kb == nv - 2) {
      u2[2 * (kb + 1)] = f[2 * (kb + 1)];
      u2[2 * (kb + 1) + 1] = f[2 * (kb + 1) + 1];
      u2[2 * kb] = u2[2 * (nv - 1)];
      u2[2 * kb + 1] = u2[2 * (nv - 1) + 1];
    } else {
      for (int k = kb + 1; k < nv; ++k) {
        u2[2 * k] = f[2 * k];
        u2[2 * k + 1] = f[2 * k + 1];
      }
    }
  }
  free(bcc);
}

void or_momentum(or_mesh* m, const or_params* p) {
  const int ns = m->ns, nv = m->nvrt;
  const double dt = p->dt, th = p->theta, rx = p->rmaxvel, ry = p->rmaxvel * 1.013, dfv0 = p->dfv0;
  double* a = xcalloc(nv, sizeof(double));
  double* b = xcalloc(nv, sizeof(double));
  double* c = xcalloc(nv, sizeof(double));
  double* r = xcalloc(2 * (size_t)nv, sizeof(double));
  double* u = xcalloc(2 * (size_t)nv, sizeof(double));
  double* gam = xcalloc(nv, sizeof(double));
  for (int s = 0; s < ns; ++s) {
    double* u2 = m->su2 + 2 * (size_t)s * nv;
    const double* bt = m->btu + 2 * (size_t)s * nv;
    const double* zc = m->zc + (size_t)s * nv;
    const int kb = m->side_kb[s];
    const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
    if (m->side_dry[s]) {
      for (int k = 0; k < 2 * nv; ++k) u2[k] = 0.0;
    } else {
      double g1[2], g2[2];
      side_grads(m, s, p->h0, g1, g2);
      const double grav3 = 0.5 * (m->grav2[n1] + m->grav2[n2]);
      const double Cd = 0.5 * (m->cd[n1] + m->cd[n2]);
      if (kb == nv - 2) {
        const double htot = zc[nv - 1] - zc[kb];
        const double hhat = m->hhat[s];
        double hg[2];
        for (int l = 0; l < 2; ++l)
          hg[l] = bt[2 * (nv - 1) + l] + dt * (0.0 - grav3 * ((1.0 - th) * g1[l] + th * g2[l]));
        const double vx = clamp_v(hg[0] * hhat / htot, rx), vy = clamp_v(hg[1] * hhat / htot, ry);
        for (int k = 0; k < nv; ++k) {
          u2[2 * k] = vx;
          u2[2 * k + 1] = vy;
        }
      } else {
        const int nr = nv - kb - 1;
        for (int k = kb + 1; k < nv; ++k) {
          const int kin = k - (kb + 1);
          a[kin] = 0.0;
          c[kin] = 0.0;
          b[kin] = 0.0;
          const double cff1 = 1.0;
          if (k < nv - 1) {
            const double dzz = zc[k + 1] - zc[k];
            const double tmp = dt * dfv0 / dzz;
            c[kin] += cff1 * dzz / 6.0 - tmp;
            b[kin] += cff1 * dzz / 3.0 + tmp;
          }
          if (k > kb + 1) {
            const double dzz = zc[k] - zc[k - 1];
            const double tmp = dt * dfv0 / dzz;
            a[kin] += cff1 * dzz / 6.0 - tmp;
            b[kin] += cff1 * dzz / 3.0 + tmp;
          } else {
            b[kin] += dt * m->chi[s];
          }
        }
        double G[2];
        for (int l = 0; l < 2; ++l) G[l] = grav3 * (th * g2[l] + (1.0 - th) * g1[l]);
        for (int k = kb + 1; k < nv; ++k) {
          const int kin = k - kb - 1;
          const double dz_0 = zc[k] - zc[k - 1];
          const double dz_1 = (k < nv - 1) ? zc[k + 1] - zc[k] : 0.0;
          double dz = 0.0;
          dz += dz_1;
          if (k > kb + 1) dz += dz_0;
          for (int l = 0; l < 2; ++l) {
            double v = -(0.5 * dz) * dt * G[l];
            if (k < nv - 1) v += (dz_1 / 6.0) * (2.0 * bt[2 * k + l] + bt[2 * (k + 1) + l]);
            if (k > kb + 1) v += (dz_0 / 6.0) * (2.0 * bt[2 * k + l] + bt[2 * (k - 1) + l]);
            r[2 * kin + l] = v;
          }
        }
        /* tridag_sch<2> */
        double bet = b[0];
        u[0] = r[0] / bet;
        u[1] = r[1] / bet;
        for (int j = 1; j < nr; ++j) {
          gam[j] = c[j - 1] / bet;
          bet = b[j] - a[j] * gam[j];
          for (int l = 0; l < 2; ++l) u[2 * j + l] = (r[2 * j + l] - a[j] * u[2 * (j - 1) + l]) / bet;
        }
        for (int j = nr - 2; j >= 0; --j)
          for (int l = 0; l < 2; ++l) u[2 * j + l] = u[2 * j + l] - gam[j + 1] * u[2 * (j + 1) + l];
        for (int k = kb + 1; k < nv; ++k) {
          const int kin = k - kb - 1;
          u2[2 * k] = clamp_v(u[2 * kin], rx);
          u2[2 * k + 1] = clamp_v(u[2 * kin + 1], ry);
        }
      }
      /* bottom condition, imm = 0 */
      if (Cd == 0.0) {
        u2[2 * kb] = u2[2 * (kb + 1)];
        u2[2 * kb + 1] = u2[2 * (kb + 1) + 1];
      } else if (kb < nv - 2) {
        u2[2 * kb] = 0.0;
        u2[2 * kb + 1] = 0.0;
      }
      for (int k = 0; k < kb; ++k) u2[2 * k] = u2[2 * k + 1] = 0.0;
    }
    /* applyMomentumBC: land sides, free slip (islip = 0): u2 = (u2 . t) t, t = (-n.y, n.x) */
    if (m->side_type[s] == 1) {
      const double tx = -m->sny[s], ty = m->snx[s];
      for (int k = kb; k < nv; ++k) {
        const double vtan = u2[2 * k] * tx + u2[2 * k + 1] * ty;
        u2[2 * k] = vtan * tx;
        u2[2 * k + 1] = vtan * ty;
      }
    }
  }
  free(a);
  free(b);
  free(c);
  free(r);
  free(u);
  free(gam);
  if (p->ishapiro == 1) shapiro_filter(m, p->shapiro0);
}

/* ------------------------------------------------------------------ vertical velocity */
/* Element::n (meshdata.cu:559-631), triangles, ics=1: the unit normal of the element's face at
 * level k (node levels max(k, kb) of its nodes) and that face's area */
static void elem_face_normal(const or_mesh* m, int e, int k, double n[3], double* area3d) {
  const int nv = m->nvrt;
  const int32_t* nd = m->elnode + 3 * e;
  double px[3], py[3], pz[3];
  for (int l = 0; l < 3; ++l) {
    const int kk = k > m->kb[nd[l]] ? k : m->kb[nd[l]];
    px[l] = m->x[nd[l]];
    py[l] = m->y[nd[l]];
    pz[l] = m->nz[(size_t)nd[l] * nv + kk];
  }
  const double n20[3] = {px[2] - px[0], py[2] - py[0], pz[2] - pz[0]};
  const double n10[3] = {px[1] - px[0], py[1] - py[0], pz[1] - pz[0]};
  const double nx = (n10[1]) * (n20[2]) - (n20[1]) * (n10[2]);
  const double ny = (n20[0]) * (n10[2]) - (n10[0]) * (n20[2]);
  const double nzz = 2.0 * m->area[e];
  const double len = sqrt(nx * nx + ny * ny + nzz * nzz);
  *area3d = 0.5 * len;
  const double inv_len = 1.0 / len;
  n[0] = nx * inv_len;
  n[1] = ny * inv_len;
  n[2] = nzz * inv_len;
}

/* computeVerticalVelocity (meshdata.cu:5978-6102), triangles, ics=1, imm=0: w of each element
 * from the flux balance of its prisms, bottom up. w = 0 for dry elements and below kb;
 * w[kb] = -grad(depth) . mean_j u2_j[kb_j]; for l = kb .. nvrt-2, with the side fluxes
 * F_l = sum_j ssign_j (0.5 (u2_j[max(l+1, kb_j)] + u2_j[max(l, kb_j)]) . n_j) len_j dz_j(l),
 * top = -F_l - bottom, w[l+1] = (top - (mean_j u2_j[max(l+1, kb_j)] . n_xy) A) / (n_z A)
 * (n, A the unit normal and area of the face at level l+1) and bottom = -top. */
void or_vertical_velocity(or_mesh* m) {
  const int ne = m->ne, nv = m->nvrt;
  for (int e = 0; e < ne; ++e) {
    double* w = m->ew + (size_t)e * nv;
    for (int l = 0; l < nv; ++l) w[l] = 0.0;
    if (m->elem_dry[e]) continue;
    const int kb = m->elem_kb[e];
    const int32_t* es = m->elside + 3 * e;
    double gh[2];
    elem_grad(m, e, m->depth, gh);
    double ub[2] = {0.0, 0.0};
    {
      double v[3][2];
      for (int j = 0; j < 3; ++j) {
        const int s = es[j];
        const double* u = m->su2 + 2 * ((size_t)s * nv + m->side_kb[s]);
        v[j][0] = u[0];
        v[j][1] = u[1];
      }
      for (int c = 0; c < 2; ++c) ub[c] = (v[0][c] + v[1][c] + v[2][c] + 0.0) / 3;
    }
    const double bottomW = -(gh[0] * ub[0] + gh[1] * ub[1]);
    w[kb] = 0.0 + bottomW;
    double bottomFlux = 0.0;
    for (int l = kb; l < nv - 1; ++l) {
      double tn[3], ta;
      elem_face_normal(m, e, l + 1, tn, &ta);
      double sideFlux = 0.0;
      double ut[2] = {0.0, 0.0};
      double vt[3][2];
      for (int j = 0; j < 3; ++j) {
        const int s = es[j];
        const int kbs = m->side_kb[s];
        const int kt = l + 1 > kbs ? l + 1 : kbs, kl = l > kbs ? l : kbs;
        const double* u2 = m->su2 + 2 * (size_t)s * nv;
        const double* zc = m->zc + (size_t)s * nv;
        const double ucx = 0.5 * (u2[2 * kt] + u2[2 * kl]), ucy = 0.5 * (u2[2 * kt + 1] + u2[2 * kl + 1]);
        const double sideArea = m->slen[s] * (zc[kt] - zc[kl]);
        const double sign = m->ssign[3 * e + j];
        sideFlux += sign * (ucx * m->snx[s] + ucy * m->sny[s]) * sideArea;
        vt[j][0] = u2[2 * kt];
        vt[j][1] = u2[2 * kt + 1];
      }
      for (int c = 0; c < 2; ++c) ut[c] = (vt[0][c] + vt[1][c] + vt[2][c] + 0.0) / 3;
      const double topFlux = -sideFlux - bottomFlux;
      const double topHorizontalFlux = (ut[0] * tn[0] + ut[1] * tn[1]) * ta;
      const double topVerticalFlux = topFlux - topHorizontalFlux;
      w[l + 1] = topVerticalFlux / (tn[2] * ta);
      bottomFlux = -topFlux;
    }
  }
}

/* ------------------------------------------------------------------ backtracking */
typedef struct {
  double x, y;
} P2;

/* Element::getNodeInOrder (meshdata.cu:646-649) */
static int node_in_order(const or_mesh* m, int e, int idx) {
  const int s = m->elside[3 * e + idx];
  return (m->ssign[3 * e + idx] > 0) ? m->isidenode[2 * s] : m->isidenode[2 * s + 1];
}

/* Element::getLocalNodePosition (meshdata.cu:635-644), ics=1 */
static P2 local_node_position(const or_mesh* m, int e, int idx) {
  const int idx2 = (idx == 3) ? 2 : idx; /* triangles */
  const int nd = node_in_order(m, e, idx2);
  P2 r = {m->x[nd], m->y[nd]};
  return r;
}

static double dot2(P2 a, P2 b) { return a.x * b.x + a.y * b.y; }
static double cross2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
static P2 sub2(P2 a, P2 b) {
  P2 r = {a.x - b.x, a.y - b.y};
  return r;
}
static P2 add2(P2 a, P2 b) {
  P2 r = {a.x + b.x, a.y + b.y};
  return r;
}
static P2 scale2(P2 v, double s) {
  P2 r = {v.x * s, v.y * s};
  return r;
}
static double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

/* Element::computePredicates (meshdata.cu:651-746) for a triangle; returns the valid mask. */
static unsigned predicates(const or_mesh* m, int e, P2 x0, P2 xt, double* fac) {
  const P2 C[4] = {local_node_position(m, e, 0), local_node_position(m, e, 1), local_node_position(m, e, 2),
                   local_node_position(m, e, 3)};
  const P2 vt0 = sub2(xt, x0);
  double maxProjDist = fmax(dot2(vt0, sub2(C[0], x0)), dot2(vt0, sub2(C[1], x0)));
  maxProjDist = fmax(maxProjDist, dot2(vt0, sub2(C[2], x0)));
  /* triangles: no fourth corner */
  *fac = fmax(1.00001 * *fac, fmin(1.0, (maxProjDist / dot2(vt0, vt0)) * 1.00001));
  const P2 ixt = add2(x0, scale2(sub2(xt, x0), *fac));
  const int i3 = nxq(3, 0, 3);
  const P2 s0 = sub2(C[1], C[0]);
  const P2 s1 = sub2(C[2], C[1]);
  const P2 s2 = sub2(C[i3], C[2]);
  const P2 s3 = sub2(C[0], C[i3]);
  unsigned isLeft = 0, isObtuse = 0, isBisector = 0;
  isLeft |= (cross2(s0, sub2(ixt, C[0])) > -small) << 0;
  isLeft |= (cross2(s1, sub2(ixt, C[1])) > -small) << 1;
  isLeft |= (cross2(s2, sub2(ixt, C[2])) > -small) << 2;
  isLeft |= ((cross2(s3, sub2(ixt, C[i3])) > -small) || 1) << 3;
  const P2 s3_tri = s0;
  isObtuse |= (dot2(s0, s1) > 0.0) << 0;
  isObtuse |= (dot2(s1, s2) > 0.0) << 1;
  isObtuse |= (dot2(s2, s3_tri) > 0.0) << 2;
  isObtuse |= (dot2(s3, s0) > 0.0) << 3;
  isBisector |= (dot2(sub2(ixt, C[1]), add2(s0, s1)) < 0.0) << 0;
  isBisector |= (dot2(sub2(ixt, C[2]), add2(s1, s2)) < 0.0) << 1;
  isBisector |= (dot2(sub2(ixt, C[i3]), add2(s2, s3_tri)) < 0.0) << 2;
  isBisector |= (dot2(sub2(ixt, C[0]), add2(s3, s0)) < 0.0) << 3;
  const unsigned notLeft = (~isLeft) & 0x0F, notObtuse = (~isObtuse) & 0x0F;
  return notLeft & (notObtuse | isBisector);
}

static int other_elem(const or_mesh* m, int s, int e) {
  return (m->isdel[2 * s] == e) ? m->isdel[2 * s + 1] : m->isdel[2 * s];
}

/* celestialWalk (meshdata.cu:85-127) */
static int celestial_walk(const or_mesh* m, int* cur, P2 x0, P2 xt, int* lastCross, int* status) {
  int iter = 0;
  double fac = 0.00001;
  while (iter < 40) {
    iter++;
    const unsigned valid = predicates(m, *cur, x0, xt, &fac);
    if (valid == 0) return 1;
    int first = 0;
    while (!((valid >> first) & 1u)) ++first;
    *lastCross = m->elside[3 * *cur + first];
    const int prev = *cur;
    *cur = other_elem(m, *lastCross, *cur);
    if (*cur == -1 || m->elem_dry[*cur]) {
      *cur = prev;
      return 0;
    }
  }
  *status |= BT_WALK_CAP;
  return 1;
}

/* Element::computeShapeCoords with fiAnywhere, triangles (meshdata.cu:990-1043) */
static void shape_coords(const or_mesh* m, int e, P2 xt, double shp[4]) {
  const int* nd = m->elnode + 3 * e;
  const P2 x0 = {m->x[nd[0]], m->y[nd[0]]}, x1 = {m->x[nd[1]], m->y[nd[1]]}, x2 = {m->x[nd[2]], m->y[nd[2]]};
  const double area = m->area[e];
  shp[0] = signed_area(xt.x, xt.y, x1.x, x1.y, x2.x, x2.y) / area;
  shp[1] = signed_area(x0.x, x0.y, xt.x, xt.y, x2.x, x2.y) / area;
  shp[2] = 1.0 - shp[0] - shp[1];
  shp[3] = 0.0;
  const double tmpmin = fmin(shp[0], fmin(shp[1], shp[2]));
  if (tmpmin < 0) {
    const double a = -tmpmin + 1.0e-8;
    const double b = 1.0 / (shp[0] + shp[1] + shp[2] + 3.0 * a);
    shp[0] = (shp[0] + a) * b;
    shp[1] = (shp[1] + a) * b;
    shp[2] = (shp[2] + a) * b;
  }
}

/* Element::interpolate (meshdata.h:1139-1144) */
static double interp4(const double shp[4], const double v[4]) {
  return shp[0] * v[0] + shp[1] * v[1] + shp[2] * v[2] + shp[3] * v[3];
}

static void nodal_z(const or_mesh* m, int e, int level, double h0, double out[4]) {
  const int* nd = m->elnode + 3 * e;
  for (int j = 0; j < 3; ++j) out[j] = comp_z(m, nd[j], level > m->kb[nd[j]] ? level : m->kb[nd[j]], h0);
  out[3] = 0.0;
}

/* Element::interpolateVelocities (meshdata.cu:1209-1291) */
static void interpolate_velocities(const or_mesh* m, int e, double px, double py, double pz, double h0,
                                   double out[3]) {
  double shp[4];
  P2 xt = {px, py};
  shape_coords(m, e, xt, shp);
  int L = m->elem_kb[e] - 1;
  int R = m->nvrt - 1;
  double zl[4], zr[4], zm[4];
  nodal_z(m, e, L, h0, zl);
  nodal_z(m, e, R, h0, zr);
  double zL = interp4(shp, zl), zR = interp4(shp, zr);
  double w1, w2;
  if (pz <= zL) {
    w1 = 1.0;
    w2 = 0.0;
  } else if (pz >= zR) {
    w1 = 0.0;
    w2 = 1.0;
  } else {
    while (R - L > 1) {
      const int M = (int)((L + R) / 2.0);
      nodal_z(m, e, M, h0, zm);
      const double zM = interp4(shp, zm);
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
  }
  const int* nd = m->elnode + 3 * e;
  const int nv = m->nvrt;
  double uL[4], vL[4], wL[4], uR[4], vR[4], wR[4];
  for (int j = 0; j < 3; ++j) {
    const int kL = L > m->kb[nd[j]] ? L : m->kb[nd[j]];
    const int kR = R > m->kb[nd[j]] ? R : m->kb[nd[j]];
    const double* a = m->uvw + 3 * ((size_t)nd[j] * nv + kL);
    const double* b = m->uvw + 3 * ((size_t)nd[j] * nv + kR);
    uL[j] = a[0], vL[j] = a[1], wL[j] = a[2];
    uR[j] = b[0], vR[j] = b[1], wR[j] = b[2];
  }
  uL[3] = vL[3] = wL[3] = uR[3] = vR[3] = wR[3] = 0.0;
  out[0] = w1 * interp4(shp, uL) + w2 * interp4(shp, uR);
  out[1] = w1 * interp4(shp, vL) + w2 * interp4(shp, vR);
  out[2] = w1 * interp4(shp, wL) + w2 * interp4(shp, wR);
}

/* Side::btrack (meshdata.cu:1449-1629) */
static void side_btrack(const or_mesh* m, const or_params* p, int s, int jlev, double dtbk, double u_int[3],
                        int* foot, int* status) {
  const int nv = m->nvrt;
  const double h0 = p->h0, dt = p->dt;
  const double velmin2 = p->velmin_btrack * p->velmin_btrack;
  double x0[3] = {m->scx[s], m->scy[s], m->zc[(size_t)s * nv + jlev]};
  double time_rm = dt, time_rm_1 = 1e10, time_rm_2 = 2e10, time_rm_3 = 3e10;
  const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
  u_int[0] = m->su2[2 * ((size_t)s * nv + jlev)];
  u_int[1] = m->su2[2 * ((size_t)s * nv + jlev) + 1];
  u_int[2] = 0.5 * (m->uvw[3 * ((size_t)n1 * nv + jlev) + 2] + m->uvw[3 * ((size_t)n2 * nv + jlev) + 2]);
  const int e0 = m->isdel[2 * s], e1 = m->isdel[2 * s + 1];
  int cur = (!m->elem_dry[e0]) ? e0 : e1;
  int iter = 0;
  double xt[3] = {-1.0e25, -1.0e25, -1.0e25};
  for (;;) {
    const double dtb = fmin(dtbk, time_rm);
    xt[0] = x0[0] - dtb * u_int[0];
    xt[1] = x0[1] - dtb * u_int[1];
    xt[2] = x0[2] - dtb * u_int[2];
    double t, sp;
    int lastCross = -1;
    const P2 x02 = {x0[0], x0[1]}, xt2 = {xt[0], xt[1]};
    const int found = celestial_walk(m, &cur, x02, xt2, &lastCross, status);
    if (time_rm_3 <= time_rm) {
      xt[0] = x0[0], xt[1] = x0[1], xt[2] = x0[2];
      *status |= BT_STUCK;
      break;
    }
    if (!found) {
      const int sn1 = m->isidenode[2 * lastCross], sn2 = m->isidenode[2 * lastCross + 1];
      const P2 p1 = {m->x[sn1], m->y[sn1]}, p2 = {m->x[sn2], m->y[sn2]};
      const P2 p0 = {x0[0], x0[1]}, pt = {xt[0], xt[1]};
      const P2 v1 = sub2(pt, p0), v2 = sub2(p2, p1);
      const double denom = cross2(v2, v1);
      if (fabs(denom) < small) {
        t = clampd(dot2(sub2(pt, p1), v2) / dot2(v2, v2), 0.0, 1.0);
        if (t < 0.0000000001) t = 0.0;
        if (t > 0.9999999999) t = 1.0;
        sp = clampd(dot2(sub2(add2(p1, scale2(v2, t)), p0), v1) / dot2(v1, v1), 0.0, 1.0);
      } else {
        t = clampd(cross2(sub2(p0, p1), v1) / denom, 0.0, 1.0);
        if (t < 0.0000000001) t = 0.0;
        if (t > 0.9999999999) t = 1.0;
        sp = clampd(dot2(sub2(add2(p1, scale2(v2, t)), p0), v1) / dot2(v1, v1), 0.0, 1.0);
        if (sp < 0.0000000001) sp = 0.0;
      }
      time_rm -= sp * dtb;
      const P2 hit = add2(p1, scale2(v2, t));
      const double hz = x0[2] + sp * (xt[2] - x0[2]);
      xt[0] = hit.x, xt[1] = hit.y, xt[2] = hz;
      if (m->side_type[lastCross] == 2) {
        *status |= BT_OPEN_EXIT;
        break;
      }
      const double len = sqrt(v2.x * v2.x + v2.y * v2.y);
      const P2 tangent = {v2.x / len, v2.y / len};
      double u[3];
      interpolate_velocities(m, cur, xt[0], xt[1], xt[2], h0, u);
      const P2 u2D = {u[0], u[1]};
      const P2 ut = scale2(tangent, dot2(u2D, tangent));
      u_int[0] = ut.x, u_int[1] = ut.y, u_int[2] = u[2];
      *status |= BT_SLIDE;
    } else {
      time_rm -= dtb;
      interpolate_velocities(m, cur, xt[0], xt[1], xt[2], h0, u_int);
    }
    if (time_rm <= 0.0) break;
    const double vmag2 = u_int[0] * u_int[0] + u_int[1] * u_int[1];
    if (vmag2 <= velmin2) {
      *foot = cur;
      return;
    }
    x0[0] = xt[0], x0[1] = xt[1], x0[2] = xt[2];
    iter++;
    if (iter == 38) {
      *status |= BT_ITER38;
      *foot = cur;
      return;
    }
    time_rm_3 = time_rm_2;
    time_rm_2 = time_rm_1;
    time_rm_1 = time_rm;
  }
  interpolate_velocities(m, cur, xt[0], xt[1], xt[2], h0, u_int);
  *foot = cur;
}

/* backtrackAllNodes (meshdata.cu:3547-3660), for one (side, level) */
static void backtrack_one(or_mesh* m, const or_params* p, int s, int k, int ndelt_min, int ndelt_max) {
  const int nv = m->nvrt;
  const size_t idx = (size_t)s * nv + k;
  const double* u2 = m->su2 + 2 * idx;
  double bx = u2[0], by = u2[1];
  int foot = -1, status = 0;
  const double velmin2 = p->velmin_btrack * p->velmin_btrack;
  if (!m->side_dry[s] && k >= m->side_kb[s]) {
    const double vmag2 = u2[0] * u2[0] + u2[1] * u2[1];
    if (!(vmag2 <= velmin2)) {
      double suma = 0.0;
      int icount = 0;
      for (int l = 0; l < 2; ++l) {
        const int e = m->isdel[2 * s + l];
        if (e == -1) continue;
        icount++;
        double utmp[3], vtmp[3];
        for (int j = 0; j < 3; ++j) {
          const int nd = m->elnode[3 * e + j];
          utmp[j] = m->uvw[3 * ((size_t)nd * nv + k)];
          vtmp[j] = m->uvw[3 * ((size_t)nd * nv + k) + 1];
        }
        const double* g = m->gradphi + 6 * (size_t)e;
        /* Element::grad2(double*) (meshdata.h:1094-1103) */
        const double gux = utmp[0] * g[0] + utmp[1] * g[2] + utmp[2] * g[4];
        const double guy = utmp[0] * g[1] + utmp[1] * g[3] + utmp[2] * g[5];
        const double gvx = vtmp[0] * g[0] + vtmp[1] * g[2] + vtmp[2] * g[4];
        const double gvy = vtmp[0] * g[1] + vtmp[1] * g[3] + vtmp[2] * g[5];
        suma += p->dt * sqrt((gux * gux + guy * guy) + (gvx * gvx + gvy * gvy));
      }
      int nd = (int)(suma / icount) * 4;
      nd = nd > ndelt_max ? ndelt_max : nd;
      const int ndelt = nd < ndelt_min ? ndelt_min : nd;
      const double dtbk = p->dt / ndelt;
      double u_int[3] = {0.0, 0.0, 0.0};
      side_btrack(m, p, s, k, dtbk, u_int, &foot, &status);
      bx = u_int[0];
      by = u_int[1];
    }
  }
  m->btu[2 * idx] = bx;
  m->btu[2 * idx + 1] = by;
  m->foot[idx] = foot;
  m->bstatus[idx] = status;
}

void or_backtrack(or_mesh* m, const or_params* p, const int32_t* ls, const int32_t* ll, int64_t count,
                  int nthreads) {
  const int ndelt_max = (int)fmax(1.0, p->dt / p->dtb_min);
  const int ndelt_min = (int)fmax(1.0, p->dt / p->dtb_max);
  const int64_t total = ls ? count : (int64_t)m->ns * m->nvrt;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
  for (int64_t t = 0; t < total; ++t) {
    int s, k;
    if (ls) {
      s = ls[t];
      k = ll[t];
    } else {
      s = (int)(t / m->nvrt);
      k = (int)(t % m->nvrt);
    }
    backtrack_one(m, p, s, k, ndelt_min, ndelt_max);
  }
  (void)nthreads;
}

/* ------------------------------------------------------------------ RHS producers */
/* a node's elevation BC type: its own, or the parameter iettype where never set */
static int node_etype(const or_mesh* m, const or_params* p, int i) {
  return m->etype[i] < 0 ? p->iettype : m->etype[i];
}

int or_set_tides(or_mesh* m, int32_t nbfr, const double* amig, const double* ff, const double* face,
                 const double* emo, const double* efa) {
  if (nbfr < 0) {
    set_err("nbfr < 0");
    return -1;
  }
  free(m->amig);
  free(m->ff);
  free(m->face);
  free(m->emo);
  free(m->efa);
  m->nbfr = nbfr;
  m->amig = xcalloc(nbfr, sizeof(double));
  m->ff = xcalloc(nbfr, sizeof(double));
  m->face = xcalloc(nbfr, sizeof(double));
  m->emo = xcalloc((size_t)m->n * nbfr, sizeof(double));
  m->efa = xcalloc((size_t)m->n * nbfr, sizeof(double));
  if (nbfr) {
    memcpy(m->amig, amig, 8 * (size_t)nbfr);
    memcpy(m->ff, ff, 8 * (size_t)nbfr);
    memcpy(m->face, face, 8 * (size_t)nbfr);
    memcpy(m->emo, emo, 8 * (size_t)m->n * nbfr);
    memcpy(m->efa, efa, 8 * (size_t)m->n * nbfr);
  }
  return 0;
}

void or_rhs(or_mesh* m, const or_params* p) {
  const int n = m->n, ne = m->ne, ns = m->ns, nv = m->nvrt;
  const double dt = p->dt;
  /* computeElevationBoundaryConditions, node types 1-5, nramp_elev=0, no loading tide or inverse
   * barometric term (meshdata.cu:4276-4339) */
  for (int i = 0; i < n; ++i) {
    if (!m->open[i]) continue;
    const int iettype = node_etype(m, p, i);
    const double eic = 0.0;
    const double r_eic = (1.0 - p->ramp) * eic;
    const double r_eth = p->ramp * m->eth[i];
    double r_eta1_bar = 0.0;
    if (iettype == 3 || iettype == 5) {
      const double pi = 3.14159265358979323846; /* meshdata.h:12 */
      for (int jfr = 0; jfr < m->nbfr; ++jfr) {
        const double efa = m->efa[(size_t)i * m->nbfr + jfr], emo = m->emo[(size_t)i * m->nbfr + jfr];
        const int ncyc = (int)(m->amig[jfr] * p->time / (2.0 * pi));
        const double arg = m->amig[jfr] * p->time - (2.0 * pi) * ncyc + m->face[jfr] - efa;
        r_eta1_bar += p->ramp * m->ff[jfr] * emo * cos(arg);
      }
    }
    switch (iettype) {
      case 1:
      case 2:
      case 4: m->elbc[i] = r_eic + r_eth; break;
      case 3: m->elbc[i] = r_eic + r_eta1_bar; break;
      case 5: m->elbc[i] = r_eic + r_eta1_bar + r_eth; break;
      default: break;
    }
  }
  for (int s = 0; s < ns; ++s) {
    const double* u2 = m->su2 + 2 * (size_t)s * nv;
    const double* zc = m->zc + (size_t)s * nv;
    const double* bt = m->btu + 2 * (size_t)s * nv;
    const int kb = m->side_kb[s];
    const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
    /* computeVerticallyIntegratedVelocity (meshdata.cu:4459-4476) */
    double U[2] = {0.0, 0.0};
    if (!m->side_dry[s]) {
      for (int k = kb; k < nv - 1; ++k) {
#if OR_PLANT == 6
        const double uc[2] = {u2[2 * k], u2[2 * k + 1]};
#else
        const double uc[2] = {0.5 * (u2[2 * k] + u2[2 * (k + 1)]), 0.5 * (u2[2 * k + 1] + u2[2 * (k + 1) + 1])};
#endif
        const double Dz = zc[k + 1] - zc[k];
        U[0] += uc[0] * Dz;
        U[1] += uc[1] * Dz;
      }
    }
    m->sU[2 * s] = U[0];
    m->sU[2 * s + 1] = U[1];
    /* computeVegetationAndDragParameters, iveg=0 (meshdata.cu:4615-4725) */
    m->hhat[s] = 0.0;
    m->chi[s] = 0.0;
    if (m->side_dry[s]) continue;
    const double htot = 0.5 * (m->eta2[n1] + m->eta2[n2]) + 0.5 * (m->depth[n1] + m->depth[n2]);
#if OR_PLANT == 4
    const int kc = kb;
#elif OR_PLANT == 5
    const int kc = kb + 2 < nv ? kb + 2 : kb + 1;
#else
    const int kc = kb + 1;
#endif
    const double vmag1 = sqrt(bt[2 * kc] * bt[2 * kc] + bt[2 * kc + 1] * bt[2 * kc + 1]);
    const double Cd = 0.5 * (m->cd[n1] + m->cd[n2]);
    const double chi2 = Cd * vmag1;
    const double chi = chi2;
    m->chi[s] = chi;
    if (kb == nv - 2) {
      const double fricH = htot + chi2 * dt;
      m->hhat[s] = htot * htot / fricH;
    } else {
      m->hhat[s] = htot - chi * dt;
#if OR_PLANT != 7
      if (p->ihhat == 1) m->hhat[s] = fmax(0.0, m->hhat[s]);
#endif
    }
  }
  /* computeElemAverageGhat (meshdata.cu:6183-6424), barotropic subset */
  for (int e = 0; e < ne; ++e) {
    double g[2] = {0.0, 0.0};
    if (!m->elem_dry[e]) {
      const int i34 = 3;
      for (int j = 0; j < i34; ++j) {
        const int s = m->elside[3 * e + j];
        const int n1 = m->isidenode[2 * s], n2 = m->isidenode[2 * s + 1];
        const double chi = m->chi[s];
        const double htot = 0.5 * (m->eta2[n1] + m->eta2[n2]) + 0.5 * (m->depth[n1] + m->depth[n2]);
        const int kb = m->side_kb[s];
        const double* zc = m->zc + (size_t)s * nv;
        const double* bt = m->btu + 2 * (size_t)s * nv;
        double cff1, cff2, cff3;
        if (kb == nv - 2) {
          cff1 = m->hhat[s] / htot;
          cff2 = 0.0;
          cff3 = 0.0;
        } else {
          cff1 = 1.0;
          cff2 = 1.0;
          cff3 = 0.0;
        }
        double Ustar[2] = {0.0, 0.0};
        const double Ustar_alpha[2] = {0.0, 0.0};
        if (kb == nv - 2) {
          const double Dz = zc[nv - 1] - zc[nv - 2];
          Ustar[0] += bt[2 * (nv - 1)] * Dz;
          Ustar[1] += bt[2 * (nv - 1) + 1] * Dz;
        } else {
          for (int k = kb + 1; k < nv; ++k) {
            const double Dz = zc[k] - zc[k - 1];
#if OR_PLANT == 9
            Ustar[0] += bt[2 * k] * Dz;
            Ustar[1] += bt[2 * k + 1] * Dz;
#else
            Ustar[0] += (bt[2 * k] + bt[2 * (k - 1)]) / 2.0 * Dz;
            Ustar[1] += (bt[2 * k + 1] + bt[2 * (k - 1) + 1]) / 2.0 * Dz;
#endif
          }
        }
        g[0] += Ustar[0] * cff1 - Ustar_alpha[0] * cff3;
        g[1] += Ustar[1] * cff1 - Ustar_alpha[1] * cff3;
        const double tau[2] = {0.0, 0.0};
#if OR_PLANT == 8
        const double ub[2] = {bt[2 * kb], bt[2 * kb + 1]};
#else
        const double ub[2] = {bt[2 * (kb + 1)], bt[2 * (kb + 1) + 1]};
#endif
        g[0] -= ub[0] * (chi * dt * cff2) + tau[0] * (dt * cff1);
        g[1] -= ub[1] * (chi * dt * cff2) + tau[1] * (dt * cff1);
      }
      g[0] = g[0] * (1.0 / i34);
      g[1] = g[1] * (1.0 / i34);
    }
    m->ghat1[2 * e] = g[0];
    m->ghat1[2 * e + 1] = g[1];
  }
}

/* ------------------------------------------------------------------ assembly */
/* Node::addToSparseMatrix (meshdata.h:1771-1786) */
static void add_to_row(const or_mesh* m, int node1, int node, double value, double* row) {
  if (node == node1) {
    row[m->diag_slot[node1]] += value;
    return;
  }
  const int b = m->nnp_ptr[node1], nn = m->nnp_ptr[node1 + 1] - b;
  int i;
  for (i = 0; i < nn; i++)
    if (m->nnp[b + i] == node) break;
  if (i == nn) {
    fprintf(stderr, "oracle: addToSparseMatrix: node not found\n");
    abort();
  }
  row[m->csr_slot[b + i]] += value;
}

/* The boundary sides (one element) that end at node i, in ascending side index; -1 where
 * fewer than two. */
static void node_bsides(const or_mesh* m, int i, int out[2]) {
  out[0] = out[1] = -1;
  for (int j = m->nne_ptr[i]; j < m->nne_ptr[i + 1]; ++j) {
    const int e = m->nne[j];
    for (int k = 0; k < 3; ++k) {
      const int s = m->elside[3 * e + k];
      if (m->isdel[2 * s + 1] >= 0 || (m->isidenode[2 * s] != i && m->isidenode[2 * s + 1] != i)) continue;
      if (s == out[0] || s == out[1]) continue;
      if (out[0] < 0 || s < out[0]) {
        out[1] = out[0];
        out[0] = s;
      } else if (out[1] < 0 || s < out[1]) {
        out[1] = s;
      }
    }
  }
}

void or_assemble(or_mesh* m, const or_params* p, double* val, double* b, int nthreads) {
  const int n = m->n;
  const double dt = p->dt, thetai = p->theta;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
  for (int node1 = 0; node1 < n; ++node1) {
    double* row = val + m->indptr[node1];
    const int len = m->indptr[node1 + 1] - m->indptr[node1];
    /* setupCoefficientMatrixAndRHS (meshdata.cu:3182-3311) */
    for (int i = 0; i < len; ++i) row[i] = 0.0; /* clearMatrixRow */
    double qel = 0.0;
    for (int j = m->nne_ptr[node1]; j < m->nne_ptr[node1 + 1]; ++j) {
      const int e = m->nne[j];
      const int* nd = m->elnode + 3 * e;
      const int sides[3] = {m->elside[3 * e], m->elside[3 * e + 1], m->elside[3 * e + 2]};
      /* Element::mean over nodes / sides (meshdata.cu:388-446): (v0+v1+v2+v3)/i34, v3 = 0 */
      const double grav3 = (m->grav2[nd[0]] + m->grav2[nd[1]] + m->grav2[nd[2]] + 0.0) / 3;
      const double hhatb = (m->hhat[sides[0]] + m->hhat[sides[1]] + m->hhat[sides[2]] + 0.0) / 3;
      const int id1 = local_index(m, e, node1);
      const double area = m->area[e];
      const double U[2] = {(m->sU[2 * sides[0]] + m->sU[2 * sides[1]] + m->sU[2 * sides[2]] + 0.0) / 3,
                           (m->sU[2 * sides[0] + 1] + m->sU[2 * sides[1] + 1] + m->sU[2 * sides[2] + 1] + 0.0) / 3};
      const double* gp = m->gradphi + 6 * (size_t)e;
      const double gradPhi_1[2] = {gp[2 * id1], gp[2 * id1 + 1]};
      const double I_4_1 = gradPhi_1[0] * U[0] + gradPhi_1[1] * U[1];
      const double I_4_2 = gradPhi_1[0] * m->ghat1[2 * e] + gradPhi_1[1] * m->ghat1[2 * e + 1];
#if OR_PLANT == 1
      qel += dt * area * (thetai * I_4_1 + (1 - thetai) * I_4_2);
#else
      qel += dt * area * ((1 - thetai) * I_4_1 + thetai * I_4_2);
#endif
      const double fricCoeff = grav3 * thetai * (1.0 - thetai) * dt * dt * hhatb;
      const int id2 = nxq(1 + id1, 0, 3), id3 = nxq(2 + id1, 0, 3);
      const int node2 = nd[id2], node3 = nd[id3];
      const P2 x0 = {m->x[nd[id1]], m->y[nd[id1]]}, x1 = {m->x[nd[id2]], m->y[nd[id2]]},
               x2 = {m->x[nd[id3]], m->y[nd[id3]]};
      const double ipp_ii = (1.0 / 6.0) * area;
      const double ipp_ij = (1.0 / 12.0) * area;
      const double ipp_ik = (1.0 / 12.0) * area;
      const double igpgp_ii = ((1.0 / 4.0) / area) * dot2(sub2(x2, x1), sub2(x2, x1));
      const double igpgp_ij = ((1.0 / 4.0) / area) * dot2(sub2(x2, x1), sub2(x0, x2));
      const double igpgp_ik = -igpgp_ii - igpgp_ij;
      const double I1_ii = ipp_ii + (grav3 * thetai * thetai * dt * dt * hhatb) * igpgp_ii;
      const double I1_ij = ipp_ij + (grav3 * thetai * thetai * dt * dt * hhatb) * igpgp_ij;
      const double I1_ik = ipp_ik + (grav3 * thetai * thetai * dt * dt * hhatb) * igpgp_ik;
      add_to_row(m, node1, node1, I1_ii, row);
      add_to_row(m, node1, node2, I1_ij, row);
      add_to_row(m, node1, node3, I1_ik, row);
      for (int l = 0; l < 3; l++) {
        const int node = nd[l];
        const double extraEta = 0.0;
        const double ipp = ((id1 == l) ? (2.0 / 12.0) : (1.0 / 12.0)) * area;      /* intPhiPhiTri */
        const double igpgp = area * (gp[2 * id1] * gp[2 * l] + gp[2 * id1 + 1] * gp[2 * l + 1]); /* intGradPhiGradPhiTri */
#if OR_PLANT == 2
        const double isElemWet = 1.0;
#else
        const double isElemWet = m->elem_dry[e] ? 0.0 : 1.0;
#endif
        qel += (m->eta2[node] + extraEta) * ipp - m->eta2[node] * fricCoeff * igpgp * isElemWet;
      }
      qel += (dt / 3) * m->vsource[e];
    }
    /* Boundary integral I_5 at open nodes (meshdata.cu:3373-3380, 3457-3463): over the node's two
     * boundary sides (open or land: NodeBoundary::linkSideBoundary links both kinds,
     * meshdata.cu:2198-2221, meshdata.h:1544-1555), taken in ascending side index, the non-dry
     * ones subtract 0.5 (1 - theta) dt len (U . n). Essential rows overwrite it below. I_3 is
     * zero: every flow type here is ifltype = 0. */
    if (m->open[node1] && node_etype(m, p, node1) == 0) {
      int bs[2];
      node_bsides(m, node1, bs);
      for (int l = 0; l < 2; l++) {
        const int s = bs[l];
        if (s < 0 || m->side_dry[s]) continue;
        const double Unbar = m->sU[2 * s] * m->snx[s] + m->sU[2 * s + 1] * m->sny[s];
#if OR_PLANT == 3
        qel += 0.5 * (1.0 - thetai) * dt * m->slen[s] * Unbar;
#else
        qel -= 0.5 * (1.0 - thetai) * dt * m->slen[s] * Unbar;
#endif
      }
    }
    /* setEssentialBC (meshdata.cu:4400-4446): essential at open nodes of type != 0
     * (isEssentialBC = iettype != 0, meshdata.cu:2506) */
    if (m->open[node1] && node_etype(m, p, node1) != 0) {
      for (int j = 0; j < len; ++j) row[j] = 0.0;
      row[m->diag_slot[node1]] = 1.0;
      qel = m->elbc[node1];
    } else {
      const int bb = m->nnp_ptr[node1], nn = m->nnp_ptr[node1 + 1] - bb;
      for (int j = 0; j < nn; ++j) {
        const int nb = m->nnp[bb + j];
        if (m->open[nb] && node_etype(m, p, nb) != 0) {
          qel -= row[m->csr_slot[bb + j]] * m->elbc[nb];
          row[m->csr_slot[bb + j]] = 0.0;
        }
      }
    }
    b[node1] = qel;
  }
}

/* ------------------------------------------------------------------ PCG */
/* Sum of 256*R rows of a tile in the device order: per-thread sequential over its R rows,
 * wave64 shuffle tree, then (w0 + w1) + (w2 + w3). */
static double block_tree(double v[256]) {
  double w[4];
  for (int wv = 0; wv < 4; ++wv) {
    double* l = v + 64 * wv;
    for (int off = 32; off >= 1; off >>= 1)
      for (int a = 0; a < off; ++a) l[a] = l[a] + l[a + off];
    w[wv] = l[0];
  }
  return (w[0] + w[1]) + (w[2] + w[3]);
}

static double canon_final(const double* P, int64_t T) {
  double v[256];
  for (int t = 0; t < 256; ++t) {
    double u = 0.0;
    for (int64_t mm = t; mm < T; mm += 256) u += P[mm];
    v[t] = u;
  }
  return block_tree(v);
}

/* terms may be computed on the fly from (a, b): term_i = a_i * b_i */
static double canon_dot(int32_t n, const double* a, const double* b, int32_t R, double* partial) {
  const int64_t tile_rows = 256 * (int64_t)R;
  const int64_t T = (n + tile_rows - 1) / tile_rows;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t t = 0; t < T; ++t) {
    double v[256];
    for (int th = 0; th < 256; ++th) {
      double u = 0.0;
      for (int j = 0; j < R; ++j) {
        const int64_t i = t * tile_rows + (int64_t)j * 256 + th;
        if (i < n) u += a[i] * b[i];
      }
      v[th] = u;
    }
    partial[t] = block_tree(v);
  }
  return canon_final(partial, T);
}

double or_dot_canon(int32_t n, const double* a, const double* b, int32_t R) {
  const int64_t T = (n + 256 * (int64_t)R - 1) / (256 * (int64_t)R);
  double* part = xcalloc(T, sizeof(double));
  const double r = canon_dot(n, a, b, R, part);
  free(part);
  return r;
}

static void spmv(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* x,
                 double* q) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) s += val[k] * x[indices[k]];
    q[i] = s;
  }
}

int32_t or_pcg(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
               const double* h0, const double* h1, const double* h2, double rtol, int32_t maxiter,
               int32_t checkEvery, int32_t R, double* x, double* rr_last, double* atol2_out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  (void)nthreads;
  if (maxiter <= 0) maxiter = 10 * n; /* pcg_solver.py:1082-1083 */
  if (checkEvery < 1) checkEvery = 1;
  const int64_t T = (n + 256 * (int64_t)R - 1) / (256 * (int64_t)R);
  double* part = xcalloc(T, sizeof(double));
  double *invD = xcalloc(n, sizeof(double)), *r = xcalloc(n, sizeof(double)), *z = xcalloc(n, sizeof(double));
  double *p = xcalloc(n, sizeof(double)), *q = xcalloc(n, sizeof(double));
  /* computeInvDiag (pcg_solver.py:933-938) */
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
      if (indices[k] == i) invD[i] = 1.0 / val[k];
  /* computeAtol (pcg_solver.py:928-930) */
  const double atol = rtol * sqrt(canon_dot(n, b, b, R, part));
  const double atol2 = atol * atol;
  *atol2_out = atol2;
  *rr_last = -1.0;
  int32_t result;
  if (atol == 0.0) { /* pcg_solver.py:1089-1095 */
    for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    result = 0;
    goto done;
  }
  /* extrapolateGuess (pcg_solver.py:881-886) */
  for (int32_t i = 0; i < n; ++i) {
    x[i] = h0 ? (3.0 * h0[i] - 3.0 * h1[i]) + h2[i] : 0.0;
    p[i] = 0.0;
  }
  /* computeResiduals (pcg_solver.py:889-895) */
  spmv(n, indptr, indices, val, x, q);
  for (int32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  /* f_init (pcg_solver.py:172-182) */
  for (int32_t i = 0; i < n; ++i) z[i] = invD[i] * r[i];
  double rz_cur = canon_dot(n, r, z, R, part), rz_old = 0.0;
  result = maxiter;
  /* pcg_loop.cpp:367-413 with f_update_p / f_dot / f_update_xrz */
  for (int32_t it = 0; it < maxiter; ++it) {
    const int k = it;
    const double beta = (k == 0 || rz_old == 0.0) ? 0.0 : rz_cur / rz_old;
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    spmv(n, indptr, indices, val, p, q);
    const double pq = canon_dot(n, p, q, R, part);
    const double alpha = (pq != 0.0) ? rz_cur / pq : 0.0;
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      const double zi = invD[i] * ri;
      r[i] = ri;
      z[i] = zi;
    }
    const double rr = canon_dot(n, r, r, R, part);
    const double rz_new = canon_dot(n, r, z, R, part);
    rz_old = rz_cur;
    rz_cur = rz_new;
    if ((it + 1) % checkEvery == 0 || it == maxiter - 1) {
      *rr_last = rr;
      if (rr <= atol2) {
        result = it + 1;
        break;
      }
    }
  }
done:
  free(part);
  free(invD);
  free(r);
  free(z);
  free(p);
  free(q);
  return result;
}

/* ------------------------------------------------------------------ multi-RHS block
 * Jacobi PCG of M systems on one CSR pattern (values [nnz][M], b and x [n][M]) with the
 * reference's recurrences per member (as or_pcg, cold start: x = 0, r = b), every dot product in
 * the library's block order (block_kernels.hip): a tile of 512 rows, S = 256 / M slot threads per
 * member, slot s adding rows s, s + S, ... of the tile to 0.0; the tile partial adding the slot
 * sums in slot order; the total taking RS = 1024 / M slots, slot r adding tile partials r,
 * r + RS, ... to 0.0, then the slot sums in order. */
static double blk_dot(int32_t n, int32_t M, int32_t m, const double* a, const double* b, double* part) {
  const int S = 256 / M, JT = 512 / S, RS = 1024 / M;
  const int64_t T = ((int64_t)n + 511) / 512;
  for (int64_t t = 0; t < T; ++t) {
    double tp = 0.0;
    for (int s = 0; s < S; ++s) {
      double acc = 0.0;
      for (int j = 0; j < JT; ++j) {
        const int64_t i = 512 * t + s + (int64_t)S * j;
        if (i >= n) break;
        acc += a[i * M + m] * b[i * M + m];
      }
      tp += acc;
    }
    part[t] = tp;
  }
  double tot = 0.0;
  for (int r = 0; r < RS; ++r) {
    double u = 0.0;
    for (int64_t t = r; t < T; t += RS) u += part[t];
    tot += u;
  }
  return tot;
}

int or_pcg_block(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t M, const double* val,
                 const double* b, double rtol, int32_t maxiter, int32_t checkEvery, double* x, int32_t* iters,
                 double* rr_last, int nthreads) {
  if (M < 1 || M > 256 || (M & (M - 1))) {
    set_err("M must be a power of two in [1, 256]");
    return -1;
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  (void)nthreads;
  if (maxiter <= 0) maxiter = 10 * n; /* pcg_solver.py:1082-1083 */
  if (checkEvery < 1) checkEvery = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int32_t m = 0; m < M; ++m) {
    const int64_t T = ((int64_t)n + 511) / 512;
    double* part = xcalloc(T, sizeof(double));
    double *invD = xcalloc(n, sizeof(double)), *r = xcalloc((size_t)n * M, sizeof(double));
    double *z = xcalloc((size_t)n * M, sizeof(double)), *p = xcalloc((size_t)n * M, sizeof(double));
    double* q = xcalloc((size_t)n * M, sizeof(double));
    /* computeInvDiag (pcg_solver.py:933-938) */
    for (int32_t i = 0; i < n; ++i)
      for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
        if (indices[k] == i) invD[i] = 1.0 / val[(int64_t)k * M + m];
    /* computeAtol (pcg_solver.py:928-930) */
    const double atol = rtol * sqrt(blk_dot(n, M, m, b, b, part));
    const double atol2 = atol * atol;
    rr_last[m] = -1.0;
    for (int32_t i = 0; i < n; ++i) x[(int64_t)i * M + m] = 0.0;
    if (atol == 0.0) { /* pcg_solver.py:1089-1095 */
      iters[m] = 0;
    } else {
      for (int32_t i = 0; i < n; ++i) {
        const int64_t e = (int64_t)i * M + m;
        r[e] = b[e];
        z[e] = invD[i] * r[e];
      }
      double rz_cur = blk_dot(n, M, m, r, z, part), rz_old = 0.0;
      int32_t result = maxiter;
      for (int32_t it = 0; it < maxiter; ++it) {
        const double beta = (it == 0 || rz_old == 0.0) ? 0.0 : rz_cur / rz_old;
        for (int32_t i = 0; i < n; ++i) {
          const int64_t e = (int64_t)i * M + m;
          p[e] = z[e] + beta * p[e];
        }
        for (int32_t i = 0; i < n; ++i) {
          double sq = 0.0;
          for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
            sq += val[(int64_t)k * M + m] * p[(int64_t)indices[k] * M + m];
          q[(int64_t)i * M + m] = sq;
        }
        const double pq = blk_dot(n, M, m, p, q, part);
        const double alpha = (pq != 0.0) ? rz_cur / pq : 0.0;
        for (int32_t i = 0; i < n; ++i) {
          const int64_t e = (int64_t)i * M + m;
          x[e] = x[e] + alpha * p[e];
          const double ri = r[e] - alpha * q[e];
          r[e] = ri;
          z[e] = invD[i] * ri;
        }
        const double rr = blk_dot(n, M, m, r, r, part);
        const double rz_new = blk_dot(n, M, m, r, z, part);
        rz_old = rz_cur;
        rz_cur = rz_new;
        if ((it + 1) % checkEvery == 0 || it == maxiter - 1) {
          rr_last[m] = rr;
          if (rr <= atol2) {
            result = it + 1;
            break;
          }
        }
      }
      iters[m] = result;
    }
    free(part);
    free(invD);
    free(r);
    free(z);
    free(p);
    free(q);
  }
  return 0;
}

double or_true_rel(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
                   const double* x, int32_t R) {
  const int64_t T = (n + 256 * (int64_t)R - 1) / (256 * (int64_t)R);
  double* part = xcalloc(T, sizeof(double));
  double* q = xcalloc(n, sizeof(double));
  spmv(n, indptr, indices, val, x, q);
  for (int32_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
  const double rr = canon_dot(n, q, q, R, part);
  const double bb = canon_dot(n, b, b, R, part);
  free(part);
  free(q);
  const double rn = sqrt(rr), bn = sqrt(bb);
  return bn > 0.0 ? rn / bn : rn;
}

/* ------------------------------------------------------------------ multicolour SSOR PCG
 * The reference has only the Jacobi preconditioner (pcg_solver.py:933-938); the SSOR variant
 * keeps its driver (extrapolated guess, atol = rtol ||b||, stopping rule checked every
 * check_every iterations on ||b - A x||^2, pcg_solver.py:881-947, 1075-1124,
 * pcg_loop.cpp:361-413), iterates with the Chronopoulos-Gear form of CG (A. T. Chronopoulos and
 * C. W. Gear, J. Comput. Appl. Math. 25 (1989) 153-168: gamma = r.z and delta = z.Az in one
 * reduction, beta = gamma/gamma_prev, alpha = gamma/(delta - beta gamma/alpha_prev), q = Az +
 * beta q), and replaces the preconditioner by symmetric SOR
 *   M = (D/w + L) (D/w)^-1 (D/w + U)
 * in a multicolour ordering, applied with Eisenstat's trick (S. C. Eisenstat, "Efficient
 * implementation of a class of preconditioned conjugate gradient methods", SIAM J. Sci. Stat.
 * Comput. 2 (1981) 1-4): CG runs on A^ = (D~+L)^-1 A (D~+U)^-1 (D~ = D/w) with preconditioner D~,
 *   A^ v = t + (D~+L)^-1 (v - (2D~ - D) t),   t = (D~+U)^-1 v,
 * so one iteration reads the off-diagonal matrix once. Iterates map back as x = (D~+U)^-1 y and
 * the true residual is b - A x = (D~+L) r.
 *
 * Layout: nodes are coloured greedily (first fit) in smallest-last order (D. W. Matula and
 * L. L. Beck, J. ACM 30 (1983) 417-427), ties broken by the smaller index; rows are renumbered
 * colour-major, ascending original index inside a colour, and each colour block is padded
 * to a multiple of the 256*R-row reduction tile with identity rows (d = 1, b = 0). Dot products
 * use the fixed tile tree over the padded order. Row sums run over ascending new column index. */

typedef struct {
  int64_t key; /* degree << 32 | index */
} hent;

static void heap_push(hent* h, int64_t* len, int64_t key) {
  int64_t i = (*len)++;
  while (i > 0) {
    const int64_t par = (i - 1) / 2;
    if (h[par].key <= key) break;
    h[i] = h[par];
    i = par;
  }
  h[i].key = key;
}

static int64_t heap_pop(hent* h, int64_t* len) {
  const int64_t top = h[0].key;
  const int64_t last = h[--(*len)].key;
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= *len) break;
    if (c + 1 < *len && h[c + 1].key < h[c].key) c++;
    if (last <= h[c].key) break;
    h[i] = h[c];
    i = c;
  }
  if (*len > 0) h[i].key = last;
  return top;
}

int32_t or_colour(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour) {
  int32_t* deg = xcalloc(n > 0 ? n : 1, sizeof(int32_t));
  char* removed = xcalloc(n > 0 ? n : 1, 1);
  int32_t* order = xcalloc(n > 0 ? n : 1, sizeof(int32_t));
  int64_t cap = (int64_t)n + indptr[n] + 1, len = 0;
  hent* h = xcalloc(cap, sizeof(hent));
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
      if (indices[k] != i) deg[i]++;
    heap_push(h, &len, ((int64_t)deg[i] << 32) | i);
  }
  int32_t cnt = 0;
  while (len > 0) {
    const int64_t key = heap_pop(h, &len);
    const int32_t i = (int32_t)(key & 0xffffffff), d = (int32_t)(key >> 32);
    if (removed[i] || d != deg[i]) continue;
    removed[i] = 1;
    order[cnt++] = i;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j != i && !removed[j]) {
        deg[j]--;
        heap_push(h, &len, ((int64_t)deg[j] << 32) | j);
      }
    }
  }
  int32_t ncol = 0;
  int32_t* stamp = xcalloc((size_t)n + 2, sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) colour[i] = -1;
  for (int32_t m = n - 1; m >= 0; --m) {
    const int32_t i = order[m];
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j != i && colour[j] >= 0) stamp[colour[j]] = m + 1;
    }
    int32_t c = 0;
    while (stamp[c] == m + 1) c++;
    colour[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(stamp);
  free(h);
  free(order);
  free(removed);
  free(deg);
  return ncol;
}

typedef struct or_ssor {
  int32_t n, np, ncol, R;
  int32_t* cstart;  /* [ncol+1] padded block starts */
  int32_t* perm;    /* [np] new -> old, -1 for padding */
  int32_t* newpos;  /* [n] old -> new */
  int32_t *lptr, *lcol, *uptr, *ucol; /* [np+1], [*] */
  int32_t *lsrc, *usrc, *dsrc;        /* slots of the entries in the CSR value array (-1: padding row) */
  double *lval, *uval, *d;            /* off-diagonal values, diagonal [np] */
} ssor_layout;

static int cmp_pair(const void* a, const void* b) {
  const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

static int ssor_build(ssor_layout* L, int32_t n, const int32_t* indptr, const int32_t* indices, int32_t R) {
  memset(L, 0, sizeof(*L));
  L->R = R;
  const int32_t T = 256 * R;
  int32_t* colour = xcalloc(n > 0 ? n : 1, sizeof(int32_t));
  const int32_t C = n > 0 ? or_colour(n, indptr, indices, colour) : 0;
  int32_t* count = xcalloc((size_t)C + 1, sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) count[colour[i]]++;
  L->n = n;
  L->ncol = C;
  L->cstart = xcalloc((size_t)C + 1, sizeof(int32_t));
  for (int32_t c = 0; c < C; ++c) L->cstart[c + 1] = L->cstart[c] + (count[c] + T - 1) / T * T;
  const int32_t np = L->cstart[C];
  L->np = np;
  L->perm = xcalloc(np > 0 ? np : 1, sizeof(int32_t));
  L->newpos = xcalloc(n > 0 ? n : 1, sizeof(int32_t));
  for (int32_t v = 0; v < np; ++v) L->perm[v] = -1;
  int32_t* fill = xcalloc((size_t)C + 1, sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) {
    const int32_t c = colour[i];
    const int32_t v = L->cstart[c] + fill[c]++;
    L->perm[v] = i;
    L->newpos[i] = v;
  }
  L->lptr = xcalloc((size_t)np + 1, sizeof(int32_t));
  L->uptr = xcalloc((size_t)np + 1, sizeof(int32_t));
  L->d = xcalloc(np > 0 ? np : 1, sizeof(double));
  L->dsrc = xcalloc(np > 0 ? np : 1, sizeof(int32_t));
  int bad = 0;
  for (int32_t v = 0; v < np; ++v) {
    const int32_t i = L->perm[v];
    int32_t nl = 0, nu = 0;
    if (i >= 0) {
      const int32_t c = colour[i];
      int has_diag = 0;
      for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
        const int32_t j = indices[k];
        if (j == i) {
          has_diag = 1;
          continue;
        }
        const int32_t w = L->newpos[j];
        if (w < L->cstart[c]) nl++;
        else if (w >= L->cstart[c + 1]) nu++;
        else bad = 1;
      }
      if (!has_diag) bad = 1;
    }
    L->lptr[v + 1] = L->lptr[v] + nl;
    L->uptr[v + 1] = L->uptr[v] + nu;
  }
  L->lcol = xcalloc((size_t)L->lptr[np] + 1, sizeof(int32_t));
  L->ucol = xcalloc((size_t)L->uptr[np] + 1, sizeof(int32_t));
  L->lval = xcalloc((size_t)L->lptr[np] + 1, sizeof(double));
  L->uval = xcalloc((size_t)L->uptr[np] + 1, sizeof(double));
  L->lsrc = xcalloc((size_t)L->lptr[np] + 1, sizeof(int32_t));
  L->usrc = xcalloc((size_t)L->uptr[np] + 1, sizeof(int32_t));
  int64_t* tmp = xcalloc((size_t)64 + (indptr[n] > 0 ? indptr[n] : 1), sizeof(int64_t));
  for (int32_t v = 0; v < np; ++v) {
    const int32_t i = L->perm[v];
    L->dsrc[v] = -1;
    if (i < 0) continue;
    const int32_t c = colour[i];
    int32_t m = 0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j == i) {
        L->dsrc[v] = k;
        continue;
      }
      tmp[m++] = ((int64_t)L->newpos[j] << 32) | (int64_t)(k - indptr[i]);
    }
    qsort(tmp, m, sizeof(int64_t), cmp_pair);
    int32_t a = L->lptr[v], b = L->uptr[v];
    for (int32_t q = 0; q < m; ++q) {
      const int32_t w = (int32_t)(tmp[q] >> 32), k = indptr[i] + (int32_t)(tmp[q] & 0xffffffff);
      if (w < L->cstart[c]) L->lcol[a] = w, L->lsrc[a++] = k;
      else L->ucol[b] = w, L->usrc[b++] = k;
    }
  }
  free(tmp);
  free(fill);
  free(count);
  free(colour);
  if (bad) set_err("SSOR layout: a row lacks its diagonal or the colouring is invalid");
  return bad ? -1 : 0;
}

/* values of the split matrix from the CSR values (padding rows: d = 1) */
static void ssor_fill(ssor_layout* L, const double* val) {
  for (int32_t v = 0; v < L->np; ++v) L->d[v] = L->dsrc[v] >= 0 ? val[L->dsrc[v]] : 1.0;
  for (int32_t k = 0; k < L->lptr[L->np]; ++k) L->lval[k] = val[L->lsrc[k]];
  for (int32_t k = 0; k < L->uptr[L->np]; ++k) L->uval[k] = val[L->usrc[k]];
}

static void ssor_free(ssor_layout* L) {
  free(L->lsrc);
  free(L->usrc);
  free(L->dsrc);
  free(L->cstart);
  free(L->perm);
  free(L->newpos);
  free(L->lptr);
  free(L->lcol);
  free(L->uptr);
  free(L->ucol);
  free(L->lval);
  free(L->uval);
  free(L->d);
}

static inline double lsum(const ssor_layout* L, int32_t v, const double* x) {
  double acc = 0.0;
  for (int32_t k = L->lptr[v]; k < L->lptr[v + 1]; ++k) acc += L->lval[k] * x[L->lcol[k]];
  return acc;
}

static inline double usum(const ssor_layout* L, int32_t v, const double* x) {
  double acc = 0.0;
  for (int32_t k = L->uptr[v]; k < L->uptr[v + 1]; ++k) acc += L->uval[k] * x[L->ucol[k]];
  return acc;
}

or_ssor* or_ssor_new(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t R) {
  ssor_layout* L = xcalloc(1, sizeof(ssor_layout));
  if (ssor_build(L, n, indptr, indices, R) != 0) {
    ssor_free(L);
    free(L);
    return NULL;
  }
  return L;
}

void or_ssor_free(or_ssor* L) {
  if (!L) return;
  ssor_free(L);
  free(L);
}

int32_t or_ssor_ncolours(const or_ssor* L) { return L->ncol; }

int32_t or_pcg_ssor(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
                    const double* h0, const double* h1, const double* h2, double omega, double rtol,
                    int32_t maxiter, int32_t checkEvery, int32_t R, double* x, double* rr_last, double* atol2_out,
                    int32_t* ncolours, int nthreads) {
  or_ssor* L = or_ssor_new(n, indptr, indices, R);
  if (!L) return -1;
  if (ncolours) *ncolours = L->ncol;
  const int32_t it = or_ssor_solve(L, val, b, h0, h1, h2, omega, rtol, maxiter, checkEvery, x, rr_last, atol2_out,
                                   nthreads);
  or_ssor_free(L);
  return it;
}

int32_t or_ssor_solve(or_ssor* L, const double* val, const double* b, const double* h0, const double* h1,
                      const double* h2, double omega, double rtol, int32_t maxiter, int32_t checkEvery, double* x,
                      double* rr_last, double* atol2_out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  (void)nthreads;
  const int32_t n = L->n, R = L->R;
  if (maxiter <= 0) maxiter = 10 * n;
  if (checkEvery < 1) checkEvery = 1;
  ssor_fill(L, val);
  const int32_t np = L->np, C = L->ncol;
  const int64_t T = np / (256 * (int64_t)R);
  double* part = xcalloc(T > 0 ? T : 1, sizeof(double));
  double *dt = xcalloc(np + 1, sizeof(double)), *kd = xcalloc(np + 1, sizeof(double));
  double *bp = xcalloc(np + 1, sizeof(double)), *xp = xcalloc(np + 1, sizeof(double));
  double *y = xcalloc(np + 1, sizeof(double)), *r = xcalloc(np + 1, sizeof(double));
  double *z = xcalloc(np + 1, sizeof(double)), *p = xcalloc(np + 1, sizeof(double));
  double *t = xcalloc(np + 1, sizeof(double)), *s = xcalloc(np + 1, sizeof(double));
  double *q = xcalloc(np + 1, sizeof(double)), *w = xcalloc(np + 1, sizeof(double));
  double* q2 = xcalloc(np + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t v = 0; v < np; ++v) {
    dt[v] = L->d[v] / omega;
    kd[v] = 2.0 * dt[v] - L->d[v];
    const int32_t i = L->perm[v];
    bp[v] = i >= 0 ? b[i] : 0.0;
    xp[v] = (i >= 0 && h0) ? (3.0 * h0[i] - 3.0 * h1[i]) + h2[i] : 0.0;
  }
  const double atol = rtol * sqrt(canon_dot(np, bp, bp, R, part));
  const double atol2 = atol * atol;
  *atol2_out = atol2;
  *rr_last = -1.0;
  int32_t result;
  if (atol == 0.0) {
    for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    result = 0;
    goto done;
  }
  /* r_true = b - A x0 and y0 = (D~+U) x0 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t v = 0; v < np; ++v) {
    const double au = usum(L, v, xp);
    const double ax = (lsum(L, v, xp) + L->d[v] * xp[v]) + au;
    w[v] = bp[v] - ax;
    y[v] = dt[v] * xp[v] + au;
  }
  /* r0 = (D~+L)^-1 r_true, colour by colour */
  for (int32_t c = 0; c < C; ++c) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t v = L->cstart[c]; v < L->cstart[c + 1]; ++v) r[v] = (w[v] - lsum(L, v, r)) / dt[v];
  }
  /* Chronopoulos-Gear recurrences: one global reduction (gamma = r.z, delta = z.A^z) per
   * iteration, so the vector updates fold into the next backward sweep. z = D~ r. */
  for (int32_t v = 0; v < np; ++v) p[v] = 0.0, q[v] = 0.0;
  double alpha = 0.0, beta = 0.0, gamma_prev = 0.0, alpha_prev = 0.0;
  result = maxiter;
  for (int32_t it = -1; it < maxiter; ++it) {
    const int update = it >= 0;
    if (update) {
      const double gamma = canon_dot(np, r, z, R, part), delta = canon_dot(np, z, s, R, part);
      if (it == 0) {
        beta = 0.0;
        alpha = (delta != 0.0) ? gamma / delta : 0.0;
      } else {
        beta = (gamma_prev != 0.0) ? gamma / gamma_prev : 0.0;
        const double den = delta - beta * gamma / alpha_prev;
        alpha = (alpha_prev != 0.0 && den != 0.0) ? gamma / den : 0.0;
      }
      gamma_prev = gamma;
      alpha_prev = alpha;
    }
    /* backward sweep, last colour first: p = z + beta p, q = A^z + beta q, y += alpha p,
     * r -= alpha q (iterations only), then z = D~ r and t = (D~+U)^-1 z */
    for (int32_t c = C - 1; c >= 0; --c)
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
      for (int32_t v = L->cstart[c]; v < L->cstart[c + 1]; ++v) {
        if (update) {
          const double zr = dt[v] * r[v];
          p[v] = zr + beta * p[v];
          q[v] = s[v] + beta * q[v];
          y[v] = y[v] + alpha * p[v];
          r[v] = r[v] - alpha * q[v];
        }
        z[v] = dt[v] * r[v];
        t[v] = (z[v] - usum(L, v, t)) / dt[v];
      }
    /* forward sweep, first colour first: u = (D~+L)^-1 (z - (2D~ - D) t), A^z = t + u */
    for (int32_t c = 0; c < C; ++c)
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
      for (int32_t v = L->cstart[c]; v < L->cstart[c + 1]; ++v) {
        w[v] = ((z[v] - kd[v] * t[v]) - lsum(L, v, w)) / dt[v];
        s[v] = t[v] + w[v];
      }
    if (update && ((it + 1) % checkEvery == 0 || it == maxiter - 1)) {
      /* ||b - A x||^2 = ||(D~+L) r||^2 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
      for (int32_t v = 0; v < np; ++v) q2[v] = dt[v] * r[v] + lsum(L, v, r);
      const double rr = canon_dot(np, q2, q2, R, part);
      *rr_last = rr;
      if (rr <= atol2) {
        result = it + 1;
        break;
      }
    }
  }
  /* x = (D~+U)^-1 y, back to the original numbering */
  for (int32_t c = C - 1; c >= 0; --c) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t v = L->cstart[c]; v < L->cstart[c + 1]; ++v) xp[v] = (y[v] - usum(L, v, xp)) / dt[v];
  }
  for (int32_t i = 0; i < n; ++i) x[i] = xp[L->newpos[i]];
done:
  free(part);
  free(dt);
  free(kd);
  free(bp);
  free(xp);
  free(y);
  free(r);
  free(z);
  free(p);
  free(t);
  free(s);
  free(q);
  free(w);
  free(q2);
  return result;
}

/* ------------------------------------------------------------------ test hooks */
int or_geometry(const or_mesh* m, int what, double* out) {
  switch (what) {
    case 0: memcpy(out, m->area, 8 * (size_t)m->ne); return 0;
    case 1: memcpy(out, m->gradphi, 48 * (size_t)m->ne); return 0;
    case 2:
      for (int s = 0; s < m->ns; ++s) out[2 * s] = m->scx[s], out[2 * s + 1] = m->scy[s];
      return 0;
    case 3: memcpy(out, m->slen, 8 * (size_t)m->ns); return 0;
    case 4:
      for (int s = 0; s < m->ns; ++s) out[2 * s] = m->snx[s], out[2 * s + 1] = m->sny[s];
      return 0;
    default: set_err("unknown geometry array %d", what); return -1;
  }
}

void or_interp(const or_mesh* m, int32_t e, double x, double y, double z, double h0, double out[3]) {
  interpolate_velocities(m, e, x, y, z, h0, out);
}
