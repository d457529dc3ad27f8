// topology.cpp -- node balls, sides and the CSR pattern of the wave-equation matrix, built
// the way the reference's grid preprocessing defines them, plus RCM renumbering and the
// element/side geometry.
//
//   node -> elements (increasing element id)   read_and_process_grid.cpp:27-84
//   element across each side (ic3)             read_and_process_grid.cpp:272-344
//   counter-clockwise balls                    read_and_process_grid.cpp:496-574
//   node -> neighbour nodes (ball order)       read_and_process_grid.cpp:87-174
//   side numbering, isdel, isidenode           read_and_process_grid.cpp:346-437
//   CSR rows, csrSlot, diagonal slot           meshdata.cu:2094-2141, 1931-1935
//   side centre / normal / length              meshdata.cu:2274-2312
//   element area, ssign, gradPhi               meshdata.cu:2350-2440, 749-771
#include <algorithm>
#include <cmath>
#include <deque>
#include <numeric>

#include "abi_util.h"

namespace gsh {

static inline int nxq(int offset, int index, int size) { return (index + offset + 16 * size) % size; }

void build_topology(const HostMesh& m, Topology& t) {
  const int32_t n = m.n, ne = m.ne;
  const auto& en = m.elnode;
  // ---- node -> elements in increasing element order (buildListsOfElementNeighbors)
  std::vector<int32_t> cnt(n + 1, 0);
  for (int32_t e = 0; e < ne; ++e)
    for (int j = 0; j < 3; ++j) {
      const int32_t v = en[3 * e + j];
      if (v < 0 || v >= n) throw Error(GSH_E_TOPOLOGY, "element " + std::to_string(e) + " has an invalid node id");
      cnt[v + 1]++;
    }
  for (int32_t i = 0; i < n; ++i)
    if (cnt[i + 1] == 0) throw Error(GSH_E_TOPOLOGY, "hanging node " + std::to_string(i));
  t.ball_ptr.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) t.ball_ptr[i + 1] = t.ball_ptr[i] + cnt[i + 1];
  t.ball_elem.assign(t.ball_ptr[n], -1);
  {
    std::vector<int32_t> fill(t.ball_ptr.begin(), t.ball_ptr.end() - 1);
    for (int32_t e = 0; e < ne; ++e)
      for (int j = 0; j < 3; ++j) t.ball_elem[fill[en[3 * e + j]]++] = e;
  }
  auto ball = [&](int32_t v, int32_t j) { return t.ball_elem[t.ball_ptr[v] + j]; };
  auto ballSize = [&](int32_t v) { return t.ball_ptr[v + 1] - t.ball_ptr[v]; };
  auto localIdx = [&](int32_t e, int32_t v) {
    for (int l = 0; l < 3; ++l)
      if (en[3 * e + l] == v) return l;
    return -1;
  };

  // ---- element across each side (buildElementSideElementTable, first part)
  t.ic3.assign(3 * (size_t)ne, -1);
  for (int32_t ie = 0; ie < ne; ++ie)
    for (int k = 0; k < 3; ++k) {
      const int32_t n1 = en[3 * ie + nxq(1, k, 3)], n2 = en[3 * ie + nxq(2, k, 3)];
      int32_t found = -1;
      for (int32_t l = 0; l < ballSize(n1); ++l) {
        const int32_t je = ball(n1, l);
        if (je != ie && localIdx(je, n2) >= 0) {
          found = je;
          break;
        }
      }
      t.ic3[3 * ie + k] = found;
      if (found >= 0)
        for (int l = 0; l < 3; ++l)
          if (en[3 * found + l] == n1 && en[3 * found + nxq(1, l, 3)] == n2)
            throw Error(GSH_E_TOPOLOGY, "elements " + std::to_string(ie) + " and " + std::to_string(found) +
                                            " have opposite orientation");
    }

  // ---- boundary nodes: nodes on a side with no element across
  t.isbnd.assign(n, 0);
  for (int32_t ie = 0; ie < ne; ++ie)
    for (int k = 0; k < 3; ++k)
      if (t.ic3[3 * ie + k] < 0) {
        t.isbnd[en[3 * ie + nxq(1, k, 3)]] = 1;
        t.isbnd[en[3 * ie + nxq(2, k, 3)]] = 1;
      }

  // ---- counter-clockwise balls (reorderElementsCounterClockwise)
  for (int32_t i = 0; i < n; ++i) {
    int32_t* b = &t.ball_elem[t.ball_ptr[i]];
    const int32_t nb = ballSize(i);
    if (t.isbnd[i]) {
      int icount = 0;
      int32_t first = -1;
      for (int32_t l = 0; l < nb; ++l) {
        const int32_t ie = b[l];
        const int il = localIdx(ie, i);
        if (t.ic3[3 * ie + nxq(-1, il, 3)] == -1) {
          ++icount;
          first = ie;
        }
      }
      if (icount != 1)
        throw Error(GSH_E_TOPOLOGY, "illegal boundary node " + std::to_string(i) + ": " + std::to_string(icount) +
                                        " boundary elements");
      b[0] = first;
    }
    for (int32_t j = 1; j < nb; ++j) {
      const int32_t ie = b[j - 1];
      const int il = localIdx(ie, i);
      const int32_t next = t.ic3[3 * ie + nxq(-2, il, 3)];
      if (next == -1) throw Error(GSH_E_TOPOLOGY, "incomplete ball around node " + std::to_string(i));
      b[j] = next;
    }
  }
  t.ball_loc.resize(t.ball_elem.size());
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = t.ball_ptr[i]; j < t.ball_ptr[i + 1]; ++j) t.ball_loc[j] = (int8_t)localIdx(t.ball_elem[j], i);

  // ---- neighbour nodes in ball order (buildListsOfNodeNeighbors)
  t.nbr_ptr.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) t.nbr_ptr[i + 1] = t.nbr_ptr[i] + ballSize(i) + (t.isbnd[i] ? 1 : 0);
  t.nbr_node.assign(t.nbr_ptr[n], -1);
  for (int32_t i = 0; i < n; ++i) {
    int32_t c = t.nbr_ptr[i];
    const int32_t nb = ballSize(i);
    for (int32_t j = 0; j < nb; ++j) {
      const int32_t ie = ball(i, j);
      const int ii = t.ball_loc[t.ball_ptr[i] + j];
      t.nbr_node[c++] = en[3 * ie + nxq(1, ii, 3)];
      if (j == nb - 1) {
        const int32_t last = en[3 * ie + nxq(2, ii, 3)];
        if (!t.isbnd[i]) {
          if (last != t.nbr_node[t.nbr_ptr[i]]) throw Error(GSH_E_TOPOLOGY, "broken ball for node " + std::to_string(i));
        } else {
          t.nbr_node[c++] = last;
        }
      }
    }
  }

  // ---- sides (buildElementSideElementTable, second part)
  t.elside.assign(3 * (size_t)ne, -1);
  int32_t ns = 0;
  for (int32_t ie = 0; ie < ne; ++ie)
    for (int j = 0; j < 3; ++j) {
      const int32_t je = t.ic3[3 * ie + j];
      if (je != -1) {
        if (ie < je) {
          t.elside[3 * ie + j] = ns;
          int l = -1;
          for (int k = 0; k < 3; ++k)
            if (t.ic3[3 * je + k] == ie) {
              l = k;
              break;
            }
          if (l == -1) throw Error(GSH_E_TOPOLOGY, "wrong ball info for elements " + std::to_string(ie));
          t.elside[3 * je + l] = ns;
          ++ns;
        }
      } else {
        t.elside[3 * ie + j] = ns++;
      }
    }
  if (ns < ne || ns < n) throw Error(GSH_E_TOPOLOGY, "weird grid with ns < ne or ns < np");
  t.ns = ns;
  t.isdel.assign(2 * (size_t)ns, -1);
  t.isidenode.assign(2 * (size_t)ns, -1);
  for (int32_t ie = 0; ie < ne; ++ie)
    for (int j = 0; j < 3; ++j) {
      const int32_t s = t.elside[3 * ie + j];
      const int32_t n1 = en[3 * ie + nxq(1, j, 3)], n2 = en[3 * ie + nxq(2, j, 3)];
      const int32_t je = t.ic3[3 * ie + j];
      if (je == -1 || ie < je) {
        t.isdel[2 * s] = ie;
        t.isdel[2 * s + 1] = je;
        t.isidenode[2 * s] = n1;
        t.isidenode[2 * s + 1] = n2;
      } else {
        t.isdel[2 * s] = je;
        t.isdel[2 * s + 1] = ie;
        t.isidenode[2 * s] = n2;
        t.isidenode[2 * s + 1] = n1;
      }
    }
  // Side::neigh(j) (meshdata.h:1409-1422), the Shapiro filter's neighbours: in element j & 1 of
  // the side, the side before it (j < 2) or after it (j >= 2) in that element's side order
  t.side_nbr.assign(4 * (size_t)ns, -1);
  for (int32_t s = 0; s < ns; ++s) {
    if (t.isdel[2 * s + 1] < 0) continue;
    for (int j = 0; j < 4; ++j) {
      const int32_t e = t.isdel[2 * s + (j & 1)];
      const int off = 2 * (j >> 1) - 1;
      int loc = 0;
      while (t.elside[3 * e + loc] != s) ++loc;
      t.side_nbr[4 * (size_t)s + j] = t.elside[3 * e + (loc + off + 48) % 3];
    }
  }
  t.side_type.assign(ns, 0);
  for (int32_t s = 0; s < ns; ++s)
    if (t.isdel[2 * s + 1] < 0) {
      const bool open = m.open_bnd[t.isidenode[2 * s]] && m.open_bnd[t.isidenode[2 * s + 1]];
      t.side_type[s] = open ? 2 : 1;
    }
  // The boundary sides ending at each boundary node, ascending side index (-1 where fewer than
  // two): the sides of the open-boundary integral I_5 (meshdata.cu:3373-3380). An open node
  // links at most two (the assert of NodeBoundary::linkSideBoundary, meshdata.h:1544-1555).
  t.node_bside.assign(2 * (size_t)n, -1);
  for (int32_t s = 0; s < ns; ++s)
    if (t.isdel[2 * s + 1] < 0)
      for (int j = 0; j < 2; ++j) {
        const int32_t i = t.isidenode[2 * s + j];
        int32_t* b = &t.node_bside[2 * (size_t)i];
        if (b[0] < 0) {
          b[0] = s;
        } else if (b[1] < 0) {
          b[1] = s;
        } else if (m.open_bnd[i]) {
          throw Error(GSH_E_TOPOLOGY, "open boundary node " + std::to_string(i) + " ends more than two boundary sides");
        }
      }

  // ---- CSR pattern (populateNodes2): row i = sorted {i} U neighbours; indptr[i] = nbrStart + i
  t.indptr.assign(n + 1, 0);
  for (int32_t i = 0; i <= n; ++i) t.indptr[i] = t.nbr_ptr[i] + i;
  const int64_t nnz = t.indptr[n];
  t.indices.assign(nnz, -1);
  t.csr_slot.assign(t.nbr_node.size(), 0);
  t.diag_off.assign(n, 0);
  t.max_row = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t b = t.nbr_ptr[i], e = t.nbr_ptr[i + 1], nn = e - b;
    t.max_row = std::max(t.max_row, nn + 1);
    if (nn + 1 > 32767) throw Error(GSH_E_TOPOLOGY, "CSR row too long");
    int16_t d = 0;
    for (int32_t j = b; j < e; ++j) d += (t.nbr_node[j] < i);
    t.diag_off[i] = d;
    for (int32_t j = b; j < e; ++j) {
      const int32_t col = t.nbr_node[j];
      int16_t idx = (i < col) ? 1 : 0;
      for (int32_t k = b; k < e; ++k) idx += (t.nbr_node[k] < col);
      t.csr_slot[j] = idx;
    }
    const int32_t row = t.indptr[i];
    t.indices[row + d] = i;
    for (int32_t j = b; j < e; ++j) t.indices[row + t.csr_slot[j]] = t.nbr_node[j];
  }
  for (int64_t k = 0; k < nnz; ++k)
    if (t.indices[k] < 0) throw Error(GSH_E_TOPOLOGY, "duplicate neighbour in a node ball");
}

void build_geometry(const HostMesh& m, const Topology& t, Geometry& g) {
  const int32_t ne = m.ne, ns = t.ns;
  const auto& X = m.x;
  const auto& Y = m.y;
  g.area.resize(ne);
  g.gradphi.resize(6 * (size_t)ne);
  g.ssign.resize(3 * (size_t)ne);
  for (int32_t e = 0; e < ne; ++e) {
    const int32_t a = m.elnode[3 * e], b = m.elnode[3 * e + 1], c = m.elnode[3 * e + 2];
    // signedArea(p1, p2, p3) = 0.5*((p1.x-p3.x)*(p2.y-p3.y) - (p2.x-p3.x)*(p1.y-p3.y))
    const double area = 0.5 * ((X[a] - X[c]) * (Y[b] - Y[c]) - (X[b] - X[c]) * (Y[a] - Y[c]));
    if (!(area > 0.0)) throw Error(GSH_E_TOPOLOGY, "element " + std::to_string(e) + " is not counter-clockwise");
    g.area[e] = area;
    for (int j = 0; j < 3; ++j) {
      const int32_t p2 = m.elnode[3 * e + nxq(1, j, 3)], p3 = m.elnode[3 * e + nxq(2, j, 3)];
      g.gradphi[6 * (size_t)e + 2 * j] = (Y[p2] - Y[p3]) / (2 * area);
      g.gradphi[6 * (size_t)e + 2 * j + 1] = (X[p3] - X[p2]) / (2 * area);
      const int32_t s = t.elside[3 * e + j];
      g.ssign[3 * (size_t)e + j] = (t.isdel[2 * s] == e) ? 1.0 : -1.0;
    }
  }
  g.side_cx.resize(ns);
  g.side_cy.resize(ns);
  g.side_len.resize(ns);
  g.side_nx.resize(ns);
  g.side_ny.resize(ns);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t n1 = t.isidenode[2 * s], n2 = t.isidenode[2 * s + 1];
    g.side_cx[s] = 0.5 * (X[n1] + X[n2]);
    g.side_cy[s] = 0.5 * (Y[n1] + Y[n2]);
    const double vx = -(Y[n1] - Y[n2]), vy = (X[n1] - X[n2]);
    const double l = std::sqrt(vx * vx + vy * vy);
    g.side_nx[s] = vx / l;
    g.side_ny[s] = vy / l;
    const double dx = X[n2] - X[n1], dy = Y[n2] - Y[n1], dz = 0.0;
    g.side_len[s] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
}

// ------------------------------------------------------------------ RCM
static void node_graph(int32_t n, int32_t ne, const std::vector<int32_t>& en, std::vector<int32_t>& ptr,
                       std::vector<int32_t>& adj) {
  std::vector<std::pair<int32_t, int32_t>> edges;
  edges.reserve(6 * (size_t)ne);
  for (int32_t e = 0; e < ne; ++e)
    for (int j = 0; j < 3; ++j) {
      const int32_t a = en[3 * e + j], b = en[3 * e + (j + 1) % 3];
      edges.emplace_back(a, b);
      edges.emplace_back(b, a);
    }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  ptr.assign(n + 1, 0);
  for (auto& p : edges) ptr[p.first + 1]++;
  for (int32_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  adj.resize(edges.size());
  for (size_t k = 0; k < edges.size(); ++k) adj[k] = edges[k].second;
}

static int32_t bfs_levels(int32_t root, const std::vector<int32_t>& ptr, const std::vector<int32_t>& adj,
                          std::vector<int32_t>& level, std::vector<int32_t>& last_level) {
  std::fill(level.begin(), level.end(), -1);
  std::vector<int32_t> q{root};
  level[root] = 0;
  int32_t depth = 0;
  for (size_t h = 0; h < q.size(); ++h) {
    const int32_t v = q[h];
    for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k)
      if (level[adj[k]] < 0) {
        level[adj[k]] = level[v] + 1;
        depth = std::max(depth, level[v] + 1);
        q.push_back(adj[k]);
      }
  }
  last_level.clear();
  for (int32_t v : q)
    if (level[v] == depth) last_level.push_back(v);
  return depth;
}

std::vector<int32_t> rcm_order(int32_t n, int32_t ne, const std::vector<int32_t>& elnode) {
  std::vector<int32_t> ptr, adj;
  node_graph(n, ne, elnode, ptr, adj);
  auto deg = [&](int32_t v) { return ptr[v + 1] - ptr[v]; };
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<int32_t> level(n), last;
  for (int32_t s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    // pseudo-peripheral start (George-Liu) within this component
    int32_t root = s0;
    int32_t ecc = bfs_levels(root, ptr, adj, level, last);
    for (int it = 0; it < 8; ++it) {
      int32_t cand = last[0];
      for (int32_t v : last)
        if (deg(v) < deg(cand) || (deg(v) == deg(cand) && v < cand)) cand = v;
      std::vector<int32_t> last2;
      const int32_t e2 = bfs_levels(cand, ptr, adj, level, last2);
      if (e2 <= ecc) break;
      root = cand;
      ecc = e2;
      last.swap(last2);
    }
    const size_t start = order.size();
    order.push_back(root);
    seen[root] = 1;
    std::vector<int32_t> nb;
    for (size_t h = start; h < order.size(); ++h) {
      const int32_t v = order[h];
      nb.clear();
      for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k)
        if (!seen[adj[k]]) nb.push_back(adj[k]);
      std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg(a) < deg(b) || (deg(a) == deg(b) && a < b); });
      for (int32_t w : nb) {
        seen[w] = 1;
        order.push_back(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

void renumber(HostMesh& m, const std::vector<int32_t>& perm, std::vector<int32_t>& elem_perm) {
  const int32_t n = m.n, ne = m.ne;
  std::vector<int32_t> newid(n);
  for (int32_t k = 0; k < n; ++k) newid[perm[k]] = k;
  auto apply = [&](auto& v) {
    auto old = v;
    for (int32_t k = 0; k < n; ++k) v[k] = old[perm[k]];
  };
  apply(m.x);
  apply(m.y);
  apply(m.depth);
  apply(m.open_bnd);
  apply(m.kb);
  if (!m.sig_node.empty()) {  // rows of the per-node sigma follow their nodes
    const size_t nv = (size_t)m.nvrt;
    std::vector<double> old = m.sig_node;
    for (int32_t k = 0; k < n; ++k)
      std::copy(old.begin() + perm[k] * nv, old.begin() + (perm[k] + 1) * nv, m.sig_node.begin() + k * nv);
  }
  std::vector<int32_t> en(3 * (size_t)ne);
  for (size_t k = 0; k < en.size(); ++k) en[k] = newid[m.elnode[k]];
  elem_perm.resize(ne);
  std::iota(elem_perm.begin(), elem_perm.end(), 0);
  std::vector<int64_t> key(ne);
  for (int32_t e = 0; e < ne; ++e) {
    const int32_t a = en[3 * e], b = en[3 * e + 1], c = en[3 * e + 2];
    const int32_t lo = std::min(a, std::min(b, c));
    const int32_t mid = std::max(std::min(a, b), std::min(std::max(a, b), c));
    key[e] = ((int64_t)lo << 32) | (uint32_t)mid;
  }
  std::stable_sort(elem_perm.begin(), elem_perm.end(), [&](int32_t p, int32_t q) { return key[p] < key[q]; });
  m.elnode.resize(3 * (size_t)ne);
  for (int32_t k = 0; k < ne; ++k)
    for (int j = 0; j < 3; ++j) m.elnode[3 * k + j] = en[3 * elem_perm[k] + j];
}

void setup_vertical(HostMesh& m, const VGridIn& vg) {
  const int32_t nv = m.nvrt;
  auto require = [](bool ok, const std::string& what) {
    if (!ok) throw Error(1 /* GSH_E_ARG */, "vertical grid: " + what);
  };
  if (vg.ivcor == 1) {
    require(vg.sigma_node != nullptr, "ivcor 1 needs sigma_node");
    m.sig_node.assign(vg.sigma_node, vg.sigma_node + (size_t)m.n * nv);
    for (int32_t i = 0; i < m.n; ++i) {
      const double* s = m.sig_node.data() + (size_t)i * nv;
      // the levels above a node's bed must increase towards the surface (its kb level is the bed)
      for (int k = m.kb[i] + 2; k < nv; ++k)
        require(s[k] > s[k - 1], "sigma must increase above the bottom level of node " + std::to_string(i));
    }
    m.vmode = 1;
    return;
  }
  require(vg.ivcor == 2, "ivcor must be 1 or 2");
  // readVGrid's checks (read_and_process_grid.cpp:1412-1486)
  require(vg.nz >= 1 && vg.nz <= nv - 1, "nz must be in [1, nvrt-1]");
  require(vg.h_s >= 6.0, "h_s needs to be larger");
  require(vg.h_c >= 5.0 && vg.h_c < vg.h_s, "h_c needs to be in [5, h_s)");
  require(vg.theta_b >= 0.0 && vg.theta_b <= 1.0, "wrong theta_b");
  require(vg.theta_f > 0.0, "wrong theta_f");
  require(vg.ztot != nullptr && vg.sigma_s != nullptr, "ivcor 2 needs ztot and sigma_s");
  const int32_t nz = vg.nz, nsig = nv - nz + 1;
  for (int k = 0; k < nz - 1; ++k) {
    require(vg.ztot[k] < -vg.h_s, "illegal Z level " + std::to_string(k + 1));
    require(k == 0 || vg.ztot[k] > vg.ztot[k - 1], "Z levels inverted at " + std::to_string(k + 1));
  }
  require(vg.ztot[nz - 1] == -vg.h_s, "the last Z level must be -h_s");
  // the first and last S levels are the bed (-1) and the surface (0); compZCoord never reads them
  for (int k = 1; k < nsig; ++k)
    require(vg.sigma_s[k] > vg.sigma_s[k - 1] && vg.sigma_s[k] >= -1.0 && vg.sigma_s[k] <= 0.0,
            "S levels must increase within [-1, 0]");
  m.nz = nz;
  m.h_s = vg.h_s;
  m.h_c = vg.h_c;
  m.theta_b = vg.theta_b;
  m.theta_f = vg.theta_f;
  m.ztot.assign(vg.ztot, vg.ztot + nz);
  m.ssig.assign(vg.sigma_s, vg.sigma_s + nsig);
  // the stretching function of readVGrid (read_and_process_grid.cpp:1497-1502)
  m.cs.resize(nsig);
  const double tb = vg.theta_b, tf = vg.theta_f;
  for (int k = 0; k < nsig; ++k)
    m.cs[k] = (1.0 - tb) * std::sinh(tf * m.ssig[k]) / std::sinh(tf) +
              tb * (std::tanh(tf * (m.ssig[k] + 0.5)) - std::tanh(tf * 0.5)) / 2.0 / std::tanh(tf * 0.5);
  m.vmode = 2;
  sz_bottom_levels(m);
}

void sz_bottom_levels(HostMesh& m) {
  const int32_t izlev = m.nz - 1;
  for (int32_t i = 0; i < m.n; ++i) {
    const double d = m.depth[i];
    int32_t kb = izlev;
    if (d > m.h_s) {
      // the Z level k < nz - 1 with ztot[k] <= -d < ztot[k + 1]
      if (!(-d >= m.ztot[0]))
        throw Error(1, "vertical grid: depth " + std::to_string(d) + " at node " + std::to_string(i) +
                           " lies below the deepest Z level");
      kb = 0;
      while (kb + 1 < izlev && m.ztot[kb + 1] <= -d) ++kb;
    }
    m.kb[i] = kb;
  }
}

}  // namespace gsh
