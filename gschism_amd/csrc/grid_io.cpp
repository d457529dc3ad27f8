// grid_io.cpp -- SCHISM grid files on the host: hgrid.gr3 (nodes, triangles, open and land
// boundary segments) and vgrid.in (ivcor=1 with one or per-node sigma columns, ivcor=2 SZ),
// parsed with the rules of the reference's readers (readHGrid, read_and_process_grid.cpp:855-1070;
// readVGrid, :1313-1520) into the arrays that gsh_mesh_create takes.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "abi_util.h"

struct gsh_grid {
  std::string name;
  int32_t n = 0, ne = 0, nvrt = 0;
  std::vector<double> xy, depth;
  std::vector<int32_t> elnode;                     // 3 per element, 0-based
  std::vector<std::vector<int32_t>> open_seg;      // 0-based node ids per open segment
  std::vector<std::vector<int32_t>> land_seg;      // 0-based node ids per land segment
  std::vector<int8_t> land_island;                 // island flag of each land segment
  std::vector<int32_t> kb;                         // vgrid: bottom level (0-based) per node (ivcor 1)
  std::vector<double> sigma;                       // vgrid: one sigma per level (ivcor 1, nodes agree)
  int32_t ivcor = 0, per_node = 0;                 // vgrid: 1 or 2; ivcor 1 with differing node sigmas
  std::vector<double> sigma_node;                  // vgrid: [n][nvrt] sigma of every node (ivcor 1)
  int32_t nz = 0;                                  // vgrid, ivcor 2: kz (Z levels incl. -h_s)
  double h_s = 0, h_c = 0, theta_b = 0, theta_f = 0;
  std::vector<double> ztot, sigma_s;               // vgrid, ivcor 2: [nz], [nvrt - nz + 1]
};

namespace gsh {
namespace {

// The whole file in memory, read line by line.
struct LineReader {
  std::string buf;
  size_t pos = 0;
  int64_t line_no = 0;
  std::string path;

  explicit LineReader(const std::string& p) : path(p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) throw Error(GSH_E_IO, "could not open " + p + ": " + std::strerror(errno));
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != buf.size()) throw Error(GSH_E_IO, "could not read " + p);
  }
  bool eof() const { return pos >= buf.size(); }
  // next line without its end-of-line characters; false at end of file
  bool next(const char*& b, const char*& e) {
    if (eof()) return false;
    const size_t nl = buf.find('\n', pos);
    const size_t end = nl == std::string::npos ? buf.size() : nl;
    b = buf.data() + pos;
    e = buf.data() + end;
    if (e > b && e[-1] == '\r') --e;
    pos = end + 1;
    ++line_no;
    return true;
  }
  [[noreturn]] void fail(const std::string& what) const {
    throw Error(GSH_E_IO, path + ":" + std::to_string(line_no) + ": " + what);
  }
  // readInt (misc.h:32-46): the first integer of the next line, the rest ignored
  int64_t first_int(const char* what) {
    const char *b, *e;
    if (!next(b, e)) fail(std::string("end of file before ") + what);
    std::string s(b, e);
    char* end = nullptr;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (end == s.c_str()) fail(std::string("expected an integer (") + what + ")");
    return v;
  }
};

// Numbers of the line [b, e), in order (parsing stops at the first token that is not a number).
// The buffer behind e is NUL-terminated, and a number that starts past e (strtod skips the line
// end as white space) belongs to the next line.
std::vector<double> line_numbers(const char* b, const char* e) {
  std::vector<double> out;
  const char* p = b;
  while (p < e) {
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    if (p >= e) break;
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p || end > e) break;
    out.push_back(v);
    p = end;
  }
  return out;
}

int32_t node_id(LineReader& r, int64_t v, int32_t n) {
  if (v < 1 || v > n) r.fail("node id " + std::to_string(v) + " out of range 1.." + std::to_string(n));
  return (int32_t)(v - 1);
}

void read_hgrid(gsh_grid& g, const std::string& path) {
  LineReader r(path);
  const char *b, *e;
  if (!r.next(b, e)) r.fail("empty file");
  g.name.assign(b, e);
  if (!r.next(b, e)) r.fail("end of file before the element and node counts");
  const std::vector<double> cnt = line_numbers(b, e);
  if (cnt.size() < 2 || cnt[0] < 1 || cnt[1] < 3) r.fail("expected 'ne np'");
  g.ne = (int32_t)cnt[0];
  g.n = (int32_t)cnt[1];
  g.xy.resize(2 * (size_t)g.n);
  g.depth.resize(g.n);
  for (int32_t i = 0; i < g.n; ++i) {
    if (!r.next(b, e)) r.fail("end of file in the node table");
    const std::vector<double> v = line_numbers(b, e);
    if (v.size() < 4) r.fail("expected 'id x y depth'");
    g.xy[2 * (size_t)i] = v[1];
    g.xy[2 * (size_t)i + 1] = v[2];
    g.depth[i] = v[3];
  }
  g.elnode.resize(3 * (size_t)g.ne);
  for (int32_t k = 0; k < g.ne; ++k) {
    if (!r.next(b, e)) r.fail("end of file in the element table");
    const std::vector<double> v = line_numbers(b, e);
    if (v.size() < 2) r.fail("expected 'id i34 nodes...'");
    const int i34 = (int)v[1];
    if (i34 == 4) r.fail("quadrilateral elements are not supported");
    if (i34 != 3 || v.size() < 5) r.fail("expected a triangle 'id 3 n1 n2 n3'");
    for (int j = 0; j < 3; ++j) g.elnode[3 * (size_t)k + j] = node_id(r, (int64_t)v[2 + j], g.n);
  }
  // open boundary segments; a file without boundary sections has none
  if (r.eof()) return;
  const int64_t nope = r.first_int("the number of open boundaries");
  const int64_t neta = r.first_int("the number of open boundary nodes");
  if (nope < 0 || neta < 0) r.fail("negative open boundary counts");
  int64_t total = 0;
  for (int64_t s = 0; s < nope; ++s) {
    const int64_t nn = r.first_int("the node count of an open boundary");
    if (nn < 1) r.fail("empty open boundary segment");
    std::vector<int32_t> seg(nn);
    for (int64_t i = 0; i < nn; ++i) seg[i] = node_id(r, r.first_int("an open boundary node"), g.n);
    if (seg.front() == seg.back())  // read_and_process_grid.cpp:981-985
      r.fail("looped open boundary " + std::to_string(s + 1));
    total += nn;
    g.open_seg.push_back(std::move(seg));
  }
  if (total != neta)
    r.fail("neta (" + std::to_string(neta) + ") != total # of open boundary nodes (" + std::to_string(total) + ")");
  // land boundary segments: 'nn island' then nn node lines (read_and_process_grid.cpp:998-1047)
  if (r.eof()) return;
  const int64_t nland = r.first_int("the number of land boundaries");
  const int64_t nvel = r.first_int("the number of land boundary nodes");
  if (nland < 0 || nvel < 0) r.fail("negative land boundary counts");
  total = 0;
  for (int64_t s = 0; s < nland; ++s) {
    if (!r.next(b, e)) r.fail("end of file before a land boundary header");
    const std::vector<double> h = line_numbers(b, e);
    if (h.empty() || h[0] < 1) r.fail("expected 'nn island' of a land boundary");
    const int64_t nn = (int64_t)h[0];
    const int8_t island = (h.size() > 1 && h[1] == 1.0) ? 1 : 0;
    std::vector<int32_t> seg(nn);
    for (int64_t i = 0; i < nn; ++i) seg[i] = node_id(r, r.first_int("a land boundary node"), g.n);
    total += nn;
    g.land_seg.push_back(std::move(seg));
    g.land_island.push_back(island);
  }
  if (total != nvel)
    r.fail("nvel (" + std::to_string(nvel) + ") != total # of land boundary nodes (" + std::to_string(total) + ")");
}

// vgrid.in (readVGrid, read_and_process_grid.cpp:1313-1520).
//
// ivcor=1 (:1341-1395): ivcor; nvrt; kbp of every node; then nvrt rows 'k sigma(node 1) ...
// sigma(node np)'. Every node's column is kept; when the nodes agree on sigma at every level
// above their bottom level (a node's own bottom level is its bed, z = -h, whatever its sigma
// there) the grid also has one sigma per level.
//
// ivcor=2 (:1397-1503): 'nvrt kz h_s'; a comment line; the Z levels 'k z(k)' for k = 1..kz, of
// which only 1..kz-1 are read (z(kz) is -h_s); a comment line; 'h_c theta_b theta_f'; the S levels
// 'k sigma(k)' for k = 1..nvrt-kz+1, of which the first (-1) is not read; with readVGrid's checks.
// readVGrid takes one line for the kz-th Z level only when kz = 1 (for kz > 1 its line skips
// leave that line to be read as the S-levels comment); here every kz reads that format.
void read_vgrid(gsh_grid& g, const std::string& path) {
  LineReader r(path);
  const int64_t ivcor = r.first_int("ivcor");
  if (ivcor != 1 && ivcor != 2) r.fail("ivcor must be 1 or 2");
  g.ivcor = (int32_t)ivcor;
  if (ivcor == 2) {
    const char *b, *e;
    if (!r.next(b, e)) r.fail("end of file before 'nvrt kz h_s'");
    const std::vector<double> head = line_numbers(b, e);
    if (head.size() < 3) r.fail("expected 'nvrt kz h_s'");
    const int64_t nvrt = (int64_t)head[0], kz = (int64_t)head[1];
    const double h_s = head[2];
    if (nvrt < 2 || nvrt > 1024) r.fail("nvrt must be in [2, 1024]");
    if (kz < 1 || kz > nvrt - 1) r.fail("kz must be in [1, nvrt-1]");
    if (h_s < 6.0) r.fail("h_s needs to be larger: " + std::to_string(h_s));
    if (!r.next(b, e)) r.fail("end of file before the Z levels");  // "Z levels"
    g.ztot.assign((size_t)kz, 0.0);
    for (int64_t k = 0; k < kz; ++k) {
      if (!r.next(b, e)) r.fail("end of file in the Z levels");
      const std::vector<double> v = line_numbers(b, e);
      if (v.size() < 2 || (int64_t)v[0] != k + 1) r.fail("invalid level number in the Z levels");
      if (k == kz - 1) break;  // z(kz) is -h_s
      g.ztot[k] = v[1];
      if (g.ztot[k] >= -h_s) r.fail("illegal Z level " + std::to_string(k + 1));
      if (k > 0 && g.ztot[k] <= g.ztot[k - 1]) r.fail("Z levels inverted at " + std::to_string(k + 1));
    }
    g.ztot[kz - 1] = -h_s;
    if (!r.next(b, e)) r.fail("end of file before the S levels");  // "S levels"
    if (!r.next(b, e)) r.fail("end of file before 'h_c theta_b theta_f'");
    const std::vector<double> sp = line_numbers(b, e);
    if (sp.size() < 3) r.fail("expected 'h_c theta_b theta_f'");
    const double h_c = sp[0], theta_b = sp[1], theta_f = sp[2];
    if (h_c < 5.0 || h_c >= h_s) r.fail("h_c needs to be larger to avoid 2nd type abnormality");
    if (theta_b < 0.0 || theta_b > 1.0) r.fail("wrong theta_b");
    if (theta_f <= 0.0) r.fail("wrong theta_f");
    const int64_t nsig = nvrt - kz + 1;
    g.sigma_s.assign((size_t)nsig, 0.0);
    g.sigma_s[0] = -1.0;
    g.sigma_s[nsig - 1] = 0.0;
    for (int64_t k = 0; k < nsig; ++k) {
      if (!r.next(b, e)) r.fail("end of file in the S levels");
      const std::vector<double> v = line_numbers(b, e);
      if (v.size() < 2 || (int64_t)v[0] != k + 1) r.fail("invalid level number in the S levels");
      if (k == 0) continue;  // sigma(1) = -1
      g.sigma_s[k] = v[1];
      if (g.sigma_s[k] <= g.sigma_s[k - 1] || g.sigma_s[k] > 0.0 || g.sigma_s[k] < -1.0)
        r.fail("invalid sigma levels at " + std::to_string(k + 1));
    }
    g.nvrt = (int32_t)nvrt;
    g.nz = (int32_t)kz;
    g.h_s = h_s;
    g.h_c = h_c;
    g.theta_b = theta_b;
    g.theta_f = theta_f;
    return;
  }
  const int64_t nvrt = r.first_int("nvrt");
  if (nvrt < 2 || nvrt > 1024) r.fail("nvrt must be in [2, 1024]");
  // the rest is a stream of numbers
  const char* p = r.buf.c_str() + std::min(r.pos, r.buf.size());
  auto num = [&](const char* what) {
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p) throw Error(GSH_E_IO, path + ": expected a number (" + what + ")");
    p = end;
    return v;
  };
  g.nvrt = (int32_t)nvrt;
  g.kb.resize(g.n);
  for (int32_t i = 0; i < g.n; ++i) {
    const double v = num("kbp");
    const int32_t kb = (int32_t)v - 1;
    if (kb < 0 || kb > nvrt - 2)
      throw Error(GSH_E_IO, path + ": node " + std::to_string(i + 1) + " has kbp " + std::to_string((int)v) +
                                " outside 1.." + std::to_string(nvrt - 1));
    g.kb[i] = kb;
  }
  g.sigma_node.assign((size_t)g.n * nvrt, 0.0);
  g.sigma.assign((size_t)nvrt, 0.0);
  g.sigma[0] = -1.0;
  bool shared = true;
  for (int32_t k = 0; k < nvrt; ++k) {
    const double row = num("level number");
    if ((int64_t)row != k + 1) throw Error(GSH_E_IO, path + ": level rows out of order at level " + std::to_string(k + 1));
    bool have = false;
    for (int32_t i = 0; i < g.n; ++i) {
      const double s = num("sigma");
      g.sigma_node[(size_t)i * nvrt + k] = s;
      if (k == 0 ? g.kb[i] != 0 : k <= g.kb[i]) continue;  // at or below the node's bed
      if (!have) {
        g.sigma[k] = s;
        have = true;
      } else if (s != g.sigma[k]) {
        shared = false;
      }
    }
  }
  for (int32_t i = 0; i < g.n && shared; ++i)
    for (int32_t k = std::max(1, g.kb[i] + 1); k < nvrt; ++k)
      if (g.sigma_node[(size_t)i * nvrt + k] != g.sigma[k]) shared = false;
  if (shared) {
    // levels 1..min(kb) lie at or below every node's bed, so no node gives their sigma and z never
    // reads it: spread them evenly between -1 and the first level some node uses
    const int32_t m = *std::min_element(g.kb.begin(), g.kb.end());
    for (int32_t k = 1; k <= m; ++k) g.sigma[k] = -1.0 + (g.sigma[m + 1] + 1.0) * k / (m + 1);
  }
  if (shared)
    for (int32_t k = 1; k < nvrt; ++k)
      if (!(g.sigma[k] > g.sigma[k - 1])) throw Error(GSH_E_IO, path + ": sigma must increase with the level");
  g.per_node = shared ? 0 : 1;
  if (shared) {
    g.sigma_node.clear();
  } else {
    g.sigma.clear();
    for (int32_t i = 0; i < g.n; ++i)
      for (int32_t k = g.kb[i] + 2; k < nvrt; ++k)
        if (!(g.sigma_node[(size_t)i * nvrt + k] > g.sigma_node[(size_t)i * nvrt + k - 1]))
          throw Error(GSH_E_IO, path + ": sigma must increase with the level above the bed of node " +
                                    std::to_string(i + 1));
  }
}

}  // namespace
}  // namespace gsh

extern "C" {

int gsh_grid_read(const char* hgrid_path, const char* vgrid_path, gsh_grid** out) {
  GSH_API_BODY({
    GSH_REQUIRE(hgrid_path && out, GSH_E_ARG, "NULL argument");
    *out = nullptr;
    auto g = std::make_unique<gsh_grid>();
    gsh::read_hgrid(*g, hgrid_path);
    if (vgrid_path) gsh::read_vgrid(*g, vgrid_path);
    *out = g.release();
  })
}

int gsh_grid_info(const gsh_grid* g, int32_t* n, int32_t* ne, int32_t* nvrt, int32_t* n_open_seg,
                  int32_t* n_land_seg) {
  GSH_API_BODY({
    GSH_REQUIRE(g, GSH_E_ARG, "grid is NULL");
    if (n) *n = g->n;
    if (ne) *ne = g->ne;
    if (nvrt) *nvrt = g->nvrt;
    if (n_open_seg) *n_open_seg = (int32_t)g->open_seg.size();
    if (n_land_seg) *n_land_seg = (int32_t)g->land_seg.size();
  })
}

int gsh_grid_get(const gsh_grid* g, double* xy, double* depth, int32_t* elnode, int8_t* open_bnd, int32_t* kb,
                 double* sigma) {
  GSH_API_BODY({
    GSH_REQUIRE(g, GSH_E_ARG, "grid is NULL");
    GSH_REQUIRE(!(kb || sigma) || g->nvrt > 0, GSH_E_STATE, "no vgrid.in was read");
    GSH_REQUIRE(!kb || g->ivcor == 1, GSH_E_STATE, "an SZ grid (ivcor 2) has no bottom levels in the file");
    GSH_REQUIRE(!sigma || (g->ivcor == 1 && !g->per_node), GSH_E_STATE,
                "the vertical grid has no single sigma column (gsh_grid_vgrid)");
    if (xy) std::memcpy(xy, g->xy.data(), 8 * g->xy.size());
    if (depth) std::memcpy(depth, g->depth.data(), 8 * g->depth.size());
    if (elnode) std::memcpy(elnode, g->elnode.data(), 4 * g->elnode.size());
    if (open_bnd) {
      std::memset(open_bnd, 0, (size_t)g->n);
      for (const auto& s : g->open_seg)
        for (int32_t i : s) open_bnd[i] = 1;
    }
    if (kb) std::memcpy(kb, g->kb.data(), 4 * g->kb.size());
    if (sigma) std::memcpy(sigma, g->sigma.data(), 8 * g->sigma.size());
  })
}

int gsh_grid_vgrid(const gsh_grid* g, int32_t* ivcor, int32_t* per_node, int32_t* nz, double* params,
                   double* ztot, double* sigma_s, double* sigma_node) {
  GSH_API_BODY({
    GSH_REQUIRE(g, GSH_E_ARG, "grid is NULL");
    GSH_REQUIRE(g->nvrt > 0, GSH_E_STATE, "no vgrid.in was read");
    if (ivcor) *ivcor = g->ivcor;
    if (per_node) *per_node = g->per_node;
    if (nz) *nz = g->nz;
    if (g->ivcor == 2) {
      if (params) params[0] = g->h_s, params[1] = g->h_c, params[2] = g->theta_b, params[3] = g->theta_f;
      if (ztot) std::memcpy(ztot, g->ztot.data(), 8 * g->ztot.size());
      if (sigma_s) std::memcpy(sigma_s, g->sigma_s.data(), 8 * g->sigma_s.size());
    }
    if (sigma_node) {
      GSH_REQUIRE(g->ivcor == 1 && g->per_node, GSH_E_STATE, "sigma_node: the grid has no per-node sigma");
      std::memcpy(sigma_node, g->sigma_node.data(), 8 * g->sigma_node.size());
    }
  })
}

int gsh_grid_segment(const gsh_grid* g, int land, int32_t seg, int32_t* nodes_out, int32_t cap, int32_t* count,
                     int32_t* island) {
  GSH_API_BODY({
    GSH_REQUIRE(g && count, GSH_E_ARG, "NULL argument");
    const auto& segs = land ? g->land_seg : g->open_seg;
    GSH_REQUIRE(seg >= 0 && seg < (int32_t)segs.size(), GSH_E_ARG, "segment index out of range");
    const auto& s = segs[seg];
    *count = (int32_t)s.size();
    if (island) *island = land ? g->land_island[seg] : 0;
    if (nodes_out) {
      GSH_REQUIRE(cap >= (int32_t)s.size(), GSH_E_ARG, "output too small");
      std::memcpy(nodes_out, s.data(), 4 * s.size());
    }
  })
}

int gsh_grid_destroy(gsh_grid* g) {
  delete g;
  return GSH_OK;
}

}  // extern "C"
