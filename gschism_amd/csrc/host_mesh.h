// host_mesh.h -- host-side mesh containers, RNG and topology builders shared by the
// generator, the topology builder and the device-upload code of libgsh.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gsh {

// Error carried to the C ABI as a status code + message.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// splitmix64-seeded xoshiro256** with Box-Muller normals. Used for every synthetic input so
// that two processes (or the device path and the CPU oracle) see identical data.
struct Rng {
  uint64_t s[4];
  bool has_spare = false;
  double spare = 0.0;
  explicit Rng(uint64_t seed);
  uint64_t next();
  double uniform();        // [0, 1) with 53 random bits
  double uniform_open();   // (0, 1]
  double normal();         // N(0, 1)
};

// Triangular mesh as handed to the library (after optional renumbering).
struct HostMesh {
  int32_t n = 0, ne = 0, nvrt = 0;
  std::vector<double> x, y, depth, sigma;
  std::vector<int32_t> elnode;     // [ne][3]
  std::vector<int8_t> open_bnd;    // [n]
  std::vector<int32_t> kb;         // [n]
  // vertical coordinates: vmode 0 = ivcor 1 with `sigma` at every node, 1 = ivcor 1 with a sigma
  // per node and level, 2 = ivcor 2 (SZ)
  int32_t vmode = 0;
  std::vector<double> sig_node;    // [n][nvrt] (vmode 1)
  int32_t nz = 0;                  // Z levels incl. the one at -h_s (vmode 2)
  double h_s = 0, h_c = 0, theta_b = 0, theta_f = 0;
  std::vector<double> ztot;        // [nz] (vmode 2)
  std::vector<double> ssig, cs;    // [nvrt - nz + 1] S levels and their stretching C(sigma) (vmode 2)
};

// Vertical grid of a gsh_vgrid (validated as readVGrid does, read_and_process_grid.cpp:1397-1503)
// into m (input node order; before renumbering). Throws Error(GSH_E_ARG) on a bad grid.
struct VGridIn {
  int32_t ivcor;
  const double* sigma_node;
  int32_t nz;
  double h_s, h_c, theta_b, theta_f;
  const double* ztot;
  const double* sigma_s;
};
void setup_vertical(HostMesh& m, const VGridIn& vg);
// SZ bottom level of every node from its depth (Node::updateKb, meshdata.cu:1676-1718): the Z
// level below the bed, or nz-1 where the depth is within h_s. Throws where the bed lies below
// the deepest Z level.
void sz_bottom_levels(HostMesh& m);

// Topology of HostMesh following read_and_process_grid.cpp and meshdata.cu:2094-2141.
struct Topology {
  int32_t ns = 0;
  std::vector<int32_t> ic3;        // [ne][3] element across side j, -1 = none
  std::vector<int32_t> elside;     // [ne][3]
  std::vector<int32_t> isdel;      // [ns][2]
  std::vector<int32_t> isidenode;  // [ns][2]
  std::vector<int8_t> side_type;   // [ns] 0 internal, 1 land, 2 open
  std::vector<int32_t> side_nbr;   // [ns][4] Side::neigh(j) of internal sides (meshdata.h:1409-1422), -1 else
  std::vector<int8_t> isbnd;       // [n] 1 = node on the mesh boundary
  std::vector<int32_t> node_bside; // [n][2] boundary sides ending at the node, ascending, -1 = none
  std::vector<int32_t> ball_ptr;   // [n+1]
  std::vector<int32_t> ball_elem;  // CCW ball
  std::vector<int8_t> ball_loc;    // local index of the node in each ball element
  std::vector<int32_t> nbr_ptr;    // [n+1] neighbour-node ranges (nodeNeighNodes)
  std::vector<int32_t> nbr_node;
  std::vector<int32_t> indptr;     // [n+1] CSR
  std::vector<int32_t> indices;    // [nnz]
  std::vector<int16_t> csr_slot;   // [nbr count] slot of neighbour j within its row
  std::vector<int32_t> diag_off;   // [n] slot of the diagonal within its row
  int32_t max_row = 0;
};

// Geometry derived from HostMesh + Topology (populateSides / populateElements).
struct Geometry {
  std::vector<double> area;        // [ne]
  std::vector<double> gradphi;     // [ne][3][2]
  std::vector<double> ssign;       // [ne][3]
  std::vector<double> side_cx, side_cy, side_len, side_nx, side_ny;  // [ns]
};

void build_topology(const HostMesh& m, Topology& t);
void build_geometry(const HostMesh& m, const Topology& t, Geometry& g);
// Reverse Cuthill-McKee of the node graph: perm[new] = old.
std::vector<int32_t> rcm_order(int32_t n, int32_t ne, const std::vector<int32_t>& elnode);
// Applies a node renumbering and orders elements by their smallest new node id.
void renumber(HostMesh& m, const std::vector<int32_t>& node_perm, std::vector<int32_t>& elem_perm);

}  // namespace gsh
