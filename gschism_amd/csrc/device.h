// device.h -- device-resident data layout (SoA in HBM) and launch wrappers of libgsh.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gsh {

// Wavefront and workgroup geometry of the CDNA kernels.
constexpr int kWave = 64;
constexpr int kTpb = 256;              // 4 waves per workgroup
constexpr int kSpmvLdsCap = 4096;      // products staged per 256-row CSR chunk (32 KiB FP64)

// Mesh state on the device. Per-level arrays are [entity][level].
struct DMesh {
  int32_t n, ne, ns, nvrt;
  // nodes
  const double* x;
  const double* y;
  const double* depth;
  const int32_t* kb;
  const double* sigma;      // [nvrt] (vmode 0)
  // vertical coordinates of Node::compZCoord (meshdata.cu:1744-1807): vmode 0 = ivcor 1 with one
  // sigma per level, 1 = ivcor 1 with a sigma per node and level, 2 = ivcor 2 (SZ)
  int32_t vmode;
  int32_t bt_stage;         // backtracking stages each workgroup's element patch in LDS (0: off)
  int32_t izlev;            // nz - 1: the level at -min(h, h_s) (vmode 2)
  double h_s, h_c;          // (vmode 2)
  const double* sig_node;   // [n][nvrt] (vmode 1)
  const double* ztot;       // [nz] Z levels (vmode 2)
  const double* ssig;       // [nvrt - nz + 1] S levels (vmode 2)
  const double* cs;         // [nvrt - nz + 1] C(sigma) of the S levels (vmode 2)
  const int8_t* open_bnd;
  const int2* node_bside;   // [n] boundary sides ending at the node, ascending side index, -1 = none
  double* eta2;
  double* eta1;
  double* grav2;
  double* cd;
  double* eth;
  double* elbc;
  double* uvw;              // [n][nvrt][4]: u, v, w, 0
  double2* nrec;            // [n][2]: (x, y), (eta2, depth), packed for the backtracking gathers
  double* eta_lev;          // [n] eta2 of the last level update (the node z levels follow it)
  int8_t* node_wet;
  int8_t* etype;            // [n] elevation BC type of each open node (-1: the step parameter iettype)
  int32_t nbfr;             // tidal constituents of the open boundary (0: none)
  const double* tide;       // [3][nbfr]: angular frequency, nodal factor, equilibrium argument
  const double* emo;        // [n][nbfr] amplitude at each open node
  const double* efa;        // [n][nbfr] phase at each open node
  // elements
  const int4* elem;         // [ne][2]: (n0, n1, n2, s0), (s1, s2, -, -)
  const double* area;
  const double* gradphi;    // [ne][3][2]
  int8_t* elem_wet;
  int32_t* elem_kb;
  double* ghat1;            // [ne][2]
  double* vsource;
  double* econ;             // [ne][3][8]: assembly terms of each (element, corner)
  double* ew;               // [ne][nvrt] element vertical velocity (allocated at first use)
  // sides
  const int4* side;         // [ns]: (n1, n2, e0, e1)
  const int8_t* side_type;  // 0 internal, 1 land, 2 open
  const int32_t* side_nbr;  // [ns][4] the Shapiro filter's neighbour sides (internal sides)
  const double* side_cx;
  const double* side_cy;
  const double* side_len;
  int32_t* side_kb;
  int8_t* side_wet;
  double* zc;               // [ns][nvrt]
  double* su2;              // [ns][nvrt][2]
  double* btu;              // [ns][nvrt][2]
  int32_t* foot;            // [ns][nvrt]
  int32_t* bstatus;         // [ns][nvrt]
  double* sU;               // [ns][2]
  double* side_us;          // [ns][2] vertical integral of btu of each side (computeElemAverageGhat)
  double* hhat;
  double* chi;
  // balls and CSR pattern
  const int32_t* ball_ptr;
  const int32_t* ball_elem;
  const int8_t* ball_loc;
  const int16_t* ball_slot;  // [ball][2]: CSR slots of the two other element nodes
  const int32_t* nbr_ptr;
  const int32_t* nbr_node;
  const int16_t* csr_slot;
  const int32_t* indptr;
  const int32_t* diag_off;
};

struct StepArgs {
  double dt, theta, h0, ramp, dtb_min, dtb_max, velmin2, dfv0, rmaxvel, time, shapiro0;
  int32_t ihhat, iettype, ndelt_min, ndelt_max, ishapiro;
};

// ---------------------------------------------------------------- PCG
// Scalar state of one system, device resident. k / kk are the iteration counters read by
// the two per-iteration phases (each phase reads one and its workgroup 0 writes the other).
struct PcgState {
  int32_t k, kk, done, zero;
  int32_t maxiter, check_every, pad0, pad1;
  double rtol, atol2, rr_last, bb;
  double rho[2];    // r.z used by iteration k (slot k & 1)
  double beta;      // SSOR: beta of iteration kk, for the sweeps after the first
  double pad2;
  double alpha[2];  // SSOR: alpha of iteration k (slot k & 1)
};

// One linear system of a (possibly batched) PCG launch. VT is the arithmetic type of the
// matrix/vector updates; dot products are accumulated in FP64.
template <class VT>
struct PcgSys {
  int32_t n, ntiles;
  const int32_t* rowptr;
  const int32_t* col;
  const VT* val;
  const int32_t* diag;       // absolute index of the diagonal entry of each row
  const int16_t* cdel;       // column - (row & ~63) when every offset fits int16, else NULL
  const double* b;           // FP64 right-hand side (MIXED inner solve: FP64 residual)
  VT* x;
  VT* r;
  VT* z;
  VT* pa;
  VT* pb;
  VT* q;
  VT* invd;
  const double* h0;          // extrapolation history x0, x1, x2 (FP64), NULL = zero guess
  const double* h1;
  const double* h2;
  double* part_rr;
  double* part_rz;
  double* part_pq;
  double* part_bb;
  PcgState* st;
};

// Multicolour SSOR system: rows in colour-major order, each colour block padded to whole
// tiles of 256*R rows with identity rows; the off-diagonal part split into the strictly lower
// part (columns of lower colours) and the strictly upper part, stored ELL per colour (entry m of
// colour-offset row r at off + m * rows + r, column -1 where a row has fewer entries).
struct SsorColour {
  int32_t start, rows, wl, wu;   // first row, padded row count, ELL widths
  int32_t offl, offu, pad0, pad1;  // ELL offsets of the lower / upper part
};

// VT is the type of the matrix values and iteration vectors (double; float for the inner solves
// of mixed precision); b, the history, the dot-product partials and the scalars stay FP64.
template <class VT>
struct SsorSysT {
  int32_t n, np, ntiles, ncol;
  double omega;
  int32_t unit_omega;                        // omega == 1: 2 D~ - D equals D~ exactly
  int32_t pre_t0;                            // first tile of the last colour
  const SsorColour* colours;                 // [ncol]
  const int32_t* tile_colour;                // [ntiles]
  const int32_t *lcol, *ucol;                // ELL columns (-1 unused)
  const int32_t *lsrc, *usrc, *dsrc;         // slots of each entry in the solver's CSR values (-1 unused)
  const int32_t* perm;                       // [np] new row -> original row, -1 for padding
  const VT* aval;                            // the solver's CSR values (original order)
  const double* b;                           // right-hand side (original order)
  const double *h0, *h1, *h2;                // extrapolation history (original order), NULL = zero guess
  VT* xout;                                  // solution (original order)
  VT *lval, *uval;                           // gathered ELL values (0 where unused)
  VT *d, *dt, *kd, *bp;                      // diagonal, D/omega, 2 D/omega - D, permuted b
  VT *y, *r, *p, *q, *t, *u, *s, *w;         // iteration vectors [np]: CG y r p q, sweeps t u, s = A^ z
  double *part_rr, *part_bb;                 // [ntiles]
  double *part_rz, *part_pq;                 // [2][ntiles]: gamma and delta partials, slot = iteration & 1
  double* pre;                               // [2][2][256]: per-thread gamma sums and delta sums below
                                             // pre_t0, slot = iteration & 1 (NULL: not kept)
  PcgState* st;
};
using SsorSys = SsorSysT<double>;

// Batched SSOR launch table entry: the workgroup's system, tile, colour and sweep flags (1 first
// sweep of the iteration, 2 colour 0, 4 last sweep, 8 set-up pass, 16 the system's lead
// workgroup, which writes its scalar state).
struct SsorRef {
  int32_t sys, tile, c, flags;
};

// Launch table entry: which system and which tile a workgroup works on.
struct TileRef {
  int32_t sys, tile;
};

// ---------------------------------------------------------------- multi-RHS block (block_kernels.hip)
constexpr int kBlkTile = 512;  // rows per workgroup tile of the block kernels

// Scalar state of one member of a block solve.
struct BlkState {
  int32_t done, zero, pad0, pad1;  // done: iteration count once stopped, -1 before
  double atol2, rz_cur, rz_old, alpha, beta, rr_last;
};

// M systems on one CSR pattern, member-interleaved: vectors [n][M], values [nnz][M].
template <class VT>
struct BlkSys {
  int32_t n, M, ntiles, pad;
  const int32_t* rowptr;
  const int32_t* col;
  const int32_t* diag;     // absolute index of each row's diagonal entry
  const VT* val;           // [nnz][M]
  const double* b;         // [n][M]
  VT *x, *r, *z, *pa, *pb, *q, *invd;
  double* part_a;          // [ntiles][M]
  double* part_b;          // [ntiles][M]
  BlkState* st;            // [M]
};

// ---------------------------------------------------------------- launch wrappers
void launch_levels(const DMesh& m, const StepArgs& a, hipStream_t s);
void launch_backtrack(const DMesh& m, const StepArgs& a, const int32_t* list_side, const int32_t* list_lev,
                      int32_t count, hipStream_t s);
void launch_rhs(const DMesh& m, const StepArgs& a, hipStream_t s);
void launch_nodal_velocity(const DMesh& m, double rmaxvel, hipStream_t s);
void launch_momentum(const DMesh& m, const StepArgs& a, hipStream_t s);
void launch_shapiro(const DMesh& m, double shapiro0, double* tmp, hipStream_t s);
void launch_vertical_velocity(const DMesh& m, double h0, hipStream_t s);
void launch_assemble(const DMesh& m, const StepArgs& a, double* val, double* b, hipStream_t s);
void launch_update_elevation(const DMesh& m, const double* x, hipStream_t s);

template <class VT>
void launch_pcg_reset(const PcgSys<VT>* sys, int nsys, double rtol, int maxiter, int check_every, hipStream_t s,
                      const double* rtols = nullptr);
template <class VT>
void launch_pcg_init(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s);
template <class VT>
void launch_pcg_iter(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s);
template <class VT>
void launch_pcg_phase(int which, const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
template <class VT>
void launch_pcg_check(const PcgSys<VT>* sys, int nsys, hipStream_t s);
template <class VT>
void launch_pcg_finish(const PcgSys<VT>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s);
// q = A x (FP64), then per-tile partials of (b - q).(b - q) and b.b for checkSolution.
void launch_true_residual(const PcgSys<double>* sys, const TileRef* tiles, int ntiles, int rpt, const double* xv,
                          double* rvec, hipStream_t s);
void launch_reduce_pair(const double* pa, const double* pb, int ntiles, double* out2, hipStream_t s);

// multi-RHS block (block_kernels.hip)
template <class VT>
void launch_blk_init(const BlkSys<VT>& S, hipStream_t st);
template <class VT>
void launch_blk_reduce_init(const BlkSys<VT>& S, double rtol, hipStream_t st);
template <class VT>
void launch_blk_iter(const BlkSys<VT>& S, int it, double rtol, int maxiter, int check_every, hipStream_t st);
void launch_blk_put(double* vblk, double* bblk, const double* v, const double* b, int64_t nnz, int n, int M, int m,
                    hipStream_t st);
template <class VT>
void launch_blk_get(double* x, const VT* xblk, int n, int M, int m, hipStream_t st);

// Multicolour SSOR PCG (one system). Tiles of colour c are [tile0, tile0 + ntc).
template <class VT>
void launch_ssor_reset(const SsorSysT<VT>* sys, double rtol, int maxiter, int check_every, hipStream_t s);
template <class VT>
void launch_ssor_setup(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s, bool gather = true);
template <class VT>
void launch_ssor_residual(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s);
template <class VT>
void launch_ssor_r0(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int rpt, hipStream_t s);
// flags: 1 first sweep of the iteration, 2 colour 0 (starts the forward sweep), 4 last sweep
// of the iteration, 8 the set-up pass (no vector update)
template <class VT>
void launch_ssor_bwd(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int flags, int rpt, hipStream_t s,
                     hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
template <class VT>
void launch_ssor_fwd(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int flags, int rpt, hipStream_t s,
                     hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
template <class VT>
void launch_ssor_check(const SsorSysT<VT>* sys, hipStream_t s);
template <class VT>
void launch_ssor_xsolve(const SsorSysT<VT>* sys, int tile0, int ntc, int c, int rpt, hipStream_t s);
template <class VT>
void launch_ssor_scatter(const SsorSysT<VT>* sys, int ntiles, int rpt, hipStream_t s);
// Batched sweeps over a table of nrefs entries (sys is an array of descriptors), the batched
// stopping test (one workgroup per system) and the states of count systems gathered into out.
template <class VT>
void launch_ssor_bwd_batch(const SsorSysT<VT>* sys, const SsorRef* refs, int nrefs, int or_flags, int rpt,
                           hipStream_t s);
template <class VT>
void launch_ssor_fwd_batch(const SsorSysT<VT>* sys, const SsorRef* refs, int nrefs, int or_flags, int rpt,
                           hipStream_t s);
template <class VT>
void launch_ssor_check_batch(const SsorSysT<VT>* sys, int count, hipStream_t s);
template <class VT>
void launch_ssor_gather_states(const SsorSysT<VT>* sys, int count, PcgState* out, hipStream_t s);

// Mixed precision helpers.
void launch_cast_values(const double* in, float* out, int64_t count, hipStream_t s);
// Batched mixed-precision steps over a launch table: r = b - A x (FP64, into sys.r) with r.r / b.b
// partials; per-system totals of those partials; x64 += d32.
void launch_residual_batch(const PcgSys<double>* sys, const TileRef* tiles, int ntiles, int rpt, hipStream_t s);
void launch_reduce_systems(const PcgSys<double>* sys, int nsys, double* out, hipStream_t s);
void launch_refine_batch(const PcgSys<double>* sys64, const PcgSys<float>* sys32, const TileRef* tiles, int ntiles,
                         int rpt, hipStream_t s);
void launch_refine_update(double* x64, const float* d32, int32_t n, hipStream_t s);
void launch_copy_f64(double* dst, const double* src, int64_t count, hipStream_t s);
void launch_fill_f64(double* dst, double v, int64_t count, hipStream_t s);

}  // namespace gsh
